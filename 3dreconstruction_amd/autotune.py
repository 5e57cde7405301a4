"""Autotuner of the engine's launch shapes on the GPU it runs on.

Sweeps, per benchmark workload, the two layout parameters of mba_options --
``point_rows_per_lane`` (rows each lane of the point-major kernels takes,
which fixes the lanes per point of every track-length class) and
``camera_chunk_rows`` (rows per camera-major chunk, one workgroup's partial
sums) -- and times with HIP events on the context's stream:

* the mixed-precision Jacobian + Schur pass (point pass, camera pass,
  finalize), median of ``reps`` passes;
* one implicit-Schur PCG iteration (point matvec, camera matvec, finalize
  and scalar updates), from a short mixed LM solve: linear-solver time over
  CG iterations.

The score of a shape is ``pass + cg_per_pass * cg`` (the mixed LM runs
about ``cg_per_pass`` CG iterations per Jacobian pass at the benchmark's
settings).  The best shape per size class is written as the table the
library compiles in (``csrc/mba_tuning.h``) and the full sweep as JSON::

    python 3dreconstruction_amd/autotune.py --out profiles/autotune_mi355x.json \\
        --header 3dreconstruction_amd/csrc/mba_tuning.h

Rebuild the library afterwards for the new defaults to take effect.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

ROWS_PER_LANE = (1, 2, 4, 8)
CHUNK_ROWS = (256, 384, 512, 768, 1024, 1536)
# (name, scene preset, Huber threshold, upper bound of the size class in
# observations for the emitted table)
WORKLOADS = [("adjoint_check_697k", 2, 0.0, 1 << 21),
             ("mixed_precision_4.0M", 3, 4.0, 1 << 23),
             ("large_scene_25M", 5, 0.0, None)]


def _time_shape(mba, d, huber, rpl, cr, reps, cg_iters):
    def problem():
        return mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"])
    opts = mba.default_options(precision=mba.PRECISION_MIXED, huber_a=huber, linear_solver=mba.IMPLICIT_SCHUR_PCG,
                               point_rows_per_lane=rpl, camera_chunk_rows=cr)
    with mba.Context(problem(), opts) as ctx:
        ctx.jacobian_schur(1e4)
        ctx.set_timing(True)
        t = []
        for _ in range(reps):
            ctx.jacobian_schur(1e4)
            t.append(ctx.last_timing()[:3])
        ctx.set_timing(False)
    pass_ms = statistics.median(x[2] for x in t)
    point_ms = statistics.median(x[0] for x in t)
    cam_ms = statistics.median(x[1] for x in t)
    # PCG iteration: a short LM solve with the Nash-Sofer test off and a
    # fixed CG count per step.
    o = mba.default_options(precision=mba.PRECISION_MIXED, huber_a=huber, linear_solver=mba.IMPLICIT_SCHUR_PCG,
                            point_rows_per_lane=rpl, camera_chunk_rows=cr, max_num_iterations=2, eta=1e-30,
                            min_linear_solver_iterations=cg_iters, max_linear_solver_iterations=cg_iters)
    s = mba.solve(problem(), o)
    cg_ms = 1e3 * s.linear_solver_time_s / max(1, s.num_linear_solver_iterations)
    return {"point_rows_per_lane": rpl, "camera_chunk_rows": cr, "pass_ms": pass_ms, "point_pass_ms": point_ms,
            "camera_pass_ms": cam_ms, "cg_iteration_ms": cg_ms}


def tune(reps=15, cg_iters=20, cg_per_pass=10.0, workloads=None, log=print):
    mba = importlib.import_module("3dreconstruction_amd")
    out = {"score": f"pass_ms + {cg_per_pass} * cg_iteration_ms", "workloads": []}
    for name, preset, huber, bound in workloads or WORKLOADS:
        d = mba.generate_scene(preset)
        nobs = int(d["obs"].shape[0])
        log(f"autotune: {name} ({nobs} observations)")
        rows = []
        for rpl in ROWS_PER_LANE:
            for cr in CHUNK_ROWS:
                t0 = time.perf_counter()
                r = _time_shape(mba, d, huber, rpl, cr, reps, cg_iters)
                r["score"] = r["pass_ms"] + cg_per_pass * r["cg_iteration_ms"]
                rows.append(r)
                log(f"  rows/lane {rpl} chunk {cr:5d}: pass {r['pass_ms']:.4f} ms (point {r['point_pass_ms']:.4f}, "
                    f"camera {r['camera_pass_ms']:.4f}), CG iteration {r['cg_iteration_ms']:.4f} ms, "
                    f"score {r['score']:.4f} [{time.perf_counter() - t0:.1f} s]")
        best = min(rows, key=lambda r: r["score"])
        out["workloads"].append({"name": name, "observations": nobs, "size_class_max_rows": bound,
                                 "best": best, "sweep": rows})
        log(f"  best: rows/lane {best['point_rows_per_lane']}, chunk {best['camera_chunk_rows']}")
        del d
    return out


def write_header(result, path):
    lines = []
    for w in sorted(result["workloads"], key=lambda w: w["observations"]):
        bound = w["size_class_max_rows"]
        b = w["best"]
        lim = "INT64_MAX" if bound is None else f"int64_t({bound})"
        lines.append(f"    {{{lim}, {b['point_rows_per_lane']}, {b['camera_chunk_rows']}}},  // {w['name']}")
    text = f"""// Launch shapes of the point-major and camera-major kernels by problem size,
// as chosen on MI355X by 3dreconstruction_amd/autotune.py (the sweep and
// its timings are recorded in profiles/autotune_mi355x.json).  Used when
// mba_options leaves point_rows_per_lane / camera_chunk_rows at 0.
#ifndef MBA_TUNING_H_
#define MBA_TUNING_H_

#include <cstdint>

namespace mba {{

struct TunedShape {{
  int64_t max_rows;             // applies to problems with at most this many observations
  int32_t point_rows_per_lane;  // rows per lane of the point-major kernels
  int32_t camera_chunk_rows;    // rows per camera-major chunk (0: ~2048 chunks of 64..2048 rows)
}};

constexpr TunedShape kTunedShapes[] = {{
    {{int64_t(1) << 18, 4, 0}},
{chr(10).join(lines)}
}};

inline TunedShape tuned_shape(int64_t rows) {{
  for (const TunedShape& t : kTunedShapes)
    if (rows <= t.max_rows) return t;
  return kTunedShapes[sizeof(kTunedShapes) / sizeof(kTunedShapes[0]) - 1];
}}

}}  // namespace mba

#endif  // MBA_TUNING_H_
"""
    with open(path, "w") as f:
        f.write(text)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="JSON file for the sweep")
    ap.add_argument("--header", default=None, help="write the tuned table (csrc/mba_tuning.h)")
    ap.add_argument("--reps", type=int, default=15)
    ap.add_argument("--cg-iters", type=int, default=20)
    ap.add_argument("--only", default=None, help="comma-separated workload names")
    a = ap.parse_args()
    wl = WORKLOADS if not a.only else [w for w in WORKLOADS if w[0] in a.only.split(",")]
    res = tune(reps=a.reps, cg_iters=a.cg_iters, workloads=wl, log=lambda m: print(m, flush=True))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)
    if a.header:
        write_header(res, a.header)


if __name__ == "__main__":
    main()
