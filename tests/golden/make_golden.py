"""Writes the golden fixtures of this directory from the reference's own Ceres
Solver 1.10 (built from the reference sources by oracle/ref/Makefile into
oracle/_ref/libceres_ref.so).  Runs only in a container that has the reference
checkout; the fixtures it writes are plain numbers (inputs are regenerated
from the seeded scene generator, so only seeds, indices and outputs are
stored).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

# Small scenes whose structure exercises the ordering beyond "all points E":
# "tiny" has a camera of degree 2 that the independent set picks as an E
# block; "const" fixes one camera (SetParameterBlockConstant).
HUBER_SPEC = dict(num_cameras=30, num_points=3000, track_min=2, track_max=10, power_alpha=1.5,
                  outlier_frac=0.05, seed=11)


# The incremental engine's layout: cameras [aa, t], three shared intrinsic
# groups [f, ppx, ppy, k1, k2, k3], residual blocks (intrinsics, camera,
# point) (libs/sfm/src/sfm_incremental_engine.cpp:1413-1446).
GROUP_SPEC = dict(num_cameras=24, num_points=1500, track_min=2, track_max=8, outlier_frac=0.05, seed=31)
GROUPS = 3
GROUP_CONST = 2  # SetParameterBlockConstant on this camera's extrinsic block
# SubsetParameterization masks per group (bit k = entry k of [f, ppx, ppy,
# k1, k2, k3] held constant) and per-group constant flags.
GROUP_SUBSET = np.array([0x3E, 0x38, 0x00], np.uint8)
GROUP_INTR_CONST = np.array([0, 1, 0], np.uint8)


def group_scene():
    return O.generate_groups(O.make_spec(**GROUP_SPEC), GROUPS)


def group_const():
    c = np.zeros(GROUP_SPEC["num_cameras"], np.uint8)
    c[GROUP_CONST] = 1
    return c


GLOBAL_FOCAL = 500.0
TWO_VIEW_PAIRS = [(0, 1), (2, 3), (4, 5), (6, 7), (8, 9), (10, 11), (12, 13), (14, 15)]


def global_scene():
    """The global engine's BA problems (ComputeResidual: pinhole, one fixed
    focal, no distortion; libs/sfm/include/mvg/sfm/pinhole_ceres_functor.h)
    on the 16-camera scene's geometry: observations are that pinhole's
    projections of the true points plus the scene's own noise and outliers."""
    d = O.generate(O.preset(1))
    r_true = O.evaluate(O.Scene(d, cameras=d["cam_true"], points=d["pt_true"]), jacobians=False)[0]
    pin = d["cam_true"].copy()
    pin[:, 6] = GLOBAL_FOCAL
    pin[:, 7:] = 0.0
    r_pin = O.evaluate(O.Scene(d, cameras=pin, points=d["pt_true"]), jacobians=False)[0]
    obs = (r_pin + d["obs"]) - r_true  # pinhole prediction + (observed - packed prediction)
    return dict(obs=obs, cam_idx=d["cam_idx"], pt_idx=d["pt_idx"], pt_init=d["pt_init"],
                cam_true=d["cam_true"][:, :6].copy(), cam_init=d["cam_init"][:, :6].copy(), focal=GLOBAL_FOCAL)


def two_view_problems(g):
    """One 2-view problem per camera pair (the global engine's pair loop,
    sfm_global_engine.cpp:985-1174): the points both cameras see and their
    observations in both, cameras renumbered 0 and 1."""
    out = []
    for a, b in TWO_VIEW_PAIRS:
        ia, ib = g["cam_idx"] == a, g["cam_idx"] == b
        common = np.intersect1d(g["pt_idx"][ia], g["pt_idx"][ib])
        rows = np.flatnonzero((ia | ib) & np.isin(g["pt_idx"], common))
        remap = {int(q): k for k, q in enumerate(common)}
        out.append(dict(obs=g["obs"][rows].copy(), cam_idx=(g["cam_idx"][rows] == b).astype(np.int32),
                        pt_idx=np.array([remap[int(q)] for q in g["pt_idx"][rows]], np.int32),
                        pt_init=g["pt_init"][common].copy(), cam_init=g["cam_init"][[a, b]].copy()))
    return out


def tiny_scene():
    """Hand-built: 4 cameras, 6 points; camera 3 sees only points 0 and 1."""
    spec = O.make_spec(3, 6, 2, 3, seed=5)
    d = O.generate(spec)
    cams = np.vstack([d["cam_init"], d["cam_init"][0] * np.array([1, 1, 1, 1.1, 1, 1, 1, 1, 1])])
    extra_obs = np.array([[10.0, -5.0], [3.0, 7.0]])
    d2 = dict(cam_init=cams, pt_init=d["pt_init"],
              obs=np.vstack([d["obs"], extra_obs]),
              cam_idx=np.concatenate([d["cam_idx"], [3, 3]]).astype(np.int32),
              pt_idx=np.concatenate([d["pt_idx"], [0, 1]]).astype(np.int32))
    return d2


def solve_fixture(d, huber, linear_solver, iters, const=None):
    sc = O.Scene(d, cam_constant=const)
    r = O.ref_solve(sc, huber_a=huber, linear_solver=linear_solver, max_iters=iters)
    return dict(initial_cost=r["initial_cost"], final_cost=r["final_cost"], successful=r["successful"],
                unsuccessful=r["unsuccessful"], termination=r["termination"], iterations=r["iterations"],
                cameras=sc.cameras, points=sc.points)


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libceres_ref.so is not built (make -C oracle/ref)")
    out = {}
    # 1. LM trajectories (SPARSE_SCHUR + CX_SPARSE as libs/sfm sets it, and
    #    ITERATIVE_SCHUR + JACOBI) on the 16-camera scene, 20 iterations.
    d1 = O.generate(O.preset(1))
    for name, ls in [("sparse", 0), ("iterative", 1)]:
        f = solve_fixture(d1, 0.0, ls, 20)
        for k, v in f.items():
            out[f"c1_{name}_{k}"] = np.asarray(v)
    # 2. Huber (a = 4 px) with 5% outliers, both solvers; a constant camera.
    dh = O.generate(O.make_spec(**HUBER_SPEC))
    for name, ls in [("sparse", 0), ("iterative", 1)]:
        f = solve_fixture(dh, 4.0, ls, 20)
        for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
            out[f"huber_{name}_{k}"] = np.asarray(f[k])
    const = np.zeros(30, np.uint8)
    const[3] = 1
    f = solve_fixture(dh, 4.0, 0, 20, const)
    for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
        out[f"const_{k}"] = np.asarray(f[k])
    # 3. Structure: Schur ordering, residual order, S block pairs.
    for name, d, cc in [("c1", d1, None), ("huber", dh, const), ("tiny", tiny_scene(), None)]:
        ne, order, res, pairs = O.ref_schur_ordering(O.Scene(d, cam_constant=cc))
        out[f"struct_{name}_num_e"] = np.asarray(ne)
        out[f"struct_{name}_order"] = order
        out[f"struct_{name}_residual_order"] = res.astype(np.int32)
        out[f"struct_{name}_pairs"] = pairs
    # 4. Ceres autodiff Jets on 400 seeded observations of the 700k scene.
    d2 = O.generate(O.preset(2))
    idx = np.sort(np.random.default_rng(2).choice(d2["num_obs"], 400, replace=False))
    R, JC, JP = [], [], []
    for i in idx:
        c = d2["cam_init"][d2["cam_idx"][i]]
        p = d2["pt_init"][d2["pt_idx"][i]]
        r, jc, jp = O.ref_evaluate_obs(0, c, p, d2["obs"][i])
        R.append(r)
        JC.append(jc)
        JP.append(jp)
    out["jets_c2_index"] = idx
    out["jets_c2_residual"] = np.array(R)
    out["jets_c2_jac_camera"] = np.array(JC)
    out["jets_c2_jac_point"] = np.array(JP)
    # 5. Grouped intrinsics (layout of the incremental engine).
    dg = group_scene()
    ne, order, res, pairs = O.ref_schur_ordering(O.group_scene(dg, cam_constant=group_const()))
    out["grp_struct_num_e"] = np.asarray(ne)
    out["grp_struct_order"] = order
    out["grp_struct_residual_order"] = res.astype(np.int32)
    out["grp_struct_pairs"] = pairs
    gidx = np.sort(np.random.default_rng(3).choice(dg["num_obs"], 200, replace=False))
    R, JI, JC, JP = [], [], [], []
    for i in gidx:
        r, ji, jc, jp = O.ref_evaluate_obs(1, dg["cam_init"][dg["cam_idx"][i]], dg["pt_init"][dg["pt_idx"][i]],
                                           dg["obs"][i], dg["intr_init"][dg["intr_idx"][i]])
        R.append(r)
        JI.append(ji)
        JC.append(jc)
        JP.append(jp)
    out["grp_jets_index"] = gidx
    out["grp_jets_residual"] = np.array(R)
    out["grp_jets_jac_intrinsics"] = np.array(JI)
    out["grp_jets_jac_camera"] = np.array(JC)
    out["grp_jets_jac_point"] = np.array(JP)
    for name, huber, ls, cc in [("sparse", 4.0, 0, group_const()), ("iterative", 0.0, 1, None),
                                ("iterative_const", 4.0, 1, group_const())]:
        sc = O.group_scene(dg, cam_constant=cc)
        r = O.ref_solve(sc, huber_a=huber, linear_solver=ls, max_iters=20)
        for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
            out[f"grp_{name}_{k}"] = np.asarray(r[k])
        out[f"grp_{name}_cameras"] = sc.cameras
        out[f"grp_{name}_intrinsics"] = sc.intrinsics
        out[f"grp_{name}_points"] = sc.points
    # 6. The incremental engine's intrinsic options on the grouped scene:
    #    SubsetParameterization masks (0x3E = focal only, its default;
    #    0x38 = k1..k3 held) and a constant group, Huber 4 px, the constant
    #    camera, 20 iterations.
    for name, ls, sub, ic in [("subset_sparse", 0, GROUP_SUBSET, None),
                              ("subset_iterative", 1, np.full(GROUPS, 0x3E, np.uint8), None),
                              ("const_group_sparse", 0, None, GROUP_INTR_CONST)]:
        sc = O.group_scene(dg, cam_constant=group_const(), intr_constant=ic, intr_subset=sub)
        r = O.ref_solve(sc, huber_a=4.0, linear_solver=ls, max_iters=20)
        for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
            out[f"grp_{name}_{k}"] = np.asarray(r[k])
        out[f"grp_{name}_cameras"] = sc.cameras
        out[f"grp_{name}_intrinsics"] = sc.intrinsics
        out[f"grp_{name}_points"] = sc.points
    # 7. The global engine's functors: PinholeReprojectionError_Rt <2,6,3>
    #    (final BA, HuberLoss(Square(2.0)), SPARSE_SCHUR), PinholeReprojectionError_t
    #    <2,3,3> (translation only, rotations fixed at the true ones,
    #    HuberLoss(Square(4.0))), and the 2-view DENSE_SCHUR problems.
    g = global_scene()
    foc = np.array([g["focal"]])
    sc = O.Scene(dict(cam_init=g["cam_init"], pt_init=g["pt_init"], obs=g["obs"], cam_idx=g["cam_idx"],
                      pt_idx=g["pt_idx"]), layout=2, intrinsics=foc)
    r = O.ref_solve(sc, huber_a=4.0, linear_solver=0, max_iters=20)
    for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
        out[f"rt_sparse_{k}"] = np.asarray(r[k])
    out["rt_sparse_cameras"] = sc.cameras
    sc = O.Scene(dict(cam_init=g["cam_init"][:, 3:].copy(), pt_init=g["pt_init"], obs=g["obs"],
                      cam_idx=g["cam_idx"], pt_idx=g["pt_idx"]), layout=3,
                 intrinsics=np.concatenate([foc, g["cam_true"][:, :3].ravel()]))
    r = O.ref_solve(sc, huber_a=16.0, linear_solver=0, max_iters=20)
    for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
        out[f"tonly_sparse_{k}"] = np.asarray(r[k])
    out["tonly_sparse_translations"] = sc.cameras
    costs, succ, term, nit = [], [], [], []
    for k, tv in enumerate(two_view_problems(g)):
        sc = O.Scene(tv, layout=2, intrinsics=foc)
        r = O.ref_solve(sc, huber_a=4.0, linear_solver=2, max_iters=50)
        costs.append([r["initial_cost"], r["final_cost"]])
        succ.append([r["successful"], r["unsuccessful"]])
        term.append(r["termination"])
        nit.append(len(r["iterations"]))
        out[f"twoview_{k}_iteration_costs"] = r["iterations"][:, 1].copy()
        out[f"twoview_{k}_cameras"] = sc.cameras
    out["twoview_costs"] = np.array(costs)
    out["twoview_steps"] = np.array(succ)
    out["twoview_termination"] = np.array(term)
    path = os.path.join(HERE, "ceres_ref.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
