"""Writes the golden fixtures of this directory from the reference's own Ceres
Solver 1.10 (built from the reference sources by oracle/ref/Makefile into
oracle/_ref/libceres_ref.so).  Runs only in a container that has the reference
checkout; the fixtures it writes are plain numbers (inputs are regenerated
from the seeded scene generator, so only seeds, indices and outputs are
stored).

    python tests/golden/make_golden.py          # ceres_ref.npz
    python tests/golden/make_golden.py --large  # ceres_ref_large.npz
    python tests/golden/make_golden.py --lm     # ceres_ref_lm.npz
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


LIBMV_DIR = os.path.join(HERE, "libmv")
LIBMV_FILES = ["problem_01.bin", "problem_02.bin", "problem_03.bin"]
# libmv_bundle_adjuster.cc:817-836 with pixel markers refines f, k1, k2:
# ppx, ppy, k3 (and p1, p2, zero here) held -> SubsetParameterization mask.
LIBMV_SUBSET = 0x26
BAL_FILE = os.path.join(HERE, "bal_small.txt")


def libmv_problem(path):
    """A libmv problem file as grouped arrays, built from an independent
    parse of the file; rotations converted by the reference's
    RotationMatrixToAngleAxis.  Cameras by ascending image, points by
    ascending track (the address order of libmv's vectors), markers in file
    order; the first marker's camera is held constant
    (libmv_bundle_adjuster.cc:700-725)."""
    f = O.parse_libmv(path)
    cam = {int(c["image"]): c for c in f["cams"] if c["image"] >= 0}
    pt = {int(p["track"]): p for p in f["pts"] if p["track"] >= 0}
    images, tracks = sorted(cam), sorted(pt)
    ci = {im: k for k, im in enumerate(images)}
    pi = {tr: k for k, tr in enumerate(tracks)}
    keep = [m for m in f["markers"] if int(m["image"]) in ci and int(m["track"]) in pi]
    R = np.array([np.asarray(cam[im]["R"], np.float64).reshape(3, 3).T for im in images])  # column-major file
    aa = O.ref_rotation_matrix_to_angle_axis(R)
    t = np.array([np.asarray(cam[im]["t"], np.float64) for im in images])
    const = np.zeros(len(images), np.uint8)
    const[ci[int(keep[0]["image"])]] = 1
    return dict(cam_init=np.concatenate([aa, t], axis=1), intr_init=f["intr8"][None, :6].copy(),
                pt_init=np.array([np.asarray(pt[tr]["X"], np.float64) for tr in tracks]),
                obs=np.array([[m["x"], m["y"]] for m in keep], np.float64),
                cam_idx=np.array([ci[int(m["image"])] for m in keep], np.int32),
                pt_idx=np.array([pi[int(m["track"])] for m in keep], np.int32),
                intr_idx=np.zeros(len(keep), np.int32), const=const, R=R)


def rotation_cases():
    """Rotation matrices for the angle-axis conversion: random, identity,
    tiny angles, exactly and nearly pi about axes (the branches of the
    quaternion extraction)."""
    rng = np.random.default_rng(8)
    def rodrigues(w):
        th = np.linalg.norm(w)
        if th == 0:
            return np.eye(3)
        k = w / th
        K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
        return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K
    ws = [rng.normal(size=3) * s for s in (0.1, 1.0, 2.0, 3.0) for _ in range(8)]
    ws += [np.zeros(3), np.array([1e-9, -2e-9, 3e-9]), np.array([np.pi, 0, 0]), np.array([0, np.pi, 0]),
           np.array([0, 0, np.pi]), np.array([0, 0, np.pi - 1e-7]), np.array([np.pi / 2, 0, 0])]
    ws += [rng.normal(size=3) / np.linalg.norm(rng.normal(size=3)) for _ in range(2)]
    axis = rng.normal(size=3)
    ws.append(axis / np.linalg.norm(axis) * (np.pi - 1e-9))
    return np.array([rodrigues(w) for w in ws])


def write_bal_scene():
    """A small packed scene as a BAL file (BAL cameras look down -z, so the
    BAL observations are the negated libs/sfm ones)."""
    d = O.generate(O.make_spec(8, 300, 2, 5, outlier_frac=0.05, seed=21))
    with open(BAL_FILE, "w") as fo:
        fo.write(f"{d['num_cameras']} {d['num_points']} {d['num_obs']}\n")
        for i in range(d["num_obs"]):
            fo.write("%d %d %.17g %.17g\n" % (d["cam_idx"][i], d["pt_idx"][i], -d["obs"][i, 0], -d["obs"][i, 1]))
        for v in np.concatenate([d["cam_init"].ravel(), d["pt_init"].ravel()]):
            fo.write("%.17g\n" % v)


# Trust-region problem dumps (TEXTFILE): small scenes so that the A files
# stay a few thousand entries.  Packed: Huber 4 px, 10% outliers, camera 2
# constant.  Grouped: three groups with GROUP_SUBSET masks, group 1
# constant, camera 2 constant.
DUMP_SPEC = dict(num_cameras=6, num_points=60, track_min=2, track_max=5, outlier_frac=0.1, seed=43)
DUMP_GROUP_SPEC = dict(num_cameras=8, num_points=50, track_min=2, track_max=5, outlier_frac=0.1, seed=44)
DUMP_ITERATIONS = [0, 1, 3]
DUMP_CONST_CAMERA = 2


def dump_scene():
    return O.generate(O.make_spec(**DUMP_SPEC))


def dump_group_scene():
    return O.generate_groups(O.make_spec(**DUMP_GROUP_SPEC), GROUPS)


def dump_const(n):
    c = np.zeros(n, np.uint8)
    c[DUMP_CONST_CAMERA] = 1
    return c


def read_dump(base):
    """The reference's TEXTFILE dump files as arrays, the .m text as is."""
    a = np.loadtxt(base + "_A.txt", ndmin=2)
    return dict(rows=a[:, 0].astype(np.int32), cols=a[:, 1].astype(np.int32), values=a[:, 2],
                D=np.loadtxt(base + "_D.txt"), b=np.loadtxt(base + "_b.txt"), x=np.loadtxt(base + "_x.txt"),
                m=open(base + ".m").read())


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
    # 8. Boundary conversions and file formats.
    Rs = rotation_cases()
    out["rot_R"] = Rs
    out["rot_aa"] = O.ref_rotation_matrix_to_angle_axis(Rs)
    write_bal_scene()
    import hashlib
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        tmp = os.path.join(td, "rt.txt")
        O.ref_bal_roundtrip(BAL_FILE, tmp)
        out["bal_write_sha256"] = np.array(hashlib.sha256(open(tmp, "rb").read()).hexdigest())
    r = O.ref_bal_solve(BAL_FILE, huber_a=4.0, linear_solver=0, max_iters=20)
    for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations", "cameras",
              "points"]:
        out[f"bal_sparse_{k}"] = np.asarray(r[k])
    O.ref_set_points_first(True)
    try:
        for k, fn in enumerate(LIBMV_FILES):
            lp = libmv_problem(os.path.join(LIBMV_DIR, fn))
            out[f"libmv_{k}_aa"] = lp["cam_init"][:, :3].copy()
            sc = O.group_scene(lp, cam_constant=lp["const"], intr_subset=np.array([LIBMV_SUBSET], np.uint8))
            ne, order, res, pairs = O.ref_schur_ordering(sc)
            out[f"libmv_{k}_struct_num_e"] = np.asarray(ne)
            out[f"libmv_{k}_struct_order"] = order
            out[f"libmv_{k}_struct_residual_order"] = res.astype(np.int32)
            # ~10^5 pairs (every camera couples through the few points):
            # their count and digest instead of the list.
            out[f"libmv_{k}_struct_num_pairs"] = np.asarray(pairs.shape[0])
            out[f"libmv_{k}_struct_pairs_sha256"] = np.array(
                hashlib.sha256(np.ascontiguousarray(pairs, np.int32).tobytes()).hexdigest())
            sc = O.group_scene(lp, cam_constant=lp["const"], intr_subset=np.array([LIBMV_SUBSET], np.uint8))
            r = O.ref_solve(sc, huber_a=0.0, linear_solver=0, max_iters=20)
            for key in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
                out[f"libmv_{k}_{key}"] = np.asarray(r[key])
            out[f"libmv_{k}_cameras"] = sc.cameras
            out[f"libmv_{k}_intrinsics"] = sc.intrinsics
            out[f"libmv_{k}_points"] = sc.points
    finally:
        O.ref_set_points_first(False)
    # 9. ITERATIVE_SCHUR with SCHUR_JACOBI (exact diagonal blocks of S).
    O.ref_set_schur_jacobi(True)
    try:
        for name, d, huber, cc in [("c1_schur_jacobi", d1, 0.0, None), ("huber_schur_jacobi", dh, 4.0, const)]:
            sc = O.Scene(d, cam_constant=cc)
            r = O.ref_solve(sc, huber_a=huber, linear_solver=1, max_iters=20)
            for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
                out[f"{name}_{k}"] = np.asarray(r[k])
            out[f"{name}_cameras"] = sc.cameras
    finally:
        O.ref_set_schur_jacobi(False)
    # 10. Trust-region problem dumps (TEXTFILE) of SPARSE_SCHUR solves, Huber
    #     4 px: the scaled, robustified Jacobian of the reduced program, the LM
    #     diagonal, the residuals and the step at DUMP_ITERATIONS; the .m
    #     loader with the directory replaced by {DIR}.
    import tempfile
    for name, sc in [("dump", O.Scene(dump_scene(), cam_constant=dump_const(DUMP_SPEC["num_cameras"]))),
                     ("dump_grp", O.group_scene(dump_group_scene(),
                                                cam_constant=dump_const(DUMP_GROUP_SPEC["num_cameras"]),
                                                intr_constant=GROUP_INTR_CONST, intr_subset=GROUP_SUBSET))]:
        tmp = tempfile.mkdtemp()
        O.ref_set_dump(tmp, DUMP_ITERATIONS)
        try:
            r = O.ref_solve(sc, huber_a=4.0, linear_solver=0, max_iters=6)
        finally:
            O.ref_set_dump(None, [])
        for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
            out[f"{name}_{k}"] = np.asarray(r[k])
        for it in DUMP_ITERATIONS:
            dd = read_dump(os.path.join(tmp, f"ceres_solver_iteration_{it:03d}"))
            for k in ["rows", "cols", "values", "D", "b", "x"]:
                out[f"{name}_{it}_{k}"] = dd[k]
            out[f"{name}_{it}_m"] = np.asarray(dd["m"].replace(tmp, "{DIR}"))
    # 11. CLUSTER_JACOBI / CLUSTER_TRIDIAGONAL: the reference's Ceres is built
    #     without SuiteSparse, so Solve rejects them (solver.cc:135-149) for
    #     either linear solver: termination type and message.
    dc = dump_scene()
    for kind, pname in [(2, "cluster_jacobi"), (3, "cluster_tridiagonal")]:
        O.ref_set_preconditioner(kind)
        try:
            for ls in (0, 1):
                r = O.ref_solve(O.Scene(dc), huber_a=4.0, linear_solver=ls, max_iters=5)
                out[f"{pname}_{ls}_termination"] = np.asarray(r["termination"])
                out[f"{pname}_{ls}_usable"] = np.asarray(r["usable"])
                out[f"{pname}_{ls}_num_iterations"] = np.asarray(len(r["iterations"]))
                out[f"{pname}_{ls}_message"] = np.asarray(O.ref_last_message())
        finally:
            O.ref_set_preconditioner(0)
    path = os.path.join(HERE, "ceres_ref.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


# Full-size trajectories of the benchmark scenes (one thread, so the run is
# reproducible bit for bit): the adjoint-check scene (200 cameras, 100,000
# points, 697k observations) with SPARSE_SCHUR + CX_SPARSE, no loss, 20 LM
# iterations (the libs/sfm solver), and the mixed-precision scene (1,000
# cameras, 500,000 points, 4.0M observations, 5% 20-px outliers) with
# Huber 4 px and ITERATIVE_SCHUR, JACOBI and SCHUR_JACOBI, capped at 5 LM
# iterations (CXSparse on its dense 9,000^2 S would take hours).
LARGE_CASES = [("c2_sparse", 2, 0.0, 0, 20, 0), ("c3_iterative", 3, 4.0, 1, 5, 0),
               ("c3_schur_jacobi", 3, 4.0, 1, 5, 1)]
LARGE_POINT_SAMPLE = 2000


def main_large():
    """Writes ceres_ref_large.npz: per case the summary, the iteration
    records (iteration, cost, cost_change, radius, successful, valid, CG
    iterations, relative_decrease, gradient_max_norm), the final cameras
    and the final positions of a seeded sample of points."""
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libceres_ref.so is not built (make -C oracle/ref)")
    out = {}
    for name, preset, huber, ls, iters, sj in LARGE_CASES:
        d = O.generate(O.preset(preset))
        sc = O.Scene(d)
        O.ref_set_schur_jacobi(bool(sj))
        try:
            r = O.ref_solve(sc, huber_a=huber, linear_solver=ls, max_iters=iters, num_threads=1)
        finally:
            O.ref_set_schur_jacobi(False)
        for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
            out[f"{name}_{k}"] = np.asarray(r[k])
        idx = np.sort(np.random.default_rng(preset).choice(d["num_points"], LARGE_POINT_SAMPLE, replace=False))
        out[f"{name}_cameras"] = sc.cameras
        out[f"{name}_point_index"] = idx
        out[f"{name}_points"] = sc.points[idx]
        print(name, r["final_cost"], len(r["iterations"]), flush=True)
    path = os.path.join(HERE, "ceres_ref_large.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


# Whole LM solves to convergence at the benchmark's LM settings (PCG forcing
# eta 1e-2, function tolerance 1e-12, at most 300 iterations) on the
# adjoint-check scene (200 cameras, 100,000 points, 697k observations, no
# loss) with ITERATIVE_SCHUR and JACOBI or SCHUR_JACOBI, one thread.
LM_CASES = [("c2_lm_jacobi", 2, 0.0, 0), ("c2_lm_schur_jacobi", 2, 0.0, 1)]
LM_ETA, LM_FUNCTION_TOLERANCE, LM_MAX_ITERATIONS = 1e-2, 1e-12, 300


def main_lm():
    """Writes ceres_ref_lm.npz: per case the summary and the iteration
    records (as ceres_ref_large.npz) of the whole solve."""
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libceres_ref.so is not built (make -C oracle/ref)")
    out = {}
    for name, preset, huber, precond in LM_CASES:
        sc = O.Scene(O.generate(O.preset(preset)))
        O.ref_set_preconditioner(precond)
        O.ref_set_lm_tolerances(LM_ETA, LM_FUNCTION_TOLERANCE)
        try:
            r = O.ref_solve(sc, huber_a=huber, linear_solver=1, max_iters=LM_MAX_ITERATIONS, num_threads=1)
        finally:
            O.ref_set_preconditioner(0)
            O.ref_set_lm_tolerances()
        for k in ["initial_cost", "final_cost", "successful", "unsuccessful", "termination", "iterations"]:
            out[f"{name}_{k}"] = np.asarray(r[k])
        print(name, r["final_cost"], len(r["iterations"]), r["termination"], flush=True)
    path = os.path.join(HERE, "ceres_ref_lm.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    if "--large" in sys.argv:
        main_large()
    elif "--lm" in sys.argv:
        main_lm()
    else:
        main()
