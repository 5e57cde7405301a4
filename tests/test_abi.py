"""C ABI checks that need no GPU: the library loads, exports every symbol the
header declares, keeps the reference defaults, and its host-side
preprocessing reproduces the reference's Schur structure bit for bit."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, ba_problem, grouped_problem


def header_symbols():
    text = open(os.path.join(ROOT, "include", "mba.h")).read()
    return sorted(set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(mba_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(mba):
    lib = mba.load_library()
    syms = header_symbols()
    assert len(syms) >= 18
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(mba.EXPORTS)
    assert lib.mba_abi_version() == 4
    # mba_problem: 4 int32, 1 int64, 10 pointers, 2 int32 (x86-64 layout).
    assert ctypes.sizeof(mba._Problem) == 4 * 4 + 8 + 10 * 8 + 2 * 4


def test_struct_layouts_match_the_header(mba, tmp_path):
    """sizeof and field offsets of the ctypes mirrors against the C compiler's
    reading of include/mba.h."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    checks = {"mba_options": (mba.Options, ["eta", "preconditioner", "dump_directory", "dump_iterations",
                                            "num_dump_iterations", "point_rows_per_lane", "camera_chunk_rows",
                                            "pcg_graph_launch"]),
              "mba_problem": (mba._Problem, []), "mba_summary": (mba._Summary, ["num_linear_solver_iterations"]),
              "mba_iteration": (mba.Iteration, ["cumulative_time_s"])}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "mba.h"', "int main(void) {"]
    for cname, (_, fields) in checks.items():
        src.append(f'  printf("%zu\\n", sizeof({cname}));')
        for f in fields:
            src.append(f'  printf("%zu\\n", offsetof({cname}, {f}));')
    src.append("  return 0;\n}")
    (tmp_path / "l.c").write_text("\n".join(src))
    exe = tmp_path / "l"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(tmp_path / "l.c"), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]
    want = []
    for cname, (cls, fields) in checks.items():
        want.append(ctypes.sizeof(cls))
        want.extend(getattr(cls, f).offset for f in fields)
    assert got == want


def test_default_options_are_the_reference_defaults(mba):
    """include/ceres/solver.h:62-129."""
    o = mba.default_options()
    assert o.max_num_iterations == 50
    assert o.max_num_consecutive_invalid_steps == 5
    assert o.function_tolerance == 1e-6 and o.gradient_tolerance == 1e-10 and o.parameter_tolerance == 1e-8
    assert o.initial_trust_region_radius == 1e4 and o.max_trust_region_radius == 1e16
    assert o.min_trust_region_radius == 1e-32 and o.min_relative_decrease == 1e-3
    assert o.min_lm_diagonal == 1e-6 and o.max_lm_diagonal == 1e32
    assert o.eta == 0.1 and o.pcg_r_tolerance == -1.0 and o.max_linear_solver_iterations == 500
    assert o.jacobi_scaling == 1


@pytest.mark.parametrize("name", ["c1", "huber", "tiny"])
def test_schur_structure_bit_exact_with_reference(mba, golden, c1, huber_scene, tiny_scene, name):
    d = {"c1": c1, "huber": huber_scene, "tiny": tiny_scene}[name]
    const = None
    if name == "huber":
        const = np.zeros(30, np.uint8)
        const[3] = 1
    ne, order, res, pairs = mba.schur_structure(ba_problem(mba, d, const))
    assert ne == int(golden[f"struct_{name}_num_e"])
    np.testing.assert_array_equal(order, golden[f"struct_{name}_order"])
    np.testing.assert_array_equal(res, golden[f"struct_{name}_residual_order"])
    np.testing.assert_array_equal(pairs, golden[f"struct_{name}_pairs"])


@pytest.fixture(scope="module")
def group_scene(O):
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import group_const, group_scene as gs
    return gs(), group_const()


def test_grouped_schur_structure_bit_exact_with_reference(mba, golden, group_scene):
    """Residual blocks (intrinsics, camera, point) with shared groups and a
    constant camera: E/F order, residual order and S pairs are the
    reference's (block ids: groups, then cameras, then points)."""
    d, const = group_scene
    ne, order, res, pairs = mba.schur_structure(grouped_problem(mba, d, const))
    assert ne == int(golden["grp_struct_num_e"])
    np.testing.assert_array_equal(order, golden["grp_struct_order"])
    np.testing.assert_array_equal(res, golden["grp_struct_residual_order"])
    np.testing.assert_array_equal(pairs, golden["grp_struct_pairs"])


def test_grouped_problem_checks(mba, group_scene):
    d, _ = group_scene
    p = grouped_problem(mba, d)
    p.intrinsic_index = p.intrinsic_index.copy()
    p.intrinsic_index[5] = 7
    with pytest.raises(mba.MbaError) as e:
        mba.schur_structure(p)
    assert e.value.status == mba.ERR_INVALID_ARGUMENT and "intrinsic index" in str(e.value)
    # A camera observed through two groups is not the incremental engine's
    # structure.
    p = grouped_problem(mba, d)
    cam = int(p.camera_index[0])
    p.intrinsic_index = p.intrinsic_index.copy()
    rows = np.flatnonzero(p.camera_index == cam)
    p.intrinsic_index[rows[1]] = (p.intrinsic_index[rows[0]] + 1) % 3
    with pytest.raises(mba.MbaError) as e:
        mba.schur_structure(p)
    assert e.value.status == mba.ERR_UNSUPPORTED and "two intrinsic groups" in str(e.value)


def test_grouped_subset_mask_checks(mba, group_scene):
    """SubsetParameterization must hold 1..5 of the 6 entries
    (local_parameterization.cc:93-109): a full mask is an argument error."""
    d, _ = group_scene
    for bad in [0x3F, 0x40]:
        with pytest.raises(mba.MbaError) as e:
            mba.schur_structure(grouped_problem(mba, d, subset=np.array([bad, 0, 0], np.uint8)))
        assert e.value.status == mba.ERR_INVALID_ARGUMENT
    mba.schur_structure(grouped_problem(mba, d, subset=np.array([0x3E, 0x01, 0], np.uint8)))


def test_schur_structure_matches_oracle_on_700k_scene(mba, O):
    d = O.generate(O.preset(2))
    ne, order, res, pairs = mba.schur_structure(ba_problem(mba, d))
    ne2, order2, res2 = O.schur_ordering(O.Scene(d))
    assert ne == ne2
    np.testing.assert_array_equal(order, order2)
    np.testing.assert_array_equal(res, res2)
    np.testing.assert_array_equal(pairs, O.schur_block_pairs(O.Scene(d)))


@pytest.mark.parametrize("case", ["points_first", "camera_ties_points"])
def test_schur_structure_matches_oracle_on_edge_structures(mba, O, case):
    """The ordering's points-first construction (every point's degree below
    every camera's) and the general greedy path (a camera whose degree ties
    the longest tracks) against the oracle's restatement of Ceres's ordering,
    on a scene with repeated (camera, point) observations, constant cameras,
    a point seen only through constant cameras, and an unobserved camera and
    point."""
    rng = np.random.default_rng(11)
    d = dict(O.generate(O.make_spec(40, 3000, 2, 8, seed=21)))
    obs, ci, pi = d["obs"], d["cam_idx"], d["pt_idx"]
    dup = rng.choice(len(ci), 300, replace=False)  # repeated observations, shuffled in
    obs = np.concatenate([obs, obs[dup] + 0.5])
    ci = np.concatenate([ci, ci[dup]])
    pi = np.concatenate([pi, pi[dup]])
    const = np.zeros(41, np.uint8)
    const[[2, 17]] = 1
    # Point 0 seen only through constant cameras; camera 40 and point 3000 unobserved.
    rows0 = np.flatnonzero(pi == 0)
    ci[rows0] = 2
    cams = np.concatenate([d["cam_init"], d["cam_init"][:1]])
    pts = np.concatenate([d["pt_init"], d["pt_init"][:1]])
    if case == "camera_ties_points":
        # Camera 5 keeps three points (its shortest tracks; the others move to
        # camera 6): its degree no longer exceeds every point's, so the greedy
        # independent set decides, and a 2-observation point ahead of it in
        # the queue keeps it out of E.
        rows5 = np.flatnonzero(ci == 5)
        tl = np.bincount(pi)
        rows5 = rows5[np.argsort(tl[pi[rows5]], kind="stable")]
        ci[rows5[3:]] = 6
    perm = rng.permutation(len(ci))
    d2 = dict(d, cam_init=cams, pt_init=pts, obs=np.ascontiguousarray(obs[perm]),
              cam_idx=np.ascontiguousarray(ci[perm], np.int32), pt_idx=np.ascontiguousarray(pi[perm], np.int32))
    p = mba.BAProblem(cams.copy(), pts.copy(), d2["obs"], d2["cam_idx"], d2["pt_idx"], const)
    ne2, order2, res2 = O.schur_ordering(O.Scene(d2, cam_constant=const))
    try:
        ne, order, res, _ = mba.schur_structure(p)
    except mba.MbaError as e:  # a camera in E: the oracle must agree it is not the standard structure
        assert e.status == mba.ERR_UNSUPPORTED and np.any(order2[:ne2] < 41)
        return
    assert ne == ne2
    np.testing.assert_array_equal(order, order2)
    np.testing.assert_array_equal(res, res2)


def test_out_of_range_index_names_the_first_observation(mba, O):
    """The parallel index check reports the lowest offending observation."""
    d = O.generate(O.preset(2))
    p = ba_problem(mba, d)
    p.point_index = p.point_index.copy()
    p.camera_index = p.camera_index.copy()
    p.point_index[600000] = -1
    p.camera_index[123457] = 10 ** 6
    p.point_index[123458] = 10 ** 7
    with pytest.raises(mba.MbaError) as e:
        mba.schur_structure(p)
    assert e.value.status == mba.ERR_INVALID_ARGUMENT and "observation 123457 " in str(e.value)


def test_scene_generator_matches_oracle_copy(mba, O):
    for preset, window in [(1, 0), (4, 7)]:
        a = mba.generate_scene(preset, window)
        b = O.generate(O.preset(preset, window))
        for k in ["cam_true", "cam_init", "pt_true", "pt_init", "obs", "cam_idx", "pt_idx"]:
            np.testing.assert_array_equal(a[k], b[k])


def test_grouped_scene_generator_matches_oracle_copy(mba, O):
    """Same seeded scene from the library and the oracle's copy of the
    generator; projections may differ in the last bit between the two
    compilers' floating-point contraction, so observations are held to
    1e-12 px."""
    a = mba.generate_grouped_scene(6)
    b = O.generate_groups(O.preset(6), 2)
    for k in ["cam_init", "intr_init", "pt_init", "cam_idx", "intr_idx", "pt_idx"]:
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_allclose(a["obs"], b["obs"], rtol=0, atol=1e-12)


def test_scene_statistics_follow_the_configs(O):
    """Observation counts of the seeded configs (~8k, ~700k, ~4M)."""
    L = O.lib()
    for cfg, lo, hi in [(1, 7500, 8500), (2, 650_000, 750_000), (3, 3_800_000, 4_200_000)]:
        sp = O.preset(cfg)
        tl = np.zeros(sp.num_points, np.int32)
        n = L.orc_scene_count(ctypes.byref(sp), tl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        assert lo <= n <= hi, (cfg, n)


def test_errors_are_status_codes_not_aborts(mba, c1):
    p = ba_problem(mba, c1)
    p.camera_index = p.camera_index.copy()
    p.camera_index[7] = 99
    with pytest.raises(mba.MbaError) as e:
        mba.solve(p)
    assert e.value.status == mba.ERR_INVALID_ARGUMENT
    assert "out of range" in str(e.value)


def test_non_point_elimination_is_reported_unsupported(mba, tiny_scene):
    """The independent set puts a camera in E for this scene; the engine
    declines rather than solving a different system."""
    with pytest.raises(mba.MbaError) as e:
        mba.solve(ba_problem(mba, tiny_scene))
    assert e.value.status == mba.ERR_UNSUPPORTED


def test_batched_input_errors_name_the_window(mba, c1, tiny_scene):
    """Windows are validated on the host before any device work; the first
    bad window is named."""
    opts = mba.default_options(linear_solver=1)
    bad = ba_problem(mba, c1)
    bad.cameras[0, 0] = np.nan
    with pytest.raises(mba.MbaError) as e:
        mba.solve_batched([ba_problem(mba, c1), bad, ba_problem(mba, tiny_scene)], opts)
    assert e.value.status == mba.ERR_NOT_FINITE and "window 1" in str(e.value)
    with pytest.raises(mba.MbaError) as e:
        mba.solve_batched([ba_problem(mba, c1), ba_problem(mba, c1), ba_problem(mba, tiny_scene)], opts)
    assert e.value.status == mba.ERR_UNSUPPORTED and "window 2" in str(e.value)
    idx = ba_problem(mba, c1)
    idx.point_index = idx.point_index.copy()
    idx.point_index[3] = -1
    with pytest.raises(mba.MbaError) as e:
        mba.solve_batched([idx], opts)
    assert e.value.status == mba.ERR_INVALID_ARGUMENT
    assert mba.solve_batched([], opts) == []


def test_ceres_shaped_problem_builds_the_same_arrays(mba, c1):
    prob = mba.Problem()
    cams = [c1["cam_init"][i].copy() for i in range(c1["cam_init"].shape[0])]
    pts = [c1["pt_init"][i].copy() for i in range(c1["pt_init"].shape[0])]
    loss = mba.HuberLoss(4.0)
    for i in range(200):
        prob.AddResidualBlock(mba.ReprojectionError(c1["obs"][i]), loss, cams[c1["cam_idx"][i]],
                              pts[c1["pt_idx"][i]])
    prob.SetParameterBlockConstant(cams[c1["cam_idx"][0]])
    arr = prob._arrays()
    assert prob.NumResidualBlocks() == 200
    assert arr.camera_constant is not None and arr.camera_constant.sum() == 1
    np.testing.assert_array_equal(arr.observations, c1["obs"][:200])


@pytest.mark.parametrize("first_robust", [True, False])
def test_ceres_shaped_problem_refuses_mixed_loss(mba, c1, first_robust):
    """The engine applies one loss to all observations: a block without a loss
    next to one with HuberLoss (in either order) is refused, not silently
    robustified; two blocks sharing one loss, or both plain, are fine."""
    cams = [c1["cam_init"][i].copy() for i in range(2)]
    pts = [c1["pt_init"][i].copy() for i in range(2)]
    loss = mba.HuberLoss(4.0)
    prob = mba.Problem()
    prob.AddResidualBlock(mba.ReprojectionError(c1["obs"][0]), loss if first_robust else None, cams[0], pts[0])
    with pytest.raises(ValueError, match="with and without a loss"):
        prob.AddResidualBlock(mba.ReprojectionError(c1["obs"][1]), None if first_robust else loss, cams[1], pts[1])
    prob.AddResidualBlock(mba.ReprojectionError(c1["obs"][1]), loss if first_robust else None, cams[1], pts[1])
    assert prob.NumResidualBlocks() == 2


def test_ceres_shaped_grouped_problem_builds_the_same_arrays(mba, group_scene):
    """The incremental engine's calls (sfm_incremental_engine.cpp:1396-1446):
    AddResidualBlock(intrinsics, extrinsics, point), SetParameterization with
    a SubsetParameterization, SetParameterBlockConstant on a camera."""
    d, _ = group_scene
    intr = [d["intr_init"][g].copy() for g in range(3)]
    cams = [d["cam_init"][c].copy() for c in range(d["cam_init"].shape[0])]
    pts = [d["pt_init"][q].copy() for q in range(d["pt_init"].shape[0])]
    prob = mba.Problem()
    focal_only = mba.SubsetParameterization(6, [1, 2, 3, 4, 5])
    for i in range(300):
        g = d["intr_idx"][i]
        prob.AddResidualBlock(mba.ErrorFunc_Refine_Camera_3DPoints(d["obs"][i]), mba.HuberLoss(4.0), intr[g],
                              cams[d["cam_idx"][i]], pts[d["pt_idx"][i]])
        prob.SetParameterization(intr[g], focal_only)
    prob.SetParameterBlockConstant(cams[d["cam_idx"][0]])
    arr = prob._arrays()
    assert arr.layout == mba.LAYOUT_GROUPED and arr.cameras.shape[1] == 6
    assert set(arr.intrinsic_subset.tolist()) == {0x3E} and arr.intrinsic_constant.sum() == 0
    assert arr.camera_constant.sum() == 1
    np.testing.assert_array_equal(arr.intrinsics[arr.intrinsic_index], d["intr_init"][d["intr_idx"][:300]])
    with pytest.raises(ValueError):
        mba.SubsetParameterization(6, [0, 1, 2, 3, 4, 5])
    with pytest.raises(ValueError):
        prob.AddResidualBlock(mba.ReprojectionError(d["obs"][0]), None, np.zeros(9), pts[0])


def test_schur_jacobi_requirements(mba, c1, group_scene):
    """SCHUR_JACOBI is for packed cameras observing each point once; both
    are checked on the host before any device work."""
    d, _ = group_scene
    opts = mba.default_options(linear_solver=mba.IMPLICIT_SCHUR_PCG, preconditioner=mba.PRECOND_SCHUR_JACOBI)
    with pytest.raises(mba.MbaError) as e:
        mba.solve(grouped_problem(mba, d), opts)
    assert e.value.status == mba.ERR_UNSUPPORTED
    p = ba_problem(mba, c1)
    obs = np.vstack([p.observations, p.observations[:1]])
    dup = mba.BAProblem(p.cameras, p.points, obs, np.append(p.camera_index, p.camera_index[0]),
                        np.append(p.point_index, p.point_index[0]))
    with pytest.raises(mba.MbaError) as e:
        mba.solve(dup, opts)
    assert e.value.status == mba.ERR_UNSUPPORTED and "more than once" in str(e.value)
    with pytest.raises(mba.MbaError) as e:
        mba.solve(ba_problem(mba, c1), mba.default_options(preconditioner=7))
    assert e.value.status == mba.ERR_INVALID_ARGUMENT
