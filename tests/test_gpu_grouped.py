"""GPU parity of the grouped-intrinsics layout: the incremental engine's
residual blocks (intrinsics, camera, point) with intrinsic groups shared by
cameras (sfm_incremental_engine.cpp:1413-1446).  Checked against the
reference's Ceres fixtures and the FP64 oracle; every call goes through the
C ABI (libmba.so).

Tolerances (stated per test):
  FP64 evaluate -- residuals and Jacobians 1e-10 relative (absolute 1e-9 px,
                   1e-8 for extrinsic columns, whose entries reach ~1e4);
  FP64 LM       -- costs 1e-9 relative, CG iteration counts equal, as for the
                   packed layout; where the oracle itself meets Ceres only to
                   1e-6 (Huber + constant camera, long CG solves) the same
                   bars as the oracle's;
  mixed         -- per-entry |a - b| <= 1e-5 (|b| + column RMS) against FP64,
                   final cost within 1e-6 relative of the FP64 reference.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT, grouped_problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def group_scene(O):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import group_const, group_scene as gs
    return gs(), group_const()


def rel_err(a, b, axis=0):
    rms = np.sqrt(np.mean(b * b, axis=axis, keepdims=True)) + 1e-30
    return float(np.max(np.abs(a - b) / (np.abs(b) + rms)))


def test_grouped_evaluate_matches_reference_jets(mba, golden, group_scene):
    d, _ = group_scene
    r, ji, jc, jp, _ = mba.evaluate_blocks(grouped_problem(mba, d), mba.PRECISION_FP64)
    idx = golden["grp_jets_index"]
    np.testing.assert_allclose(r[idx], golden["grp_jets_residual"], rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(ji[idx], golden["grp_jets_jac_intrinsics"], rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(jc[idx], golden["grp_jets_jac_camera"], rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(jp[idx], golden["grp_jets_jac_point"], rtol=1e-10, atol=1e-9)


def test_grouped_evaluate_matches_oracle_and_closed_forms(mba, O, group_scene):
    d, _ = group_scene
    r, ji, jc, jp, cost = mba.evaluate_blocks(grouped_problem(mba, d), mba.PRECISION_FP64, huber_a=4.0)
    ro, jco, jpo, costo = O.evaluate(O.group_scene(d), huber_a=4.0)
    np.testing.assert_allclose(r, ro, rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(jc, jco, rtol=1e-10, atol=1e-8)
    np.testing.assert_allclose(jp, jpo, rtol=1e-10, atol=1e-9)
    np.testing.assert_allclose(cost, costo, rtol=1e-12)
    # Intrinsic columns [f, ppx, ppy, k1, k2, k3] from the model's closed
    # form: f dr/df = pred - pp; dr/dpp = I; dr/dk_{j+1} = r^2 dr/dk_j, so
    # (dr/dk2)^2 = dr/dk1 dr/dk3.
    intr = d["intr_init"][d["intr_idx"]]
    pred_minus_pp = r + d["obs"] - intr[:, 1:3]
    np.testing.assert_allclose(intr[:, :1] * ji[:, :, 0], pred_minus_pp, rtol=1e-12, atol=1e-9)
    np.testing.assert_array_equal(ji[:, :, 1], np.tile([1.0, 0.0], (len(r), 1)))
    np.testing.assert_array_equal(ji[:, :, 2], np.tile([0.0, 1.0], (len(r), 1)))
    np.testing.assert_allclose(ji[:, :, 4] ** 2, ji[:, :, 3] * ji[:, :, 5], rtol=1e-10, atol=1e-12)


def test_grouped_evaluate_mixed_within_fp32_rounding(mba, group_scene):
    d, _ = group_scene
    p = grouped_problem(mba, d)
    r, ji, jc, jp, _ = mba.evaluate_blocks(p, mba.PRECISION_FP64)
    r32, ji32, jc32, jp32, _ = mba.evaluate_blocks(p, mba.PRECISION_MIXED)
    # A residual is a difference of ~500 px values: its FP32 error is
    # relative to (|pred| + f), not to the residual.
    scale = np.abs(r + d["obs"]) + d["intr_init"][d["intr_idx"], :1]
    assert np.max(np.abs(r32 - r) / scale) <= 1e-5
    for a, b in [(ji32, ji), (jc32, jc), (jp32, jp)]:
        assert rel_err(a.reshape(len(r), -1), b.reshape(len(r), -1)) <= 1e-5


def test_grouped_fp64_pcg_matches_reference_ceres(mba, golden, group_scene):
    """ITERATIVE_SCHUR + JACOBI on the grouped scene, 20 iterations: the
    reference's trajectory, CG counts and result."""
    d, _ = group_scene
    p = grouped_problem(mba, d)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=1))
    ref = golden["grp_iterative_iterations"]
    assert s.IsSolutionUsable()
    assert (s.num_successful_steps, s.num_unsuccessful_steps, s.termination_type) == (
        int(golden["grp_iterative_successful"]), int(golden["grp_iterative_unsuccessful"]),
        int(golden["grp_iterative_termination"]))
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-9)
    assert [it.linear_solver_iterations for it in s.iterations] == [int(v) for v in ref[:, 6]]
    np.testing.assert_allclose(s.final_cost, float(golden["grp_iterative_final_cost"]), rtol=1e-10)
    # Weakly determined directions still moving at iteration 20: absolute
    # tolerances for poses and points (as the oracle's own check).
    np.testing.assert_allclose(p.cameras, golden["grp_iterative_cameras"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(p.intrinsics, golden["grp_iterative_intrinsics"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(p.points, golden["grp_iterative_points"], rtol=0, atol=1e-5)


def test_grouped_fp64_explicit_matches_reference_sparse_schur(mba, golden, group_scene):
    """The incremental engine's own call: SPARSE_SCHUR, Huber 4 px, a
    constant camera (explicit reduced system, dense Cholesky on the device)."""
    d, cc = group_scene
    p = grouped_problem(mba, d, cc)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=0, huber_a=4.0))
    ref = golden["grp_sparse_iterations"]
    assert s.IsSolutionUsable()
    assert (s.num_successful_steps, s.num_unsuccessful_steps, s.termination_type) == (
        int(golden["grp_sparse_successful"]), int(golden["grp_sparse_unsuccessful"]),
        int(golden["grp_sparse_termination"]))
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-9)
    np.testing.assert_allclose(s.final_cost, float(golden["grp_sparse_final_cost"]), rtol=1e-10)
    np.testing.assert_allclose(p.cameras, golden["grp_sparse_cameras"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(p.intrinsics, golden["grp_sparse_intrinsics"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(p.points, golden["grp_sparse_points"], rtol=0, atol=1e-5)
    # An FP32 S loses this problem's weakly determined directions (the mixed
    # run ended 2e-4 above FP64 after 20 iterations, 9e-6 after 50), so
    # grouped explicit solves run in FP64 whatever precision is asked for.
    sm = mba.solve(grouped_problem(mba, d, cc), mba.default_options(max_num_iterations=20, linear_solver=0,
                                                                    huber_a=4.0, precision=mba.PRECISION_MIXED))
    assert sm.final_cost == s.final_cost


def test_grouped_fp64_pcg_huber_constant_camera(mba, golden, group_scene):
    """Huber (4 px) with 5% outliers and a constant camera: long CG solves
    amplify summation-order rounding, so (as for the oracle against Ceres)
    the first four iterations are held to 1e-9 with equal CG counts for the
    first five, every cost to 1e-6 and the final cost to 1e-7."""
    d, cc = group_scene
    s = mba.solve(grouped_problem(mba, d, cc),
                  mba.default_options(max_num_iterations=20, linear_solver=1, huber_a=4.0))
    ref = golden["grp_iterative_const_iterations"]
    its = s.iterations
    assert len(its) == ref.shape[0]
    np.testing.assert_allclose([it.cost for it in its[:4]], ref[:4, 1], rtol=1e-9)
    assert [it.linear_solver_iterations for it in its[:5]] == [int(v) for v in ref[:5, 6]]
    np.testing.assert_allclose([it.cost for it in its], ref[:, 1], rtol=1e-6)
    np.testing.assert_allclose(s.final_cost, float(golden["grp_iterative_const_final_cost"]), rtol=1e-7)


@pytest.mark.parametrize("case,solver", [("subset_sparse", 0), ("subset_iterative", 1),
                                         ("const_group_sparse", 0)])
def test_grouped_subset_and_constant_group_match_reference(mba, golden, group_scene, case, solver):
    """The incremental engine's intrinsic options against the reference's
    Ceres: SubsetParameterization masks (0x3E focal-only, its default; 0x38
    k1..k3 held) and a constant group, Huber 4 px, a constant camera.  Held
    entries come back bit for bit.  SPARSE_SCHUR trajectories to 1e-9; the
    ITERATIVE_SCHUR one with the bars of the other long-CG case."""
    import make_golden
    d, cc = group_scene
    sub = ic = None
    if case == "subset_sparse":
        sub = make_golden.GROUP_SUBSET
    elif case == "subset_iterative":
        sub = np.full(3, 0x3E, np.uint8)
    else:
        ic = make_golden.GROUP_INTR_CONST
    p = grouped_problem(mba, d, cc, ic, sub)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=solver, huber_a=4.0))
    ref = golden[f"grp_{case}_iterations"]
    its = s.iterations
    assert s.IsSolutionUsable() and len(its) == ref.shape[0]
    if solver == 0:
        np.testing.assert_allclose([it.cost for it in its], ref[:, 1], rtol=1e-9)
        np.testing.assert_allclose(s.final_cost, float(golden[f"grp_{case}_final_cost"]), rtol=1e-10)
        np.testing.assert_allclose(p.intrinsics, golden[f"grp_{case}_intrinsics"], rtol=1e-8, atol=1e-8)
    else:
        np.testing.assert_allclose([it.cost for it in its[:4]], ref[:4, 1], rtol=1e-9)
        assert [it.linear_solver_iterations for it in its[:4]] == [int(v) for v in ref[:4, 6]]
        np.testing.assert_allclose([it.cost for it in its], ref[:, 1], rtol=1e-6)
        np.testing.assert_allclose(s.final_cost, float(golden[f"grp_{case}_final_cost"]), rtol=1e-7)
    for g in range(3):
        held = [k for k in range(6) if (sub is not None and (int(sub[g]) >> k) & 1) or (ic is not None and ic[g])]
        np.testing.assert_array_equal(p.intrinsics[g, held], d["intr_init"][g, held])
        np.testing.assert_array_equal(golden[f"grp_{case}_intrinsics"][g, held], d["intr_init"][g, held])


def test_incremental_engine_call_through_the_ceres_shaped_api(mba, golden, group_scene):
    """sfm_incremental_engine.cpp:1396-1472 written against this library:
    per-observation AddResidualBlock(intrinsics, extrinsics, point) with one
    shared HuberLoss(4), SetParameterization per group, the first camera held
    constant, SPARSE_SCHUR: the reference's result (subset_sparse fixture)."""
    import make_golden
    d, _ = group_scene
    intr = [d["intr_init"][g].copy() for g in range(3)]
    cams = [d["cam_init"][c].copy() for c in range(d["cam_init"].shape[0])]
    pts = [d["pt_init"][q].copy() for q in range(d["pt_init"].shape[0])]
    prob, loss = mba.Problem(), mba.HuberLoss(4.0)
    params = {0x3E: mba.SubsetParameterization(6, [1, 2, 3, 4, 5]), 0x38: mba.SubsetParameterization(6, [3, 4, 5])}
    for i in range(d["obs"].shape[0]):
        g = d["intr_idx"][i]
        prob.AddResidualBlock(mba.ErrorFunc_Refine_Camera_3DPoints(d["obs"][i]), loss, intr[g],
                              cams[d["cam_idx"][i]], pts[d["pt_idx"][i]])
        m = int(make_golden.GROUP_SUBSET[g])
        if m:
            prob.SetParameterization(intr[g], params[m])
    prob.SetParameterBlockConstant(cams[make_golden.GROUP_CONST])
    opts = mba.SolverOptions()
    opts.linear_solver_type = mba.SPARSE_SCHUR
    opts.max_num_iterations = 20
    s = mba.Solve(opts, prob)
    assert s.IsSolutionUsable()
    assert s.num_successful_steps == int(golden["grp_subset_sparse_successful"])
    np.testing.assert_allclose(s.final_cost, float(golden["grp_subset_sparse_final_cost"]), rtol=1e-10)
    np.testing.assert_allclose(np.array(intr), golden["grp_subset_sparse_intrinsics"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(np.array(cams), golden["grp_subset_sparse_cameras"], rtol=0, atol=1e-6)


def test_grouped_constant_blocks_stay_put(mba, group_scene):
    """SetParameterBlockConstant on a camera and on a group: their values are
    returned bit for bit."""
    d, cc = group_scene
    ic = np.zeros(3, np.uint8)
    ic[1] = 1
    p = grouped_problem(mba, d, cc, ic)
    s = mba.solve(p, mba.default_options(max_num_iterations=5, linear_solver=1, huber_a=4.0))
    assert s.IsSolutionUsable() and s.final_cost < s.initial_cost
    np.testing.assert_array_equal(p.cameras[cc == 1], d["cam_init"][cc == 1])
    np.testing.assert_array_equal(p.intrinsics[1], d["intr_init"][1])
    assert np.all(p.intrinsics[[0, 2]] != d["intr_init"][[0, 2]])


def test_grouped_single_shared_group_matches_oracle(mba, O):
    """One intrinsic group shared by all 40 cameras (a single-camera
    sequence): the shared block couples every camera and every point; the
    trajectory follows the oracle's."""
    d = O.generate_groups(O.make_spec(num_cameras=40, num_points=3000, track_min=2, track_max=12,
                                      outlier_frac=0.05, seed=77), 1)
    opts = dict(max_num_iterations=10, linear_solver=1, huber_a=4.0)
    so, its = O.solve(O.group_scene(d), O.default_options(**opts))
    s = mba.solve(mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"],
                                intrinsics=d["intr_init"].copy(), intrinsic_index=d["intr_idx"]),
                  mba.default_options(**opts))
    n = min(len(its), 4)
    np.testing.assert_allclose([it.cost for it in s.iterations[:n]], [it.cost for it in its[:n]], rtol=1e-9)
    np.testing.assert_allclose(s.final_cost, so.final_cost, rtol=1e-7)


def test_grouped_mixed_reaches_fp64_final_cost(mba, golden, group_scene):
    d, _ = group_scene
    s = mba.solve(grouped_problem(mba, d), mba.default_options(max_num_iterations=20, linear_solver=1,
                                                               precision=mba.PRECISION_MIXED))
    ref = float(golden["grp_iterative_final_cost"])
    assert s.IsSolutionUsable()
    assert abs(s.final_cost - ref) / ref < 1e-6


@pytest.mark.parametrize("prec", [0, 1])
def test_grouped_pcg_graph_replay_matches_direct_launches(mba, group_scene, prec):
    """Grouped layout (group-direction kernel and group-row sums inside the
    CG iteration): replaying the captured iteration equals launching its
    kernels one by one, bit for bit."""
    d, _ = group_scene
    runs = []
    for graph in (1, 0):
        p = grouped_problem(mba, d)
        s = mba.solve(p, mba.default_options(max_num_iterations=10, linear_solver=1, huber_a=4.0, precision=prec,
                                             pcg_graph_launch=graph))
        runs.append((s, p))
    (sa, pa), (sb, pb) = runs
    assert sa.final_cost == sb.final_cost
    assert sa.num_linear_solver_iterations == sb.num_linear_solver_iterations
    np.testing.assert_array_equal(pa.cameras, pb.cameras)
    np.testing.assert_array_equal(pa.intrinsics, pb.intrinsics)
