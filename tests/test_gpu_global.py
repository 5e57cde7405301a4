"""GPU parity of the global engine's BA calls (libs/sfm/src/sfm_global_engine.cpp),
expressed as packed cameras whose `camera_params` mask leaves the fixed
entries as data:
  PinholeReprojectionError_Rt <2,6,3>: [aa, t] with a fixed focal (final BA,
    :1758-1802; and the 2-view problems of the pair loop, :985-1174);
  PinholeReprojectionError_t <2,3,3>: translation only, rotation and focal
    fixed (:1548-1594).
Checked against the reference's own Ceres running the reference's functors
(tests/golden/ceres_ref.npz); every call goes through the C ABI.

Tolerances: LM costs 1e-9 relative with equal step counts and terminations;
parameters 1e-6 absolute; fixed entries bit for bit.  The batched 2-view
problems solved by PCG in shared launches (CG to eta 1e-10, so each step is
the exact dense one up to CG rounding) hold the same trajectories to 1e-8.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gscene(O):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_golden
    g = make_golden.global_scene()
    return g, make_golden.two_view_problems(g)


def packed(mba, g, cams6, params):
    cams = np.zeros((cams6.shape[0], 9))
    cams[:, :6] = cams6
    cams[:, 6] = g["focal"] if "focal" in g else 500.0
    return mba.BAProblem(cams, g["pt_init"].copy(), g["obs"], g["cam_idx"], g["pt_idx"], camera_params=params)


def check(s, golden, name):
    ref = golden[f"{name}_iterations"]
    assert s.IsSolutionUsable()
    assert (s.num_successful_steps, s.num_unsuccessful_steps, s.termination_type) == (
        int(golden[f"{name}_successful"]), int(golden[f"{name}_unsuccessful"]), int(golden[f"{name}_termination"]))
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-9)
    np.testing.assert_allclose(s.final_cost, float(golden[f"{name}_final_cost"]), rtol=1e-10)


@pytest.mark.parametrize("solver", [0, 1])
def test_rt_fixed_focal_matches_reference(mba, golden, gscene, solver):
    """Final BA of the global engine: HuberLoss(Square(2.0)), SPARSE_SCHUR
    (explicit here; the PCG run must reach the same final cost)."""
    g, _ = gscene
    p = packed(mba, g, g["cam_init"], 0x3F)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=solver, huber_a=4.0))
    if solver == 0:
        check(s, golden, "rt_sparse")
        np.testing.assert_allclose(p.cameras[:, :6], golden["rt_sparse_cameras"], rtol=0, atol=1e-6)
    else:
        np.testing.assert_allclose(s.final_cost, float(golden["rt_sparse_final_cost"]), rtol=1e-6)
    np.testing.assert_array_equal(p.cameras[:, 6:], np.tile([g["focal"], 0.0, 0.0], (p.cameras.shape[0], 1)))


def test_translation_only_matches_reference(mba, golden, gscene):
    """Translation-only BA: rotations are data (the true ones), HuberLoss(Square(4.0))."""
    g, _ = gscene
    cams6 = np.concatenate([g["cam_true"][:, :3], g["cam_init"][:, 3:]], axis=1)
    p = packed(mba, g, cams6, 0x38)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=0, huber_a=16.0))
    check(s, golden, "tonly_sparse")
    np.testing.assert_allclose(p.cameras[:, 3:6], golden["tonly_sparse_translations"], rtol=0, atol=1e-6)
    np.testing.assert_array_equal(p.cameras[:, :3], g["cam_true"][:, :3])


@pytest.mark.parametrize("solver", [0, 1])
def test_two_view_batch_matches_reference(mba, golden, gscene, solver):
    """The pair loop's 2-view DENSE_SCHUR problems as one batched call:
    explicit (windows in turn) and PCG (all windows in shared launches)."""
    g, tvs = gscene
    probs = [packed(mba, dict(tv, focal=g["focal"]), tv["cam_init"], 0x3F) for tv in tvs]
    extra = {} if solver == 0 else dict(eta=1e-10, max_linear_solver_iterations=100)
    sums = mba.solve_batched(probs, mba.default_options(linear_solver=solver, huber_a=4.0, **extra))
    tol = 1e-9 if solver == 0 else 1e-8
    for k, (s, p) in enumerate(zip(sums, probs)):
        assert s.IsSolutionUsable()
        assert (s.num_successful_steps, s.num_unsuccessful_steps) == tuple(int(v) for v in golden["twoview_steps"][k])
        assert s.termination_type == int(golden["twoview_termination"][k])
        np.testing.assert_allclose([s.initial_cost, s.final_cost], golden["twoview_costs"][k], rtol=tol)
        # Nothing is held in a 2-view problem, so poses are fixed only up to
        # a similarity (7 gauge directions); CG rounding moves along them
        # (~4e-6), so the PCG run is judged by its gauge-invariant costs.
        if solver == 0:
            np.testing.assert_allclose(p.cameras[:, :6], golden[f"twoview_{k}_cameras"], rtol=0, atol=1e-6)
