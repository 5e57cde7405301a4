"""SCHUR_JACOBI preconditioning of the implicit-Schur PCG (exact diagonal
blocks of S, schur_jacobi_preconditioner.cc:83-103) against the reference's
Ceres ITERATIVE_SCHUR + SCHUR_JACOBI runs (tests/golden/ceres_ref.npz), and in
batched windows against lone solves.  Through the C ABI.

Tolerances: costs 1e-9 relative with equal CG counts on the 16-camera scene;
with Huber and a constant camera (long CG solves, rounding-sensitive stops)
the first four iterations 1e-9 with equal CG counts, every cost 1e-6 and the
final cost 1e-7, as for the JACOBI runs of that kind; mixed precision within
1e-6 of the FP64 final cost; batched windows 1e-9 of lone solves.
"""
import numpy as np
import pytest

from conftest import ba_problem

pytestmark = pytest.mark.gpu


def sj_opts(mba, **kw):
    return mba.default_options(linear_solver=mba.IMPLICIT_SCHUR_PCG, preconditioner=mba.PRECOND_SCHUR_JACOBI, **kw)


def test_schur_jacobi_matches_reference(mba, golden, c1):
    s = mba.solve(ba_problem(mba, c1), sj_opts(mba, max_num_iterations=20))
    ref = golden["c1_schur_jacobi_iterations"]
    assert (s.num_successful_steps, s.termination_type) == (int(golden["c1_schur_jacobi_successful"]),
                                                           int(golden["c1_schur_jacobi_termination"]))
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-9)
    assert [it.linear_solver_iterations for it in s.iterations] == [int(v) for v in ref[:, 6]]
    np.testing.assert_allclose(s.final_cost, float(golden["c1_schur_jacobi_final_cost"]), rtol=1e-10)


def test_schur_jacobi_huber_constant_camera(mba, golden, huber_scene):
    const = np.zeros(30, np.uint8)
    const[3] = 1
    s = mba.solve(ba_problem(mba, huber_scene, const), sj_opts(mba, max_num_iterations=20, huber_a=4.0))
    ref = golden["huber_schur_jacobi_iterations"]
    its = s.iterations
    assert len(its) == ref.shape[0]
    np.testing.assert_allclose([it.cost for it in its[:4]], ref[:4, 1], rtol=1e-9)
    assert [it.linear_solver_iterations for it in its[:5]] == [int(v) for v in ref[:5, 6]]
    np.testing.assert_allclose([it.cost for it in its], ref[:, 1], rtol=1e-6)
    np.testing.assert_allclose(s.final_cost, float(golden["huber_schur_jacobi_final_cost"]), rtol=1e-7)


def test_schur_jacobi_mixed_and_batched(mba, golden, c1):
    sm = mba.solve(ba_problem(mba, c1), sj_opts(mba, max_num_iterations=20, precision=mba.PRECISION_MIXED))
    ref = float(golden["c1_schur_jacobi_final_cost"])
    assert abs(sm.final_cost - ref) / ref < 1e-6
    ws = [mba.generate_scene(4, w) for w in range(6)]
    opts = sj_opts(mba, max_num_iterations=10, huber_a=4.0)
    probs = [mba.BAProblem(w["cam_init"].copy(), w["pt_init"].copy(), w["obs"], w["cam_idx"], w["pt_idx"])
             for w in ws]
    sums = mba.solve_batched(probs, opts)
    for w, sb in zip(ws, sums):
        sl = mba.solve(mba.BAProblem(w["cam_init"].copy(), w["pt_init"].copy(), w["obs"], w["cam_idx"], w["pt_idx"]),
                       opts)
        assert sb.num_successful_steps == sl.num_successful_steps and sb.termination_type == sl.termination_type
        np.testing.assert_allclose(sb.final_cost, sl.final_cost, rtol=1e-9)
