"""Parity at the benchmark scales, through the C ABI.

* Full-size trajectories of the reference's own Ceres (tests/golden/
  ceres_ref_large.npz, written by make_golden.py --large on one thread):
  the 200-camera / 697k-observation scene with SPARSE_SCHUR, and the
  1,000-camera / 4.0M-observation scene (Huber 4 px, 5% outliers) with
  ITERATIVE_SCHUR + JACOBI and + SCHUR_JACOBI, against the FP64 GPU path:
  every iteration record's cost to 1e-9 relative, trust radius to 1e-9,
  step validity, success and CG iteration counts exactly, relative decrease
  to 1e-5 (a ratio of cost differences), the final cameras and a seeded
  sample of final points to 1e-6 (TrustRegionMinimizer::Minimize,
  trust_region_minimizer.cc:262-711; ConjugateGradientsSolver,
  conjugate_gradients_solver.cc:66-245).
* Whole LM solves to convergence on the 697k-observation scene at the
  benchmark's LM settings (PCG forcing eta 1e-2, function tolerance 1e-12)
  with ITERATIVE_SCHUR + JACOBI and + SCHUR_JACOBI (tests/golden/
  ceres_ref_lm.npz, make_golden.py --lm): the FP64 GPU path ends with the
  reference's termination, step counts and final cost, its costs follow
  the reference's iteration by iteration, and the mixed path ends within
  1e-6 of the reference's JACOBI final cost.
* The mixed-precision criterion at 4.0M observations: the mixed LM reaches
  the FP64 LM's final cost within 1e-6 relative.
* The 10,000-camera / 4M-point / 25M-observation scene: the pass is
  deterministic, S is symmetric and positive definite, the LM with PCG to
  relative residual 1e-6 (conjugate_gradients_solver.cc:227-237) converges,
  and mixed ends within 1e-6 of FP64; the blocks of a scene drawn from the
  same track-length law (alpha 2.0482 on [2, 100]) at oracle-friendly size
  match the FP64 oracle.
* The mixed pass the LM actually runs (FP64 cost first, its residuals
  rounded once into the FP32 kernels) against the FP64 oracle at 1e-5.
"""
import os

import numpy as np
import pytest

from conftest import ROOT, ba_problem

pytestmark = pytest.mark.gpu

LARGE = os.path.join(ROOT, "tests", "golden", "ceres_ref_large.npz")
LM = os.path.join(ROOT, "tests", "golden", "ceres_ref_lm.npz")
U24 = 2.0 ** -24  # FP32 unit roundoff


@pytest.fixture(scope="module")
def large():
    return np.load(LARGE, allow_pickle=False)


def _check_trajectory(s, large, name):
    ref = large[f"{name}_iterations"]
    its = s.iterations
    print(f"{name}: {len(its)} records, final cost gpu {s.final_cost:.12e} ceres "
          f"{float(large[name + '_final_cost']):.12e}", flush=True)
    assert len(its) == ref.shape[0]
    assert s.termination_type == int(large[f"{name}_termination"])
    assert s.num_successful_steps == int(large[f"{name}_successful"])
    assert s.num_unsuccessful_steps == int(large[f"{name}_unsuccessful"])
    np.testing.assert_allclose([it.cost for it in its], ref[:, 1], rtol=1e-9)
    np.testing.assert_allclose([it.trust_region_radius for it in its], ref[:, 3], rtol=1e-9)
    assert [it.step_is_successful for it in its] == [int(v) for v in ref[:, 4]]
    assert [it.step_is_valid for it in its] == [int(v) for v in ref[:, 5]]
    assert [it.linear_solver_iterations for it in its] == [int(v) for v in ref[:, 6]]
    np.testing.assert_allclose([it.relative_decrease for it in its], ref[:, 7], rtol=1e-5, atol=1e-12)
    np.testing.assert_allclose(s.final_cost, float(large[f"{name}_final_cost"]), rtol=1e-10)


# ------------------------------------------- reference Ceres, full size --
def test_c2_sparse_schur_trajectory_matches_reference_ceres(mba, O, large):
    """697k observations, no loss, SPARSE_SCHUR + CX_SPARSE (the explicit
    solver, FP64), 20 iterations at most."""
    d = mba.generate_scene(2)
    p = ba_problem(mba, d)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=mba.EXPLICIT_SCHUR))
    _check_trajectory(s, large, "c2_sparse")
    np.testing.assert_allclose(p.cameras, large["c2_sparse_cameras"], rtol=1e-6, atol=1e-8)
    idx = large["c2_sparse_point_index"]
    np.testing.assert_allclose(p.points[idx], large["c2_sparse_points"], rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("name,precond", [("c3_iterative", 0), ("c3_schur_jacobi", 1)])
def test_c3_iterative_schur_trajectory_matches_reference_ceres(mba, large, name, precond):
    """4.0M observations, 5% 20-px outliers, Huber 4 px, ITERATIVE_SCHUR with
    JACOBI or SCHUR_JACOBI (FP64 PCG), 5 iterations at most."""
    d = mba.generate_scene(3)
    p = ba_problem(mba, d)
    s = mba.solve(p, mba.default_options(max_num_iterations=5, linear_solver=mba.IMPLICIT_SCHUR_PCG, huber_a=4.0,
                                         preconditioner=precond))
    _check_trajectory(s, large, name)
    np.testing.assert_allclose(p.cameras, large[f"{name}_cameras"], rtol=1e-6, atol=1e-8)
    idx = large[f"{name}_point_index"]
    np.testing.assert_allclose(p.points[idx], large[f"{name}_points"], rtol=1e-6, atol=1e-8)


# ------------------------- whole LM solves at 697k observations, to convergence --
@pytest.fixture(scope="module")
def lm_ref():
    return np.load(LM, allow_pickle=False)


@pytest.fixture(scope="module")
def c2(mba):
    return mba.generate_scene(2)


@pytest.mark.parametrize("name,precond", [("c2_lm_jacobi", 0), ("c2_lm_schur_jacobi", 1)])
def test_c2_lm_to_convergence_matches_reference_ceres(mba, lm_ref, c2, name, precond):
    """ITERATIVE_SCHUR with JACOBI or SCHUR_JACOBI, eta 1e-2, function
    tolerance 1e-12, at most 300 iterations (the reference converges in 55).
    FP64: the reference's termination, successful and unsuccessful step
    counts and final cost (1e-10), every iteration's cost within 1e-8 of the
    reference's (the CG iteration counts are not compared: over 55 solves the
    rounding of the Nash-Sofer test moves a few by one), total CG iterations
    within 5%.  Mixed: converges, and ends within 1e-6 of the reference's
    JACOBI final cost."""
    ref = lm_ref[f"{name}_iterations"]
    target = float(lm_ref["c2_lm_jacobi_final_cost"])
    res = {}
    for prec, cap in ((mba.PRECISION_FP64, 300), (mba.PRECISION_MIXED, 450)):
        o = mba.default_options(precision=prec, linear_solver=mba.IMPLICIT_SCHUR_PCG, preconditioner=precond,
                                max_num_iterations=cap, eta=1e-2, function_tolerance=1e-12)
        s = mba.solve(ba_problem(mba, c2), o)
        res[prec] = s
        inv = sum(1 for it in s.iterations[1:] if not it.step_is_valid)
        print(f"{name} precision {prec}: final cost {s.final_cost:.12e} (reference {float(lm_ref[name + '_final_cost']):.12e}), "
              f"{s.num_successful_steps}+{s.num_unsuccessful_steps} steps, {inv} invalid, CG "
              f"{s.num_linear_solver_iterations} (reference {int(ref[:, 6].sum())}), termination {s.termination_type}",
              flush=True)
    f64, mix = res[mba.PRECISION_FP64], res[mba.PRECISION_MIXED]
    assert f64.termination_type == int(lm_ref[f"{name}_termination"]) == mba.CONVERGENCE
    assert f64.num_successful_steps == int(lm_ref[f"{name}_successful"])
    assert f64.num_unsuccessful_steps == int(lm_ref[f"{name}_unsuccessful"])
    assert len(f64.iterations) == ref.shape[0]
    np.testing.assert_allclose([it.cost for it in f64.iterations], ref[:, 1], rtol=1e-8)
    np.testing.assert_allclose(f64.final_cost, float(lm_ref[f"{name}_final_cost"]), rtol=1e-10)
    assert abs(f64.num_linear_solver_iterations - ref[:, 6].sum()) <= 0.05 * ref[:, 6].sum()
    assert mix.termination_type == mba.CONVERGENCE and mix.is_solution_usable
    assert abs(mix.final_cost - target) / target <= 1e-6
    assert abs(f64.final_cost - target) / target <= 1e-6


# ---------------------------------------- mixed vs FP64 at 4.0M observations --
def test_c3_mixed_reaches_fp64_final_cost(mba):
    """The north-star criterion on the 4.0M-observation scene with the LM
    settings of the benchmark (PCG forcing eta 1e-2, function tolerance
    1e-12; FP64 at most 300 iterations, mixed 450): the mixed solve's final
    cost is within 1e-6 relative of the FP64 solve's.  Invalid steps (a
    negative FP64 model cost change of an FP32 PCG step near the minimum)
    are counted and reported; Ceres treats them as rejected steps too."""
    d = mba.generate_scene(3)
    res = {}
    for prec, cap in ((mba.PRECISION_FP64, 300), (mba.PRECISION_MIXED, 450)):
        o = mba.default_options(precision=prec, huber_a=4.0, linear_solver=mba.IMPLICIT_SCHUR_PCG,
                                max_num_iterations=cap, eta=1e-2, function_tolerance=1e-12)
        res[prec] = mba.solve(ba_problem(mba, d), o)
        s = res[prec]
        inv = [it.iteration for it in s.iterations[1:] if not it.step_is_valid]
        print(f"C3 LM precision {prec}: final cost {s.final_cost:.10e}, {s.num_successful_steps}+"
              f"{s.num_unsuccessful_steps} steps, termination {s.termination_type}, {len(inv)} invalid "
              f"steps {inv[:20]}", flush=True)
    f64, mix = res[mba.PRECISION_FP64], res[mba.PRECISION_MIXED]
    assert f64.is_solution_usable and mix.is_solution_usable
    rel = abs(mix.final_cost - f64.final_cost) / f64.final_cost
    print(f"C3 mixed vs FP64 final cost: {rel:.2e}", flush=True)
    assert rel <= 1e-6


# ------------------------------------------------- 25M-observation scene --
@pytest.fixture(scope="module")
def c5(mba):
    return mba.generate_scene(5)


def test_c5_pass_deterministic_symmetric_positive(mba, c5):
    """10,000 cameras, 4M points, ~25M observations (mixed): two passes give
    bit-identical blocks, all finite; S is symmetric (x'Sy = y'Sx) and
    positive along random directions."""
    d = c5
    p = mba.BAProblem(d["cam_init"], d["pt_init"], d["obs"], d["cam_idx"], d["pt_idx"])
    with mba.Context(p, mba.default_options(precision=mba.PRECISION_MIXED,
                                            linear_solver=mba.IMPLICIT_SCHUR_PCG)) as ctx:
        ctx.jacobian_schur(1e4)
        a = ctx.read_schur()
        ctx.jacobian_schur(1e4)
        b = ctx.read_schur()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
        assert all(np.all(np.isfinite(v)) for v in a.values())
        rng = np.random.default_rng(5)
        nc = d["cam_init"].shape[0]
        x, y = rng.standard_normal((nc, 9)), rng.standard_normal((nc, 9))
        Sx, Sy = ctx.schur_matvec(x), ctx.schur_matvec(y)
        xSx, ySy = float(np.sum(x * Sx)), float(np.sum(y * Sy))
        asym = abs(np.sum(y * Sx) - np.sum(x * Sy)) / np.sqrt(xSx * ySy)
        print(f"C5 S: x'Sx {xSx:.6e}, y'Sy {ySy:.6e}, asymmetry {asym:.2e}", flush=True)
        assert asym <= 1e-4
        assert xSx > 0 and ySy > 0


def test_c5_lm_converges_and_mixed_matches_fp64(mba, c5):
    """The 25M-observation LM: block-Jacobi PCG to relative residual 1e-6
    (Nash-Sofer test off), at most 30 iterations, in both precisions: both end
    CONVERGENCE with the RMS reprojection error below 0.62 px (from 3.5 px),
    and the mixed final cost is within 1e-6 of the FP64 one."""
    d = c5
    nobs = d["obs"].shape[0]
    res = {}
    for prec in (mba.PRECISION_FP64, mba.PRECISION_MIXED):
        o = mba.default_options(precision=prec, linear_solver=mba.IMPLICIT_SCHUR_PCG, max_num_iterations=30,
                                pcg_r_tolerance=1e-6, eta=1e-10)
        s = mba.solve(ba_problem(mba, d), o)
        res[prec] = s
        print(f"C5 LM precision {prec}: {s.num_iterations - 1} iterations, {s.num_linear_solver_iterations} CG, "
              f"RMS {np.sqrt(2 * s.initial_cost / nobs):.4f} -> {np.sqrt(2 * s.final_cost / nobs):.6f} px, "
              f"termination {s.termination_type}, {s.total_time_s:.2f} s", flush=True)
        assert s.termination_type == mba.CONVERGENCE
        assert np.sqrt(2 * s.final_cost / nobs) < 0.62
    rel = abs(res[1].final_cost - res[0].final_cost) / res[0].final_cost
    print(f"C5 mixed vs FP64 final cost: {rel:.2e}", flush=True)
    assert rel <= 1e-6


@pytest.mark.parametrize("prec", [0, 1])
def test_c5_law_blocks_match_oracle(mba, O, prec):
    """A scene with the 25M scene's track-length law (power law alpha 2.0482
    on [2, 100]) at 1,000 cameras / 40,000 points (~250k observations): the
    blocks of one pass against the FP64 oracle, FP64 to 1e-10, mixed to
    1e-5 (blocks in Frobenius norm; gradient-type sums against the sum of
    their |terms|)."""
    d = O.generate(O.make_spec(num_cameras=1000, num_points=40000, track_min=2, track_max=100,
                               power_alpha=2.0482, seed=5))
    assert np.bincount(d["pt_idx"]).max() > 50
    e = _blocks_vs_oracle(mba, O, d, prec, 0.0, lm=True)
    tol = 1e-10 if prec == 0 else 1e-5
    assert max(e.values()) < tol, e


# ----------------------------------- the pass the mixed LM actually runs --
def _blocks_vs_oracle(mba, O, d, prec, huber, lm):
    """Max errors of one pass (lm: Context.lm_pass, else jacobian_schur) at
    radius 1e4 against the oracle's blocks: U, V as ||dB||_F / ||B||_F per
    block; g_c, g_p, rhs per entry against the sum of |terms| of the entry
    (|J| |r w| for the gradients, |Jc| (|r w| + |Jp| |V^-1 g_p|) for rhs)."""
    from test_gpu_parity import block_err
    with mba.Context(ba_problem(mba, d), mba.default_options(precision=prec, huber_a=huber)) as ctx:
        if lm:
            ctx.lm_pass(1e4)
        else:
            ctx.jacobian_schur(1e4)
        out = ctx.read_schur()
        cs, ps = ctx.read_scale()
    scale = np.concatenate([cs.ravel(), ps.ravel()])
    sc = O.Scene(d)
    ref0 = O.jacobian_schur(sc, huber_a=huber, scale=scale)
    ut = [i * 9 - i * (i - 1) // 2 for i in range(9)]
    D = np.sqrt(np.concatenate([np.clip(ref0["U"][:, ut], 1e-6, 1e32).ravel(),
                                np.clip(ref0["V"][:, [0, 3, 5]], 1e-6, 1e32).ravel()]) / 1e4)
    ref = O.jacobian_schur(sc, huber_a=huber, scale=scale, D=D)
    err = {"U": block_err(out["U"], ref0["U"], 9), "V": block_err(out["V"], ref0["V"], 3)}
    r, jc, jp, _ = O.evaluate(sc, huber_a=huber)
    s = np.sum(r * r, axis=1)
    w = np.where((huber > 0) & (s > huber * huber), np.sqrt(huber / np.sqrt(np.maximum(s, 1e-300))), 1.0)
    nc, npt = sc.num_cameras, sc.num_points
    Jc = np.abs(jc * w[:, None, None] * scale[:9 * nc].reshape(nc, 9)[sc.cam_idx][:, None, :])
    Jp = np.abs(jp * w[:, None, None] * scale[9 * nc:].reshape(npt, 3)[sc.pt_idx][:, None, :])
    m = np.abs(r * w[:, None])
    gp = np.zeros((npt, 3))
    np.add.at(gp, sc.pt_idx, np.einsum("nri,nr->ni", Jp, m))
    gc = np.zeros((nc, 9))
    np.add.at(gc, sc.cam_idx, np.einsum("nri,nr->ni", Jc, m))
    iu = np.triu_indices(3)
    Vd = np.zeros((npt, 3, 3))
    Vd[:, iu[0], iu[1]] = ref["V"]
    Vd[:, iu[1], iu[0]] = ref["V"]
    wv = np.linalg.solve(Vd, ref["gp"][:, :, None])[:, :, 0]
    t = m + np.einsum("nrj,nj->nr", Jp, np.abs(wv[sc.pt_idx]))
    rh = np.zeros((nc, 9))
    np.add.at(rh, sc.cam_idx, np.einsum("nri,nr->ni", Jc, t))
    for k, sc_k in (("gc", gc), ("gp", gp), ("rhs", rh)):
        err[k] = float(np.max(np.abs(out[k] - ref[k]) / (sc_k + 1e-300)))
    print(f"blocks vs oracle (precision {prec}, huber {huber}, lm {lm}): "
          + ", ".join(f"{k} {v:.2e}" for k, v in err.items()), flush=True)
    return err


@pytest.mark.parametrize("huber", [0.0, 4.0])
@pytest.mark.parametrize("cfg", [1, 2])
def test_mixed_lm_pass_blocks_match_oracle(mba, O, cfg, huber):
    """The mixed pass as the LM runs it: the FP64 cost kernel's residuals,
    rounded once to FP32, feed the FP32 Jacobian + Schur kernels (gradient,
    rhs and the Huber weights sqrt(a/|r|), loss_function.cc:47-61 with the
    Corrector of corrector.cc:41-85).  U, V, g_c, g_p and rhs are held to
    the north-star 1e-5 with and without the loss."""
    d = O.generate(O.preset(cfg))
    e = _blocks_vs_oracle(mba, O, d, 1, huber, lm=True)
    assert max(e.values()) < 1e-5, e


def test_fp64_lm_pass_equals_plain_pass(mba, O, c1):
    """An FP64 context has no FP32 residual copies: lm_pass computes the same
    blocks as jacobian_schur, bit for bit, and returns the FP64 cost."""
    opts = mba.default_options(huber_a=4.0)
    with mba.Context(ba_problem(mba, c1), opts) as ctx:
        cost = ctx.lm_pass(1e4)
        a = ctx.read_schur()
        ctx.jacobian_schur(1e4)
        b = ctx.read_schur()
    for k in a:
        np.testing.assert_array_equal(a[k], b[k])
    _, _, _, costo = O.evaluate(O.Scene(c1), huber_a=4.0)
    np.testing.assert_allclose(cost, costo, rtol=1e-12)


# ------------------------------------------------------ launch shapes --
@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 2e-5)])
def test_launch_shapes_give_the_same_blocks(mba, O, prec, tol):
    """The autotuner's knobs change only how work is laid out: rows per lane
    of the point-major kernels (1, 8) and rows per camera-major chunk (256,
    1536) give the blocks of the default shape on the 697k-observation scene
    with Huber 4 px, to summation-order rounding (FP64: 1e-12 relative per
    block in Frobenius norm; mixed: 2e-5), and the same matvec."""
    from test_gpu_parity import block_err
    d = O.generate(O.preset(2))
    out = []
    for rpl, cr in ((0, 0), (1, 256), (8, 1536)):
        o = mba.default_options(precision=prec, huber_a=4.0, linear_solver=mba.IMPLICIT_SCHUR_PCG,
                                point_rows_per_lane=rpl, camera_chunk_rows=cr)
        with mba.Context(ba_problem(mba, d), o) as ctx:
            ctx.jacobian_schur(1e4)
            blk = ctx.read_schur()
            x = np.random.default_rng(1).standard_normal((d["cam_init"].shape[0], 9))
            blk["Sx"] = ctx.schur_matvec(x)
            out.append(blk)
    for other in out[1:]:
        errs = {"U": block_err(other["U"], out[0]["U"], 9), "V": block_err(other["V"], out[0]["V"], 3)}
        for k in ("gc", "gp", "rhs", "Sx"):
            errs[k] = float(np.max(np.abs(other[k] - out[0][k])) / np.max(np.abs(out[0][k])))
        print(f"launch shapes, precision {prec}: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
        assert max(errs.values()) < tol, errs


def test_launch_shape_options_are_checked(mba, c1):
    for kw in ({"point_rows_per_lane": 65}, {"point_rows_per_lane": -1}, {"camera_chunk_rows": 32},
               {"camera_chunk_rows": 9000}):
        with pytest.raises(mba.MbaError) as e:
            mba.Context(ba_problem(mba, c1), mba.default_options(**kw))
        assert e.value.status == mba.ERR_INVALID_ARGUMENT
