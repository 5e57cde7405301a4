"""GPU parity of the HIP engine against the FP64 oracle and the reference's
Ceres fixtures.  Every call goes through the C ABI (libmba.so).

Tolerances (stated per test):
  FP64 path  -- Jets / blocks to 1e-10 relative, LM costs to 1e-9 relative;
  mixed path -- per-entry |a - b| <= 1e-5 (|b| + column RMS) against FP64
                (blocks U, V: ||dB||_F <= 1e-5 ||B||_F; gradient-type sums
                against the sum of their |terms|), final cost
                within 1e-6 relative of the FP64 reference.
"""
import numpy as np
import pytest

from conftest import axis_point_scene, ba_problem

pytestmark = pytest.mark.gpu


def rel_err(a, b, axis=0):
    """max |a - b| / (|b| + RMS of b along `axis`)."""
    rms = np.sqrt(np.mean(b * b, axis=axis, keepdims=True)) + 1e-30
    return float(np.max(np.abs(a - b) / (np.abs(b) + rms)))


def block_err(a, b, n):
    """Symmetric n x n blocks stored as upper triangles (row-major): max over
    blocks of ||A - B||_F / ||B||_F.  FP32 rounding of the weakly observed
    depth direction of a point is ~1e-7 of the block, but can be 1e-5 of a
    small diagonal entry of an ill-conditioned point block, so the blocks are
    judged as matrices; residuals and Jacobians keep the per-entry bar."""
    iu = np.triu_indices(n)
    w = np.where(iu[0] == iu[1], 1.0, 2.0)  # off-diagonals appear twice
    num = np.sqrt(np.sum(w * (a - b) ** 2, axis=1))
    den = np.sqrt(np.sum(w * b ** 2, axis=1)) + 1e-300
    return float(np.max(num / den))


def term_scales(O, d, huber, scale, ref):
    """Error scale of g_p = sum Jp'r, g_c = sum Jc'r, rhs = sum Jc'(r - Jp w):
    the per-entry residual bound (|pred| + f, see the adjoint check) carried
    linearly through each sum, plus the |terms| themselves."""
    sc = O.Scene(d)
    r, jc, jp, _ = O.evaluate(sc, huber_a=huber)
    s = np.sum(r * r, axis=1)
    w = np.where((huber > 0) & (s > huber * huber), np.sqrt(huber / np.sqrt(np.maximum(s, 1e-300))), 1.0)
    nc, npt = sc.num_cameras, sc.num_points
    cs = scale[:9 * nc].reshape(nc, 9)[sc.cam_idx]
    ps = scale[9 * nc:].reshape(npt, 3)[sc.pt_idx]
    Jc = np.abs(jc * w[:, None, None] * cs[:, None, :])
    Jp = np.abs(jp * w[:, None, None] * ps[:, None, :])
    f = sc.cameras[sc.cam_idx, 6][:, None]
    m = np.abs(r * w[:, None]) + w[:, None] * (np.abs(r + d["obs"]) + f)
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
    return {"gp": gp, "gc": gc, "rhs": rh}


def oracle_scale_D(O, d, scale, radius):
    sc = O.Scene(d)
    ref0 = O.jacobian_schur(sc, scale=scale)
    ut = [i * 9 - i * (i - 1) // 2 for i in range(9)]
    dc = np.clip(ref0["U"][:, ut], 1e-6, 1e32)
    dp = np.clip(ref0["V"][:, [0, 3, 5]], 1e-6, 1e32)
    return np.sqrt(np.concatenate([dc.ravel(), dp.ravel()]) / radius), ref0


# --------------------------------------------------------------- evaluate --
def test_evaluate_fp64_matches_oracle_jets(mba, O, c1):
    r, jc, jp, cost = mba.evaluate(ba_problem(mba, c1), mba.PRECISION_FP64)
    ro, jco, jpo, costo = O.evaluate(O.Scene(c1))
    assert np.max(np.abs(r - ro)) < 1e-9
    assert rel_err(jc, jco) < 1e-10 and rel_err(jp, jpo) < 1e-10
    np.testing.assert_allclose(cost, costo, rtol=1e-12)


def test_evaluate_fp64_matches_reference_autodiff(mba, O, golden):
    d2 = O.generate(O.preset(2))
    idx = golden["jets_c2_index"]
    sub = dict(cam_init=d2["cam_init"], pt_init=d2["pt_init"], obs=d2["obs"][idx],
               cam_idx=d2["cam_idx"][idx], pt_idx=d2["pt_idx"][idx])
    r, jc, jp, _ = mba.evaluate(ba_problem(mba, sub), mba.PRECISION_FP64)
    assert np.max(np.abs(r - golden["jets_c2_residual"])) < 1e-9
    assert rel_err(jc, golden["jets_c2_jac_camera"]) < 1e-10
    assert rel_err(jp, golden["jets_c2_jac_point"]) < 1e-10


def test_adjoint_check_700k_mixed_vs_fp64_jets(mba, O):
    """200 cameras, 100k points, ~700k observations: FP32 residual + analytic
    Jacobian against the FP64 Jets, in units of the FP32 unit roundoff
    u = 2^-24.  Residuals: an FP32 residual is the difference of two pixel
    coordinates of the image-plane scale |predicted| + f (the camera-frame
    point's cancellation carries f |X_c| / z), held to 16 u of that scale.
    Jacobian entries: per entry against |b| + column RMS, held to 128 u
    (7.6e-6; the measured worst entries are ~57 u)."""
    d2 = O.generate(O.preset(2))
    r, jc, jp, _ = mba.evaluate(ba_problem(mba, d2), mba.PRECISION_MIXED)
    ro, jco, jpo, _ = O.evaluate(O.Scene(d2))
    f = d2["cam_init"][d2["cam_idx"], 6][:, None]
    scale = np.abs(ro + d2["obs"]) + f
    u = 2.0 ** -24
    e_r = float(np.max(np.abs(r - ro) / scale)) / u
    e_c, e_p = rel_err(jc, jco) / u, rel_err(jp, jpo) / u
    print(f"700k adjoint check: residual {e_r:.2f} u of |pred| + f, camera J {e_c:.1f} u, point J {e_p:.1f} u")
    assert e_r < 16
    assert e_c < 128 and e_p < 128


def test_jacobian_matches_central_differences(mba, O):
    """Central differences, h = 1e-4 relative, on 300 observations (FP64 path)."""
    d2 = O.generate(O.preset(2))
    idx = np.random.default_rng(7).choice(d2["num_obs"], 300, replace=False)
    sub = dict(cam_init=d2["cam_init"], pt_init=d2["pt_init"], obs=d2["obs"][idx],
               cam_idx=d2["cam_idx"][idx], pt_idx=d2["pt_idx"][idx])
    _, jc, jp, _ = mba.evaluate(ba_problem(mba, sub), mba.PRECISION_FP64)
    for k in range(9):
        h = 1e-4 * np.maximum(np.abs(sub["cam_init"][:, k]), 1e-2)
        a, b = sub["cam_init"].copy(), sub["cam_init"].copy()
        a[:, k] += h
        b[:, k] -= h
        ra = mba.evaluate(mba.BAProblem(a, sub["pt_init"], sub["obs"], sub["cam_idx"], sub["pt_idx"]), jacobians=False)[0]
        rb = mba.evaluate(mba.BAProblem(b, sub["pt_init"], sub["obs"], sub["cam_idx"], sub["pt_idx"]), jacobians=False)[0]
        fd = (ra - rb) / (2 * h[sub["cam_idx"]])[:, None]
        assert rel_err(jc[:, :, k], fd) < 1e-5, k
    for k in range(3):
        h = 1e-4 * np.maximum(np.abs(sub["pt_init"][:, k]), 1e-2)
        a, b = sub["pt_init"].copy(), sub["pt_init"].copy()
        a[:, k] += h
        b[:, k] -= h
        ra = mba.evaluate(mba.BAProblem(sub["cam_init"], a, sub["obs"], sub["cam_idx"], sub["pt_idx"]), jacobians=False)[0]
        rb = mba.evaluate(mba.BAProblem(sub["cam_init"], b, sub["obs"], sub["cam_idx"], sub["pt_idx"]), jacobians=False)[0]
        fd = (ra - rb) / (2 * h[sub["pt_idx"]])[:, None]
        assert rel_err(jp[:, :, k], fd) < 1e-5, k


# -------------------------------------------------------- Jacobian + Schur --
@pytest.mark.parametrize("prec,huber,tol", [(0, 4.0, 1e-10), (1, 0.0, 1e-5), (1, 4.0, 1e-4)])
@pytest.mark.parametrize("cfg", [1, 2])
def test_jacobian_schur_blocks_match_oracle(mba, O, prec, huber, tol, cfg):
    """Blocks of one plain Jacobian + Schur pass (residuals formed as FP32
    pred - obs inside the kernels) against the FP64 oracle.  Without a loss
    function the mixed blocks meet 1e-5.  With Huber, V is held to 1e-4 by
    a cancellation argument: the FP32 residual is the difference of two
    pixel coordinates of size ~f, so its absolute error is up to ~1e-4 px,
    and an outlier row's weight sqrt(a/|r|) carries it as a relative error
    |dr| / (2 |r|) ~ 1e-5 at |r| = a = 4 px; a point seen by few rows, most
    of them outliers, passes that straight into V (measured worst 5e-5).
    The LM loop does not run this variant: its mixed passes take the FP64
    cost's residuals (Context.lm_pass, held to 1e-5 with Huber in
    test_gpu_scale.py)."""
    d = O.generate(O.preset(cfg))
    with mba.Context(ba_problem(mba, d), mba.default_options(precision=prec, huber_a=huber)) as ctx:
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
    errs = {}
    for k, b, n in [("U", ref0["U"], 9), ("V", ref0["V"], 3)]:
        errs[k] = block_err(out[k], b, n)
    ts = term_scales(O, d, huber, scale, ref)
    for k in ("gc", "gp", "rhs"):
        errs[k] = float(np.max(np.abs(out[k] - ref[k]) / (ts[k] + 1e-300)))
    print(f"blocks cfg {cfg} precision {prec} huber {huber}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    for k, e in errs.items():
        assert e < tol, (k, e)


def test_schur_matvec_matches_dense_schur_complement(mba, O):
    d = O.generate(O.make_spec(8, 150, 2, 5, outlier_frac=0.1, seed=21))
    with mba.Context(ba_problem(mba, d), mba.default_options(huber_a=4.0)) as ctx:
        ctx.jacobian_schur(10.0)
        cs, ps = ctx.read_scale()
        x = np.random.default_rng(5).standard_normal((8, 9))
        y = ctx.schur_matvec(x)
    sc = O.Scene(d)
    nc, npt, n = sc.num_cameras, sc.num_points, sc.num_obs
    r, jc, jp, _ = O.evaluate(sc, huber_a=4.0)
    scale = np.concatenate([cs.ravel(), ps.ravel()])
    J = np.zeros((2 * n, 9 * nc + 3 * npt))
    for i in range(n):
        s = r[i] @ r[i]
        w = np.sqrt(4.0 / np.sqrt(s)) if s > 16.0 else 1.0
        J[2 * i:2 * i + 2, 9 * sc.cam_idx[i]:9 * sc.cam_idx[i] + 9] = w * jc[i]
        J[2 * i:2 * i + 2, 9 * nc + 3 * sc.pt_idx[i]:9 * nc + 3 * sc.pt_idx[i] + 3] = w * jp[i]
    J *= scale[None, :]
    H = J.T @ J
    diag = np.clip(np.diag(H), 1e-6, 1e32)
    H = H + np.diag(diag / 10.0)
    f, e = slice(0, 9 * nc), slice(9 * nc, None)
    S = H[f, f] - H[f, e] @ np.linalg.solve(H[e, e], H[e, f])
    np.testing.assert_allclose(y.ravel(), S @ x.ravel(), rtol=1e-9, atol=1e-9 * np.abs(S @ x.ravel()).max())


@pytest.mark.parametrize("ncams", [200, 1500])
def test_mixed_pass_and_matvec_with_staged_camera_records(mba, O, ncams):
    """The FP32 point-major kernels read camera records staged in LDS when the
    table fits (pass up to ~2,000 cameras, matvec up to ~1,100): 200 cameras
    run both staged, 1,500 the staged pass beside the gathering matvec.  The
    mixed pass's V and g_p and the mixed implicit S x against the FP64 path's
    (which gathers): within FP32 rounding, 1e-5 of each block's scale."""
    d = O.generate(O.make_spec(ncams, 12 * ncams, 2, 8, seed=41))
    out = {}
    for prec in (0, 1):
        with mba.Context(ba_problem(mba, d), mba.default_options(precision=prec)) as ctx:
            ctx.jacobian_schur(1e4)
            blocks = ctx.read_schur()
            x = np.random.default_rng(7).standard_normal((ncams, 9))
            out[prec] = (blocks, ctx.schur_matvec(x))
    (b64, y64), (b32, y32) = out[0], out[1]
    assert block_err(b32["V"], b64["V"], 3) < 1e-5
    assert float(np.max(np.abs(b32["gp"] - b64["gp"]))) < 1e-5 * float(np.max(np.abs(b64["gp"])))
    assert float(np.max(np.abs(y32 - y64))) < 1e-5 * float(np.max(np.abs(y64)))


# -------------------------------------------------------------- LM solves --
@pytest.mark.parametrize("solver,name", [(0, "sparse"), (1, "iterative")])
def test_fp64_solve_matches_reference_ceres(mba, golden, c1, solver, name):
    p = ba_problem(mba, c1)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=solver))
    ref = golden[f"c1_{name}_iterations"]
    assert s.IsSolutionUsable()
    assert s.num_successful_steps == int(golden[f"c1_{name}_successful"])
    assert s.num_unsuccessful_steps == int(golden[f"c1_{name}_unsuccessful"])
    assert s.termination_type == int(golden[f"c1_{name}_termination"])
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-9)
    np.testing.assert_allclose(s.final_cost, float(golden[f"c1_{name}_final_cost"]), rtol=1e-10)
    assert [it.linear_solver_iterations for it in s.iterations] == [int(v) for v in ref[:, 6]]
    np.testing.assert_allclose(p.cameras, golden[f"c1_{name}_cameras"], rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(p.points, golden[f"c1_{name}_points"], rtol=1e-6, atol=1e-8)


@pytest.mark.parametrize("case", ["huber_sparse", "huber_iterative", "const"])
def test_fp64_solve_huber_and_constant_camera(mba, golden, huber_scene, case):
    const = None
    if case == "const":
        const = np.zeros(30, np.uint8)
        const[3] = 1
    solver = 1 if case == "huber_iterative" else 0
    s = mba.solve(ba_problem(mba, huber_scene, const),
                  mba.default_options(max_num_iterations=20, linear_solver=solver, huber_a=4.0))
    assert s.num_successful_steps == int(golden[f"{case}_successful"])
    np.testing.assert_allclose([it.cost for it in s.iterations], golden[f"{case}_iterations"][:, 1], rtol=1e-8)
    np.testing.assert_allclose(s.final_cost, float(golden[f"{case}_final_cost"]), rtol=1e-9)


def test_explicit_solve_with_tracks_longer_than_128(mba, O):
    """Points seen by up to 140 cameras (21 of them by more than 128): the
    explicit FP64 trajectory matches the oracle's SPARSE_SCHUR run, the mixed
    one reaches its final cost within 1e-6."""
    d = O.generate(O.make_spec(num_cameras=150, num_points=800, track_min=2, track_max=140, power_alpha=1.0,
                               seed=7))
    assert np.bincount(d["pt_idx"]).max() > 128
    so, its = O.solve(O.Scene(d), O.default_options(linear_solver=0, max_num_iterations=5, huber_a=4.0))
    s = mba.solve(ba_problem(mba, d), mba.default_options(linear_solver=0, max_num_iterations=5, huber_a=4.0))
    np.testing.assert_allclose([it.cost for it in s.iterations], [it.cost for it in its], rtol=1e-9)
    assert s.num_successful_steps == so.num_successful_steps and s.termination_type == so.termination_type
    sm = mba.solve(ba_problem(mba, d), mba.default_options(linear_solver=0, max_num_iterations=5, huber_a=4.0,
                                                           precision=mba.PRECISION_MIXED))
    assert abs(sm.final_cost - so.final_cost) <= 1e-6 * so.final_cost


def test_mixed_lm_with_fp64_residuals_on_split_layout(mba, O):
    """C2 (697k observations, Huber 4 px): the camera-major kernels run on XCD
    point ranges, and inside the LM loop the mixed Jacobian pass takes its
    residuals from the FP64 cost of the same state.  FP32 residual noise in
    the gradient used to give runs of invalid steps (negative FP64 model cost
    change) that stopped mixed solves early, well above the minimum.  The
    mixed solve must take no invalid step and end at or below the FP64
    solve's final cost (within 1e-6)."""
    d = O.generate(O.preset(2))
    res = {}
    for prec in (0, 1):
        o = mba.default_options(precision=prec, huber_a=4.0, linear_solver=1, max_num_iterations=60)
        res[prec] = mba.solve(ba_problem(mba, d), o)
    assert res[1].is_solution_usable
    assert not any(not it.step_is_valid for it in res[1].iterations[1:])
    assert res[1].final_cost <= res[0].final_cost * (1 + 1e-6), (res[1].final_cost, res[0].final_cost)


@pytest.mark.parametrize("solver", [0, 1])
def test_mixed_precision_reaches_fp64_final_cost(mba, golden, c1, solver):
    s = mba.solve(ba_problem(mba, c1), mba.default_options(max_num_iterations=20, linear_solver=solver,
                                                           precision=mba.PRECISION_MIXED))
    ref = float(golden["c1_sparse_final_cost"])
    assert s.IsSolutionUsable()
    assert abs(s.final_cost - ref) / ref < 1e-6


def test_solve_is_deterministic(mba, c1):
    opts = mba.default_options(max_num_iterations=10, linear_solver=1, precision=mba.PRECISION_MIXED)
    a, b = ba_problem(mba, c1), ba_problem(mba, c1)
    sa, sb = mba.solve(a, opts), mba.solve(b, opts)
    assert sa.final_cost == sb.final_cost
    np.testing.assert_array_equal(a.cameras, b.cameras)
    np.testing.assert_array_equal(a.points, b.points)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("precond", ["jacobi", "schur_jacobi"])
def test_pcg_graph_replay_matches_direct_launches(mba, c1, huber_scene, prec, precond):
    """A CG iteration replayed from its captured HIP graph (pcg_graph_launch) runs
    the kernels that launching them one by one runs, with the same
    arguments: bit-identical parameters, costs and iteration counts over LM
    runs whose radius changes every iteration (kernels read it from device
    memory) and, in mixed precision, alternate the row-weight slot."""
    const = np.zeros(30, np.uint8)
    const[3] = 1
    for d, k, huber in ((c1, None, 0.0), (huber_scene, const, 4.0)):
        runs = []
        for graph in (1, 0):
            p = ba_problem(mba, d, k)
            o = mba.default_options(max_num_iterations=15, linear_solver=mba.IMPLICIT_SCHUR_PCG, huber_a=huber,
                                    precision=mba.PRECISION_MIXED if prec == "mixed" else mba.PRECISION_FP64,
                                    preconditioner=mba.PRECOND_SCHUR_JACOBI if precond == "schur_jacobi"
                                    else mba.PRECOND_JACOBI, pcg_graph_launch=graph)
            runs.append((mba.solve(p, o), p))
        (sa, pa), (sb, pb) = runs
        assert sa.IsSolutionUsable()
        assert sa.final_cost == sb.final_cost
        assert (sa.num_successful_steps, sa.num_unsuccessful_steps, sa.num_linear_solver_iterations) == \
               (sb.num_successful_steps, sb.num_unsuccessful_steps, sb.num_linear_solver_iterations)
        np.testing.assert_array_equal(pa.cameras, pb.cameras)
        np.testing.assert_array_equal(pa.points, pb.points)


def test_nonfinite_input_is_an_error(mba, c1):
    p = ba_problem(mba, c1)
    p.points[3, 1] = np.inf
    with pytest.raises(mba.MbaError) as e:
        mba.solve(p)
    assert e.value.status == mba.ERR_NOT_FINITE


@pytest.mark.parametrize("solver", ["explicit", "pcg"])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_singular_point_block_makes_the_step_invalid(mba, O, solver, prec):
    """A point seen by one camera on that camera's optical axis, with the
    camera's rotation zero, has V = Jp'Jp with a zero row and column along
    the world z axis; with min_lm_diagonal = 0 its damped block V + D^2 is
    singular.  The pass flags the failed 3x3 inverse and every step built on
    it is invalid, as Ceres treats a failed linear solve
    (trust_region_minimizer.cc:323-366): five in a row end the solve with
    FAILURE (max_num_consecutive_invalid_steps) and the parameters are not
    moved; the oracle's LM (oracle/mba_oracle.c, a non-finite or failed
    solve is an invalid step) gives the same records.  With the default
    diagonal floor (1e-6) the same problem solves."""
    ls = mba.EXPLICIT_SCHUR if solver == "explicit" else mba.IMPLICIT_SCHUR_PCG
    pr = mba.PRECISION_FP64 if prec == "fp64" else mba.PRECISION_MIXED
    cams, pts, obs, ci, pi = axis_point_scene()
    p = mba.BAProblem(cams.copy(), pts.copy(), obs, ci, pi)
    s = mba.solve(p, mba.default_options(linear_solver=ls, precision=pr, min_lm_diagonal=0.0))
    print(f"{solver} {prec}: termination {s.termination_type}, records "
          f"{[(it.iteration, it.step_is_valid) for it in s.iterations]}", flush=True)
    d = dict(cam_init=cams, pt_init=pts, obs=obs, cam_idx=ci, pt_idx=pi)
    os_, oits = O.solve(O.Scene(d), O.default_options(linear_solver=0 if solver == "explicit" else 1,
                                                      min_lm_diagonal=0.0))
    assert (s.termination_type, len(s.iterations)) == (os_.termination_type, len(oits))
    assert [it.step_is_valid for it in s.iterations] == [it.step_is_valid for it in oits]
    assert s.termination_type == mba.FAILURE and not s.is_solution_usable
    assert [it.step_is_valid for it in s.iterations[1:]] == [0] * 4  # the fifth ends the solve
    assert all(it.cost == s.iterations[0].cost for it in s.iterations)
    np.testing.assert_array_equal(p.cameras, cams)
    np.testing.assert_array_equal(p.points, pts)
    q = mba.BAProblem(cams.copy(), pts.copy(), obs, ci, pi)
    s = mba.solve(q, mba.default_options(linear_solver=ls, precision=pr))
    assert s.is_solution_usable and s.iterations[1].step_is_valid
    assert s.final_cost < s.initial_cost


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_failed_triangular_solve_is_an_invalid_step(mba, c1, prec, monkeypatch):
    """The explicit solver's triangular solves hand 64-row blocks between
    workgroups through bounded waits.  With the wait limit set to zero from
    the host (MBA_FLOW_SPIN_LIMIT=0 when the context is created) every wait
    gives up at once and the solve comes out NaN (16 cameras: blocks 1 and 2
    of 3 wait); the LM treats that as a failed linear solve: invalid steps,
    FAILURE after five, parameters untouched.  A context created with the
    default limit then solves the problem exactly as before."""
    pr = mba.PRECISION_FP64 if prec == "fp64" else mba.PRECISION_MIXED
    opts = mba.default_options(linear_solver=mba.EXPLICIT_SCHUR, precision=pr)
    ref = mba.solve(ba_problem(mba, c1), opts)
    monkeypatch.setenv("MBA_FLOW_SPIN_LIMIT", "0")
    p = ba_problem(mba, c1)
    s = mba.solve(p, opts)
    print(f"{prec}: termination {s.termination_type}, records {[(it.iteration, it.step_is_valid) for it in s.iterations]}",
          flush=True)
    assert s.termination_type == mba.FAILURE and not s.is_solution_usable
    assert [it.step_is_valid for it in s.iterations[1:]] == [0] * 4  # the fifth ends the solve
    np.testing.assert_array_equal(p.cameras, c1["cam_init"])
    np.testing.assert_array_equal(p.points, c1["pt_init"])
    monkeypatch.delenv("MBA_FLOW_SPIN_LIMIT")
    again = mba.solve(ba_problem(mba, c1), opts)
    assert again.final_cost == ref.final_cost and len(again.iterations) == len(ref.iterations)
    assert ref.termination_type == mba.CONVERGENCE


def test_ceres_shaped_api_like_the_reference_unit_test(mba, O):
    """bundle_adjustment_unittest.cpp:19-122 shape: per-observation residual
    blocks, SPARSE_SCHUR, the solution is usable and the RMSE decreases; plus
    the final cost agrees with the oracle."""
    d = O.generate(O.make_spec(3, 6, 3, 3, seed=9))
    cams = [d["cam_init"][i].copy() for i in range(3)]
    pts = [d["pt_init"][i].copy() for i in range(6)]
    prob = mba.Problem()
    for i in range(d["num_obs"]):
        prob.AddResidualBlock(mba.ReprojectionError(d["obs"][i]), None, cams[d["cam_idx"][i]], pts[d["pt_idx"][i]])
    opts = mba.SolverOptions()
    opts.linear_solver_type = mba.SPARSE_SCHUR
    s = mba.Solve(opts, prob)
    n = prob.NumResidualBlocks()
    assert s.IsSolutionUsable()
    assert np.sqrt(s.initial_cost / (2 * n)) > np.sqrt(s.final_cost / (2 * n))
    so, _ = O.solve(O.Scene(d), O.default_options())
    np.testing.assert_allclose(s.final_cost, so.final_cost, rtol=1e-8, atol=1e-12)
    assert not np.allclose(cams[0], d["cam_init"][0])


# ------------------------------------------- full-size (4M) properties --
def test_full_size_jacobian_schur_properties(mba):
    """1,000 cameras, 500k points, ~4M observations (Huber 4 px): the pass is
    deterministic run to run, S is symmetric (x'Sy = y'Sx) and positive
    definite along random directions."""
    d = mba.generate_scene(3)
    p = mba.BAProblem(d["cam_init"], d["pt_init"], d["obs"], d["cam_idx"], d["pt_idx"])
    with mba.Context(p, mba.default_options(precision=mba.PRECISION_MIXED, huber_a=4.0)) as ctx:
        ctx.jacobian_schur(1e4)
        a = ctx.read_schur()
        ctx.jacobian_schur(1e4)
        b = ctx.read_schur()
        for k in a:
            np.testing.assert_array_equal(a[k], b[k])
        assert all(np.all(np.isfinite(v)) for v in a.values())
        rng = np.random.default_rng(0)
        x, y = rng.standard_normal((1000, 9)), rng.standard_normal((1000, 9))
        Sx, Sy = ctx.schur_matvec(x), ctx.schur_matvec(y)
        assert abs(np.sum(y * Sx) - np.sum(x * Sy)) <= 1e-4 * np.sqrt(np.sum(x * Sx) * np.sum(y * Sy))
        assert np.sum(x * Sx) > 0 and np.sum(y * Sy) > 0
