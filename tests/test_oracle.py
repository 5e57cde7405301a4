"""Pins the FP64 CPU oracle before it is trusted as the GPU checker.

Each part of the path is pinned by something other than a restatement of the
oracle's own formula:
  - golden fixtures written by the reference's own Ceres (tests/golden);
  - textbook routines (scipy's rotation vector -> matrix), closed forms and
    special cases of the camera model;
  - central finite differences for every derivative;
  - dense normal equations in numpy for the Schur complement;
  - a hand-worked ordering example.
"""
import os
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from conftest import ROOT, axis_point_scene

C1_ITERS = 20


def project_np(cam, X):
    """libs/sfm model with a textbook rotation (scipy)."""
    R = Rotation.from_rotvec(cam[:3]).as_matrix()
    x = R @ X + cam[3:6]
    xe, ye = x[0] / x[2], x[1] / x[2]
    r2 = xe * xe + ye * ye
    d = 1 + cam[7] * r2 + cam[8] * r2 * r2
    return np.array([cam[6] * d * xe, cam[6] * d * ye])


# ----------------------------------------------------------- A2 residual --
def test_residual_matches_textbook_rotation(O, c1):
    rng = np.random.default_rng(0)
    for i in rng.choice(c1["num_obs"], 50, replace=False):
        cam = c1["cam_init"][c1["cam_idx"][i]]
        X = c1["pt_init"][c1["pt_idx"][i]]
        r, _, _ = O.evaluate_obs(0, cam, X, c1["obs"][i])
        np.testing.assert_allclose(r, project_np(cam, X) - c1["obs"][i], rtol=0, atol=1e-9)


def test_residual_special_cases(O):
    X = np.array([0.3, -0.2, 5.0])
    obs = np.array([12.0, -7.0])
    # aa = 0: first-order branch, R = I.
    cam = np.array([0, 0, 0, 0.1, 0.2, 0.3, 500, 0.01, 0.001])
    r, _, _ = O.evaluate_obs(0, cam, X, obs)
    xe, ye = 0.4 / 5.3, 0.0 / 5.3
    r2 = xe * xe + ye * ye
    d = 1 + 0.01 * r2 + 0.001 * r2 * r2
    np.testing.assert_allclose(r, [500 * d * xe - 12.0, 500 * d * ye + 7.0], atol=1e-12)
    # tiny rotation (theta^2 <= DBL_EPSILON): R = I + [w]x exactly.
    w = np.array([1e-9, -2e-9, 3e-9])
    cam2 = np.concatenate([w, [0, 0, 0, 400, 0, 0]])
    r, _, _ = O.evaluate_obs(0, cam2, X, obs)
    Xc = X + np.cross(w, X)
    np.testing.assert_allclose(r, 400 * Xc[:2] / Xc[2] - obs, atol=1e-12)
    # k1 = k2 = 0: pure pinhole.
    cam3 = np.array([0.05, -0.02, 0.1, 0.2, -0.1, 0.5, 480, 0, 0])
    R = Rotation.from_rotvec(cam3[:3]).as_matrix()
    Xc = R @ X + cam3[3:6]
    r, _, _ = O.evaluate_obs(0, cam3, X, obs)
    np.testing.assert_allclose(r, 480 * Xc[:2] / Xc[2] - obs, atol=1e-10)
    # A point on the optical axis projects to the principal point (0, 0).
    cam4 = np.array([0, 0, 0, 0, 0, 1.0, 500, 0.3, 0.2])
    r, _, _ = O.evaluate_obs(0, cam4, np.array([0, 0, 4.0]), obs)
    np.testing.assert_allclose(r, -obs, atol=1e-12)


# ----------------------------------------------------------- A3 Jacobian --
def test_jets_match_reference_ceres_autodiff(O, golden):
    d2 = O.generate(O.preset(2))
    idx = golden["jets_c2_index"]
    for k, i in enumerate(idx):
        r, jc, jp = O.evaluate_obs(0, d2["cam_init"][d2["cam_idx"][i]], d2["pt_init"][d2["pt_idx"][i]], d2["obs"][i])
        np.testing.assert_allclose(r, golden["jets_c2_residual"][k], rtol=1e-13, atol=1e-12)
        np.testing.assert_allclose(jc, golden["jets_c2_jac_camera"][k], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(jp, golden["jets_c2_jac_point"][k], rtol=1e-12, atol=1e-12)


def test_jets_match_central_differences(O, c1):
    """autodiff_test.cc:159-233 style: FD vs AD at 1e-6 relative, h = 1e-4 relative."""
    rng = np.random.default_rng(1)
    for i in rng.choice(c1["num_obs"], 40, replace=False):
        cam = c1["cam_init"][c1["cam_idx"][i]].copy()
        X = c1["pt_init"][c1["pt_idx"][i]].copy()
        o = c1["obs"][i]
        _, jc, jp = O.evaluate_obs(0, cam, X, o)
        for blk, J in ((cam, jc), (X, jp)):
            for k in range(blk.size):
                h = 1e-4 * max(abs(blk[k]), 1e-2)
                a, b = blk.copy(), blk.copy()
                a[k] += h
                b[k] -= h
                if blk is cam:
                    fd = (O.evaluate_obs(0, a, X, o)[0] - O.evaluate_obs(0, b, X, o)[0]) / (2 * h)
                else:
                    fd = (O.evaluate_obs(0, cam, a, o)[0] - O.evaluate_obs(0, cam, b, o)[0]) / (2 * h)
                scale = np.max(np.abs(J)) + 1e-12
                np.testing.assert_allclose(J[:, k], fd, rtol=1e-6, atol=1e-6 * scale)


# ----------------------------------------------------- A4 Huber/Corrector --
def test_huber_cost_closed_form(O, huber_scene):
    sc = O.Scene(huber_scene)
    r, _, _, cost = O.evaluate(sc, huber_a=4.0, jacobians=False)
    s = np.sum(r * r, axis=1)
    rho = np.where(s > 16.0, 2 * 4.0 * np.sqrt(s) - 16.0, s)
    np.testing.assert_allclose(cost, 0.5 * rho.sum(), rtol=1e-13)
    assert np.mean(s > 16.0) > 0.03  # the outliers are in the outlier region


def test_corrected_gradient_is_gradient_of_robust_cost(O, huber_scene):
    """Corrector (corrector.cc:41-85): sqrt(rho') scaling makes J'r the exact
    gradient of 1/2 sum rho -- checked against central differences of the cost."""
    sc = O.Scene(huber_scene)
    out = O.jacobian_schur(sc, huber_a=4.0)
    rng = np.random.default_rng(3)
    for c in rng.choice(sc.num_cameras, 3, replace=False):
        for k in range(9):
            h = 1e-6 * max(abs(sc.cameras[c, k]), 1e-2)
            a = O.Scene(huber_scene, cameras=sc.cameras.copy())
            b = O.Scene(huber_scene, cameras=sc.cameras.copy())
            a.cameras[c, k] += h
            b.cameras[c, k] -= h
            fd = (O.evaluate(a, 4.0, False)[3] - O.evaluate(b, 4.0, False)[3]) / (2 * h)
            np.testing.assert_allclose(out["gc"][c, k], fd, rtol=2e-5, atol=1e-4)


# ------------------------------------------------------------ A6 ordering --
def test_ordering_hand_worked_example(O):
    """2 cameras, 3 points: degrees p1 = p2 = 1 < c0 = c1 = p0 = 2, so the
    stable degree sort gives [p1, p2, c0, p0, c1]; the greedy independent set
    is {p1, p2, p0}; residual buckets are filled back to front."""
    d = dict(cam_init=np.tile([0.0, 0, 0, 0, 0, 0, 500, 0, 0], (2, 1)),
             pt_init=np.array([[0, 0, 5.0], [1, 0, 5.0], [0, 1, 5.0]]),
             obs=np.zeros((4, 2)), cam_idx=np.array([0, 1, 0, 1], np.int32),
             pt_idx=np.array([0, 0, 1, 2], np.int32))
    ne, order, res = O.schur_ordering(O.Scene(d))
    assert ne == 3
    assert order.tolist() == [3, 4, 2, 0, 1]
    assert res.tolist() == [2, 3, 1, 0]


@pytest.mark.parametrize("name", ["c1", "huber", "tiny"])
def test_ordering_and_pairs_match_reference(O, golden, c1, huber_scene, tiny_scene, name):
    d = {"c1": c1, "huber": huber_scene, "tiny": tiny_scene}[name]
    const = None
    if name == "huber":
        const = np.zeros(30, np.uint8)
        const[3] = 1
    sc = O.Scene(d, cam_constant=const)
    ne, order, res = O.schur_ordering(sc)
    assert ne == int(golden[f"struct_{name}_num_e"])
    np.testing.assert_array_equal(order, golden[f"struct_{name}_order"])
    np.testing.assert_array_equal(res, golden[f"struct_{name}_residual_order"])
    np.testing.assert_array_equal(O.schur_block_pairs(sc), golden[f"struct_{name}_pairs"])


# ----------------------------------------------------------- A8/A9 Schur --
def test_schur_blocks_match_dense_normal_equations(O, huber_scene):
    """schur_eliminator_test.cc:79-183 style: the eliminated system against
    dense J'J + D'D built in numpy from the oracle's per-observation Jets."""
    d = O.generate(O.make_spec(8, 150, 2, 5, outlier_frac=0.1, seed=21))
    sc = O.Scene(d)
    nc, npt, n = sc.num_cameras, sc.num_points, sc.num_obs
    r, jc, jp, _ = O.evaluate(sc, huber_a=4.0)
    rng = np.random.default_rng(4)
    scale = rng.uniform(0.5, 1.5, 9 * nc + 3 * npt)
    D = rng.uniform(0.1, 1.0, 9 * nc + 3 * npt)
    # Dense scaled, corrected Jacobian; columns [cameras | points].
    J = np.zeros((2 * n, 9 * nc + 3 * npt))
    b = np.zeros(2 * n)
    for i in range(n):
        s = r[i] @ r[i]
        w = np.sqrt(4.0 / np.sqrt(s)) if s > 16.0 else 1.0
        c, p = sc.cam_idx[i], sc.pt_idx[i]
        J[2 * i:2 * i + 2, 9 * c:9 * c + 9] = w * jc[i]
        J[2 * i:2 * i + 2, 9 * nc + 3 * p:9 * nc + 3 * p + 3] = w * jp[i]
        b[2 * i:2 * i + 2] = w * r[i]
    J = J * scale[None, :]
    H = J.T @ J + np.diag(D * D)
    g = J.T @ b
    f = slice(0, 9 * nc)
    e = slice(9 * nc, None)
    Hee_inv = np.linalg.inv(H[e, e])
    rhs = g[f] - H[f, e] @ Hee_inv @ g[e]
    out = O.jacobian_schur(sc, huber_a=4.0, scale=scale, D=D)
    ut = np.triu_indices(9)
    for c in range(nc):
        np.testing.assert_allclose(out["U"][c], H[9 * c:9 * c + 9, 9 * c:9 * c + 9][ut], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(out["gc"].ravel(), g[f], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(out["rhs"].ravel(), rhs, rtol=1e-9, atol=1e-8)
    Vd = np.array([H[9 * nc + 3 * p:9 * nc + 3 * p + 3, 9 * nc + 3 * p:9 * nc + 3 * p + 3][np.triu_indices(3)]
                   for p in range(npt)])
    np.testing.assert_allclose(out["V"], Vd, rtol=1e-12, atol=1e-10)


# -------------------------------------------------------- A7/A10/A11 LM --
@pytest.mark.parametrize("solver,name", [(0, "sparse"), (1, "iterative")])
def test_lm_trajectory_matches_reference(O, golden, c1, solver, name):
    s, its = O.solve(O.Scene(c1), O.default_options(max_num_iterations=C1_ITERS, linear_solver=solver))
    ref = golden[f"c1_{name}_iterations"]
    assert len(its) == ref.shape[0]
    assert s.num_successful_steps == int(golden[f"c1_{name}_successful"])
    assert s.num_unsuccessful_steps == int(golden[f"c1_{name}_unsuccessful"])
    assert s.termination_type == int(golden[f"c1_{name}_termination"])
    for it, row in zip(its, ref):
        assert it.iteration == int(row[0])
        np.testing.assert_allclose(it.cost, row[1], rtol=1e-10)
        np.testing.assert_allclose(it.trust_region_radius, row[3], rtol=1e-12)
        assert it.step_is_successful == int(row[4])
        assert it.linear_solver_iterations == int(row[6])
    np.testing.assert_allclose(s.final_cost, float(golden[f"c1_{name}_final_cost"]), rtol=1e-10)


@pytest.mark.parametrize("case", ["huber_sparse", "huber_iterative", "const"])
def test_lm_huber_and_constant_camera_match_reference(O, golden, huber_scene, case):
    const = None
    solver = 1 if case == "huber_iterative" else 0
    if case == "const":
        const = np.zeros(30, np.uint8)
        const[3] = 1
    s, its = O.solve(O.Scene(huber_scene, cam_constant=const),
                     O.default_options(max_num_iterations=20, linear_solver=solver, huber_a=4.0))
    assert s.num_successful_steps == int(golden[f"{case}_successful"])
    assert s.num_unsuccessful_steps == int(golden[f"{case}_unsuccessful"])
    np.testing.assert_allclose(s.final_cost, float(golden[f"{case}_final_cost"]), rtol=1e-9)
    np.testing.assert_allclose([it.cost for it in its], golden[f"{case}_iterations"][:, 1], rtol=1e-9)


def test_lm_nonfinite_input_fails(O, c1):
    cams = c1["cam_init"].copy()
    pts = c1["pt_init"].copy()
    pts[5, 2] = np.nan
    s, _ = O.solve(O.Scene(c1, cameras=cams, points=pts), O.default_options(max_num_iterations=5))
    assert s.termination_type == O.FAILURE


# ------------------------------------- grouped intrinsics (layout 1) -----
@pytest.fixture(scope="module")
def group_scene():
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_golden
    return make_golden.group_scene(), make_golden.group_const()


def test_grouped_jets_match_reference_ceres_autodiff(O, golden, group_scene):
    """ErrorFunc_Refine_Camera_3DPoints (intrinsics, extrinsics, point) against
    the reference Ceres autodiff on 200 observations."""
    d, _ = group_scene
    for k, i in enumerate(golden["grp_jets_index"]):
        r, ji, jc, jp = O.evaluate_obs(1, d["cam_init"][d["cam_idx"][i]], d["pt_init"][d["pt_idx"][i]], d["obs"][i],
                                       d["intr_init"][d["intr_idx"][i]])
        np.testing.assert_allclose(r, golden["grp_jets_residual"][k], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(ji, golden["grp_jets_jac_intrinsics"][k], rtol=1e-12, atol=1e-10)
        np.testing.assert_allclose(jc, golden["grp_jets_jac_camera"][k], rtol=1e-12, atol=1e-9)
        np.testing.assert_allclose(jp, golden["grp_jets_jac_point"][k], rtol=1e-12, atol=1e-10)


def test_grouped_ordering_and_pairs_match_reference(O, golden, group_scene):
    d, const = group_scene
    sc = O.group_scene(d, cam_constant=const)
    ne, order, res = O.schur_ordering(sc)
    assert ne == int(golden["grp_struct_num_e"])
    np.testing.assert_array_equal(order, golden["grp_struct_order"])
    np.testing.assert_array_equal(res, golden["grp_struct_residual_order"])
    np.testing.assert_array_equal(O.schur_block_pairs(sc), golden["grp_struct_pairs"])


@pytest.mark.parametrize("case,huber,solver,const", [("sparse", 4.0, 0, True), ("iterative", 0.0, 1, False)])
def test_grouped_trajectory_matches_reference(O, golden, group_scene, case, huber, solver, const):
    d, cc = group_scene
    sc = O.group_scene(d, cam_constant=cc if const else None)
    s, its = O.solve(sc, O.default_options(max_num_iterations=20, linear_solver=solver, huber_a=huber))
    ref = golden[f"grp_{case}_iterations"]
    assert len(its) == ref.shape[0]
    assert (s.num_successful_steps, s.num_unsuccessful_steps, s.termination_type) == (
        int(golden[f"grp_{case}_successful"]), int(golden[f"grp_{case}_unsuccessful"]),
        int(golden[f"grp_{case}_termination"]))
    for it, row in zip(its, ref):
        np.testing.assert_allclose(it.cost, row[1], rtol=1e-10)
        assert it.linear_solver_iterations == int(row[6])
    # The run stops at 20 iterations while still stepping (|step| ~ 0.02) along
    # weakly determined directions, where rounding moves parameters by ~1e-6
    # at equal cost: absolute tolerances for the poses and points.
    np.testing.assert_allclose(sc.cameras, golden[f"grp_{case}_cameras"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(sc.intrinsics, golden[f"grp_{case}_intrinsics"], rtol=1e-8, atol=1e-8)
    np.testing.assert_allclose(sc.points, golden[f"grp_{case}_points"], rtol=0, atol=1e-5)


def test_grouped_iterative_with_constant_camera_matches_reference(O, golden, group_scene):
    """With a constant camera and Huber outliers the CG solves reach 45
    iterations and amplify summation-order rounding (the oracle and Ceres add
    the Jacobian cells of a row in different orders): the first four
    iterations agree to 1e-10 with equal CG counts; afterwards the Nash-Sofer
    stop can land one iteration apart, so every cost is held to 1e-6 and the
    final cost to 1e-7."""
    d, cc = group_scene
    s, its = O.solve(O.group_scene(d, cam_constant=cc), O.default_options(max_num_iterations=20, linear_solver=1,
                                                                           huber_a=4.0))
    ref = golden["grp_iterative_const_iterations"]
    assert len(its) == ref.shape[0]
    np.testing.assert_allclose([it.cost for it in its[:4]], ref[:4, 1], rtol=1e-10)
    assert [it.linear_solver_iterations for it in its[:5]] == [int(v) for v in ref[:5, 6]]
    np.testing.assert_allclose([it.cost for it in its], ref[:, 1], rtol=1e-6)
    np.testing.assert_allclose(s.final_cost, float(golden["grp_iterative_const_final_cost"]), rtol=1e-7)


@pytest.mark.parametrize("linear_solver", [0, 1])
def test_singular_point_block_is_an_invalid_step(O, linear_solver):
    """The axis point's Jacobian has a zero third column (x = y = 0 in the
    camera frame, R = I), so its V has a zero row and column and, with
    min_lm_diagonal = 0, V + D^2 is singular: the Schur elimination cannot
    invert it, the step is invalid, and five invalid steps in a row end the
    LM with FAILURE (trust_region_minimizer.cc:323-366).  With the default
    floor (1e-6) the LM runs normally (its steps are valid and the cost
    falls; the axis point's depth is barely damped, so it does not converge
    within the 50 iterations)."""
    cams, pts, obs, ci, pi = axis_point_scene()
    d = dict(cam_init=cams, pt_init=pts, obs=obs, cam_idx=ci, pt_idx=pi)
    _, _, jp, _ = O.evaluate(O.Scene(d))
    assert np.all(jp[-1][:, 2] == 0.0) and jp[-1][0, 0] != 0.0 and jp[-1][1, 1] != 0.0
    s, its = O.solve(O.Scene(d), O.default_options(linear_solver=linear_solver, min_lm_diagonal=0.0))
    # The fifth invalid step ends the solve before its record is kept.
    assert s.termination_type == 2
    assert [it.step_is_valid for it in its[1:]] == [0] * 4
    s, its = O.solve(O.Scene(d), O.default_options(linear_solver=linear_solver))
    assert s.termination_type != 2 and its[1].step_is_valid
    assert s.final_cost < s.initial_cost
