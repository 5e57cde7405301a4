"""Boundary conversions and problem files (host code, no GPU):
  * RotationMatrixToAngleAxis / AngleAxisToRotationMatrix against the
    reference's Ceres on random rotations and the special cases (identity,
    tiny angles, exactly and nearly pi);
  * the libmv problem files (real data, 3rdparty/ceres/data/libmv-ba-problems)
    read by the library against an independent parse, and their Schur
    structure under a points-first ordering against the reference's Ceres;
  * BAL text: the reader against a plain parse of the text, the writer
    byte for byte against the reference's BALProblem::WriteToFile (digest);
  * trust-region problem dumps (TEXTFILE): option checks as Ceres's
    (solver.cc:243-247) and the reader on files of that layout;
  * CLUSTER_JACOBI / CLUSTER_TRIDIAGONAL rejected with the reference's
    termination and message (its Ceres has no SuiteSparse).
"""
import hashlib
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def mg(O):
    sys.path.insert(0, GOLDEN)
    import make_golden
    return make_golden


def test_rotation_to_angle_axis_matches_reference(mba, golden):
    aa = mba.rotation_matrix_to_angle_axis(golden["rot_R"])
    np.testing.assert_allclose(aa, golden["rot_aa"], rtol=0, atol=2e-15)
    # Round trip and the closed forms: |aa| <= pi, R(aa) = R.
    assert np.all(np.linalg.norm(aa, axis=1) <= np.pi + 1e-12)
    np.testing.assert_allclose(mba.angle_axis_to_rotation_matrix(aa), golden["rot_R"], rtol=0, atol=1e-14)


def test_angle_axis_to_rotation_matrix_closed_forms(mba):
    # Quarter turn about z; the first-order branch below sqrt(DBL_EPSILON).
    R = mba.angle_axis_to_rotation_matrix([[0, 0, np.pi / 2], [1e-9, 0, 0]])
    np.testing.assert_allclose(R[0], [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-16)
    np.testing.assert_array_equal(R[1], [[1, 0, 0], [0, 1, -1e-9], [0, 1e-9, 1]])
    rng = np.random.default_rng(3)
    Rs = mba.angle_axis_to_rotation_matrix(rng.normal(size=(50, 3)))
    np.testing.assert_allclose(Rs @ np.transpose(Rs, (0, 2, 1)), np.tile(np.eye(3), (50, 1, 1)), atol=1e-14)
    np.testing.assert_allclose(np.linalg.det(Rs), 1.0, atol=1e-14)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_libmv_reader_matches_independent_parse(mba, golden, mg, k):
    path = os.path.join(GOLDEN, "libmv", mg.LIBMV_FILES[k])
    d = mba.read_libmv(path)
    ref = mg.libmv_problem(path)
    assert d["image_space"]
    np.testing.assert_array_equal(d["obs"], ref["obs"])
    np.testing.assert_array_equal(d["cam_idx"], ref["cam_idx"])
    np.testing.assert_array_equal(d["pt_idx"], ref["pt_idx"])
    np.testing.assert_array_equal(d["points"], ref["pt_init"])
    np.testing.assert_array_equal(d["intrinsics"], ref["intr_init"])
    np.testing.assert_array_equal(d["cameras"][:, 3:], ref["cam_init"][:, 3:])
    # Rotations: the library's conversion against the reference's.
    np.testing.assert_allclose(d["cameras"][:, :3], golden[f"libmv_{k}_aa"], rtol=0, atol=2e-15)


def libmv_ba(mba, d, ordering):
    const = np.zeros(d["cameras"].shape[0], np.uint8)
    const[d["cam_idx"][0]] = 1  # the first marker's camera (libmv_bundle_adjuster.cc:715-718)
    return mba.BAProblem(d["cameras"].copy(), d["points"].copy(), d["obs"], d["cam_idx"], d["pt_idx"], const,
                         intrinsics=d["intrinsics"].copy(), intrinsic_index=d["intr_idx"],
                         intrinsic_subset=np.array([0x26], np.uint8), ordering=ordering)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_libmv_structure_with_points_first_ordering(mba, golden, mg, k):
    """Few points seen by hundreds of cameras: the automatic independent set
    eliminates the cameras (unsupported here), a points-first ordering
    (linear_solver_ordering) reproduces the reference's structure."""
    d = mba.read_libmv(os.path.join(GOLDEN, "libmv", mg.LIBMV_FILES[k]))
    with pytest.raises(mba.MbaError) as e:
        mba.solve(libmv_ba(mba, d, mba.ORDER_AUTOMATIC), mba.default_options(max_num_iterations=1))
    assert e.value.status == mba.ERR_UNSUPPORTED
    ne, order, res, pairs = mba.schur_structure(libmv_ba(mba, d, mba.ORDER_POINTS_FIRST))
    assert ne == int(golden[f"libmv_{k}_struct_num_e"]) == d["points"].shape[0]
    np.testing.assert_array_equal(order, golden[f"libmv_{k}_struct_order"])
    np.testing.assert_array_equal(res, golden[f"libmv_{k}_struct_residual_order"])
    assert pairs.shape[0] == int(golden[f"libmv_{k}_struct_num_pairs"])
    assert hashlib.sha256(np.ascontiguousarray(pairs, np.int32).tobytes()).hexdigest() == \
        str(golden[f"libmv_{k}_struct_pairs_sha256"])


def test_bal_reader_and_writer(mba, golden, mg, tmp_path):
    d = mba.read_bal(mg.BAL_FILE)
    lines = open(mg.BAL_FILE).read().split("\n")
    nc, np_, no = (int(v) for v in lines[0].split())
    obs = np.array([[float(v) for v in ln.split()[2:]] for ln in lines[1:1 + no]])
    idx = np.array([[int(v) for v in ln.split()[:2]] for ln in lines[1:1 + no]])
    vals = np.array([float(v) for v in lines[1 + no:] if v.strip()])
    assert d["cameras"].shape == (nc, 9) and d["points"].shape == (np_, 3)
    np.testing.assert_array_equal(d["obs"], -obs)  # BAL cameras look down -z
    np.testing.assert_array_equal(d["cam_idx"], idx[:, 0])
    np.testing.assert_array_equal(d["pt_idx"], idx[:, 1])
    np.testing.assert_array_equal(np.concatenate([d["cameras"].ravel(), d["points"].ravel()]), vals)
    out = tmp_path / "w.txt"
    mba.write_bal(str(out), mba.BAProblem(d["cameras"], d["points"], d["obs"], d["cam_idx"], d["pt_idx"]))
    assert hashlib.sha256(out.read_bytes()).hexdigest() == str(golden["bal_write_sha256"])
    with pytest.raises(mba.MbaError):
        mba.read_bal(str(tmp_path / "missing.txt"))


@pytest.mark.parametrize("kind,name", [(2, "cluster_jacobi"), (3, "cluster_tridiagonal")])
@pytest.mark.parametrize("solver", [0, 1])
def test_cluster_preconditioners_rejected_as_the_reference(mba, golden, mg, kind, name, solver):
    """The reference's Solve fails option validation (FAILURE, no iterations,
    solution not usable); the library refuses before any device work with
    the same message."""
    assert int(golden[f"{name}_{solver}_termination"]) == mba.FAILURE
    assert int(golden[f"{name}_{solver}_num_iterations"]) == 0 and not bool(golden[f"{name}_{solver}_usable"])
    d = mg.dump_scene()
    p = mba.BAProblem(d["cam_init"], d["pt_init"], d["obs"], d["cam_idx"], d["pt_idx"])
    with pytest.raises(mba.MbaError) as e:
        mba.solve(p, mba.default_options(linear_solver=solver, preconditioner=kind, huber_a=4.0))
    assert e.value.status == mba.ERR_UNSUPPORTED
    assert str(e.value).endswith(str(golden[f"{name}_{solver}_message"]))


def test_dump_options_are_checked_before_device_work(mba, mg, tmp_path):
    d = mg.dump_scene()
    p = mba.BAProblem(d["cam_init"], d["pt_init"], d["obs"], d["cam_idx"], d["pt_idx"])
    for directory in ("", None):
        with pytest.raises(mba.MbaError) as e:
            mba.solve(p, mba.default_options(dump_directory=directory, dump_iterations=[0]))
        assert e.value.status == mba.ERR_INVALID_ARGUMENT
        assert str(e.value).endswith("Solver::Options::trust_region_problem_dump_directory is empty.")
    with pytest.raises(mba.MbaError) as e:
        mba.solve_batched([p, p], mba.default_options(linear_solver=mba.IMPLICIT_SCHUR_PCG,
                                                      dump_directory=str(tmp_path), dump_iterations=[0]))
    assert e.value.status == mba.ERR_UNSUPPORTED


def test_trust_region_problem_reader(mba, tmp_path):
    """Files in the layout of DumpLinearLeastSquaresProblemToTextFile
    (linear_least_squares_problems.cc:534-603; "% 10d % 10d %17f" entries)."""
    base = str(tmp_path / "ceres_solver_iteration_004")
    rows, cols, vals = [0, 0, 1, 3], [0, 2, 1, 4], [1.5, -0.25, 1e-6, 123456.789]
    with open(base + "_A.txt", "w") as f:
        for r, c, v in zip(rows, cols, vals):
            f.write("% 10d % 10d %17f\n" % (r, c, v))
    for key, arr in [("D", [0.1, 0.2, 0.3, 0.4, 0.5]), ("b", [1.0, -2.0, 3.0, -4.0]), ("x", [5, 4, 3, 2, 1])]:
        with open(f"{base}_{key}.txt", "w") as f:
            f.writelines("%17f\n" % v for v in arr)
    with open(base + ".m", "w") as f:
        f.write("function lsqp = load_trust_region_problem()\nlsqp.num_rows = 4;\nlsqp.num_cols = 5;\n")
    t = mba.read_trust_region_problem(base)
    assert (t["num_rows"], t["num_cols"]) == (4, 5)
    np.testing.assert_array_equal(t["rows"], rows)
    np.testing.assert_array_equal(t["cols"], cols)
    np.testing.assert_allclose(t["values"], vals, atol=5e-7)
    np.testing.assert_allclose(t["x"], [5, 4, 3, 2, 1])
    assert t["D"].shape == (5,) and t["b"].shape == (4,)
