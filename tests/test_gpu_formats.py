"""GPU solves of problems read from files, against the reference's Ceres on
the same files:
  * the three libmv camera-refinement problems (real data; grouped layout,
    f, k1, k2 refined, first marker's camera held, no loss, points-first
    ordering, SPARSE_SCHUR);
  * a BAL file solved by the reference with its SnavelyReprojectionError
    functor; the library reads it with negated observations (BAL cameras
    look down -z), which leaves the cost and the LM trajectory unchanged.

Tolerances: LM costs 1e-9 relative with equal step counts and terminations
for the BAL problem; the libmv problems (focal lengths of 1.7e3-6.3e3 px
refined from 26-71 points) 1e-8 per iteration; parameters as stated.
"""
import os
import sys

import numpy as np
import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def mg(O):
    sys.path.insert(0, GOLDEN)
    import make_golden
    return make_golden


@pytest.mark.parametrize("k", [0, 1, 2])
def test_libmv_problem_matches_reference(mba, golden, mg, k):
    d = mba.read_libmv(os.path.join(GOLDEN, "libmv", mg.LIBMV_FILES[k]))
    const = np.zeros(d["cameras"].shape[0], np.uint8)
    const[d["cam_idx"][0]] = 1
    p = mba.BAProblem(d["cameras"].copy(), d["points"].copy(), d["obs"], d["cam_idx"], d["pt_idx"], const,
                      intrinsics=d["intrinsics"].copy(), intrinsic_index=d["intr_idx"],
                      intrinsic_subset=np.array([mg.LIBMV_SUBSET], np.uint8), ordering=mba.ORDER_POINTS_FIRST)
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=mba.EXPLICIT_SCHUR))
    ref = golden[f"libmv_{k}_iterations"]
    assert s.IsSolutionUsable()
    assert (s.num_successful_steps, s.num_unsuccessful_steps, s.termination_type) == (
        int(golden[f"libmv_{k}_successful"]), int(golden[f"libmv_{k}_unsuccessful"]),
        int(golden[f"libmv_{k}_termination"]))
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-8)
    np.testing.assert_allclose(p.intrinsics, golden[f"libmv_{k}_intrinsics"], rtol=1e-6)
    # Held entries (ppx, ppy, k3) and the held camera bit for bit.
    np.testing.assert_array_equal(p.intrinsics[0, [1, 2, 5]], d["intrinsics"][0, [1, 2, 5]])
    np.testing.assert_array_equal(p.cameras[const == 1], d["cameras"][const == 1])


def test_bal_file_matches_reference_snavely_functor(mba, golden, mg):
    d = mba.read_bal(mg.BAL_FILE)
    p = mba.BAProblem(d["cameras"].copy(), d["points"].copy(), d["obs"], d["cam_idx"], d["pt_idx"])
    s = mba.solve(p, mba.default_options(max_num_iterations=20, linear_solver=mba.EXPLICIT_SCHUR, huber_a=4.0))
    ref = golden["bal_sparse_iterations"]
    assert (s.num_successful_steps, s.num_unsuccessful_steps, s.termination_type) == (
        int(golden["bal_sparse_successful"]), int(golden["bal_sparse_unsuccessful"]),
        int(golden["bal_sparse_termination"]))
    np.testing.assert_allclose([it.cost for it in s.iterations], ref[:, 1], rtol=1e-9)
    np.testing.assert_allclose(s.final_cost, float(golden["bal_sparse_final_cost"]), rtol=1e-10)
    np.testing.assert_allclose(p.cameras, golden["bal_sparse_cameras"], rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(p.points, golden["bal_sparse_points"], rtol=1e-6, atol=1e-8)
