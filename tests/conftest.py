import importlib
import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
GOLDEN = os.path.join(ROOT, "tests", "golden", "ceres_ref.npz")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


@pytest.fixture(scope="session")
def mba():
    return importlib.import_module("3dreconstruction_amd")


@pytest.fixture(scope="session")
def O():
    import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def golden():
    return np.load(GOLDEN, allow_pickle=False)


@pytest.fixture(scope="session")
def c1(O):
    return O.generate(O.preset(1))


@pytest.fixture(scope="session")
def huber_scene(O):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import HUBER_SPEC
    return O.generate(O.make_spec(**HUBER_SPEC))


@pytest.fixture(scope="session")
def tiny_scene(O):
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    from make_golden import tiny_scene
    return tiny_scene()


def ba_problem(mba, d, const=None):
    """Fresh (copied) parameter arrays so a solve never touches the fixture."""
    return mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"], const)


def grouped_problem(mba, d, const=None, intr_const=None, subset=None):
    """BAProblem of a grouped-intrinsics scene (fresh parameter copies)."""
    return mba.BAProblem(d["cam_init"].copy(), d["pt_init"].copy(), d["obs"], d["cam_idx"], d["pt_idx"], const,
                         intrinsics=d["intr_init"].copy(), intrinsic_index=d["intr_idx"],
                         intrinsic_constant=intr_const, intrinsic_subset=subset)


def axis_point_scene():
    """Three cameras with rotation exactly zero (R = I) at x = -1, 0, 1, 40
    points in front of them seen by all three, and one more point seen
    only by camera 1, on its optical axis (camera-frame x = y = 0)."""
    rng = np.random.default_rng(5)
    cams = np.zeros((3, 9))
    cams[:, 3] = [-1.0, 0.0, 1.0]
    cams[:, 6] = 500.0
    pts = np.column_stack([rng.uniform(-2, 2, 40), rng.uniform(-2, 2, 40), rng.uniform(5, 10, 40)])
    pts = np.vstack([pts, [[0.0, 0.0, 7.0]]])
    cam_idx, pt_idx, obs = [], [], []
    for q in range(40):
        for c in range(3):
            xc = pts[q] + cams[c, 3:6]
            cam_idx.append(c)
            pt_idx.append(q)
            obs.append(500.0 * xc[:2] / xc[2] + rng.normal(0.0, 0.5, 2))
    cam_idx.append(1)
    pt_idx.append(40)
    obs.append([0.4, -0.3])
    pts_init = pts + rng.normal(0.0, 0.01, pts.shape)
    pts_init[40] = pts[40]  # stays on the axis
    return cams, pts_init, np.array(obs), np.array(cam_idx, np.int32), np.array(pt_idx, np.int32)
