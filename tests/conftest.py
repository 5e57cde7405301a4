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
