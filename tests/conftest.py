import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
ORACLE_DIR = os.path.join(ROOT, "oracle")
if ORACLE_DIR not in sys.path:
    sys.path.insert(0, ORACLE_DIR)
TESTS_DIR = os.path.dirname(os.path.abspath(__file__))
if TESTS_DIR not in sys.path:
    sys.path.insert(0, TESTS_DIR)

REF_ROOT = "/root/reference"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhmv.so on the device)")
    config.addinivalue_line("markers", "slow: CPU test taking more than a few seconds")


@pytest.fixture(scope="session")
def H():
    import __graft_entry__ as g
    return g.load_package()


@pytest.fixture(scope="session")
def O():
    import oracle
    return oracle


@pytest.fixture(scope="session")
def gpu(H):
    """Initialises device 0; GPU tests fail (not skip) when the device is missing."""
    rc = H.hmv_lib().hmv_init(0)
    if rc != 0:
        raise RuntimeError("GPU test without a usable device: " + H.hmv_lib().hmv_last_error().decode())
    return H


@pytest.fixture(scope="session")
def sphere3(H):
    nodes, faces = H.icosphere(3)
    return H.HMatrix(nodes, faces, H.baseline_params(aca_eps=1e-4))


@pytest.fixture(scope="session")
def sphere4(H):
    """The CPU-runnable benchmark case: icosphere s=4 (5,120 triangles), leaf 20,
    eta 2, ACA+ 1e-4, 1/(4 pi r) centroid kernel."""
    nodes, faces = H.icosphere(4)
    return H.HMatrix(nodes, faces, H.baseline_params(aca_eps=1e-4))


@pytest.fixture(scope="session")
def two_spheres4(H):
    """Two disjoint s=4 icospheres with centres 3 apart (small stand-in of the
    163,840-unknown two-sphere case)."""
    m = H.merge(H.icosphere(4), H.icosphere(4, 1.0, (3.0, 0.0, 0.0)))
    return H.HMatrix(*m, H.baseline_params(aca_eps=1e-5))


@pytest.fixture(scope="session")
def ref_mesh_sample(H):
    """The reference's input_sample.pbf (600 triangles, r=0.75), ppohBEM kernel."""
    path = os.path.join(ROOT, "tests", "golden", "input_sample_mesh.npz")
    import numpy as np
    d = np.load(path)
    return H.HMatrix(d["nodes"], d["faces"], H.HParams(leaf_size=15, aca_eps=1e-6, kernel="ppohbem"))
