import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")
    config.addinivalue_line("markers", "slow: larger meshes")


@pytest.fixture(scope="session")
def mesh_dir(tmp_path_factory):
    return tmp_path_factory.mktemp("meshes")


@pytest.fixture(scope="session")
def mesh_a(mesh_dir):
    """The reference's own input mesh (mesh/uns3dA.msh: 2,800 vertices)."""
    out = os.path.join(str(mesh_dir), "uns3dA.msh")
    with gzip.open(os.path.join(GOLDEN, "uns3dA.msh.gz"), "rb") as src, open(out, "wb") as dst:
        shutil.copyfileobj(src, dst)
    return out


def lattice(mesh_dir, nx, ny, nz, seed):
    import kfun3d_amd as K
    out = os.path.join(str(mesh_dir), "lattice_%dx%dx%d_s%d.msh" % (nx, ny, nz, seed))
    if not os.path.exists(out):
        K.write_lattice(out, nx, ny, nz, seed)
    return out


@pytest.fixture(scope="session")
def mesh_c1(mesh_dir):
    """Configuration 1: 48x16x16 Kuhn lattice, seed 0 (12,288 vertices)."""
    return lattice(mesh_dir, 48, 16, 16, 0)


def perturbed_state(N, alpha_deg, seed, amp=0.05):
    """Free stream (1, cos a, sin a, 0) plus amp * U(-1,1) per entry."""
    a = alpha_deg / (180 / np.pi)
    q = np.tile(np.array([1.0, np.cos(a), np.sin(a), 0.0]), N)
    rng = np.random.default_rng(seed)
    return q + amp * (2.0 * rng.random(4 * N) - 1.0)
