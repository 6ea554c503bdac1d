"""Pins the oracle and the product's host setup to the reference itself.

tests/golden/ref_*.json were written by tests/golden/make_golden.py from the
reference's own sources compiled in the development container (one thread,
mesh/uns3dA.msh and the 48x16x16 lattice of configuration 1).  Every check
here is bit for bit (SHA-256 of the IEEE bytes), except where stated.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import GOLDEN

MESHES = ["mesh_a", "c1"]


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()


def golden(name):
    with open(os.path.join(GOLDEN, "ref_%s.json" % name)) as f:
        return json.load(f)


def splitmix_uniforms(state, n):
    """The harness's splitmix64 stream: returns n draws in [0,1) and the new state."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(state) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        new_state = np.uint64(state) + np.uint64(n) * np.uint64(0x9E3779B97F4A7C15)
    return u, int(new_state)


def kernel_inputs(N, U, V, seed=7):
    """The seeded states and vector of the golden harness (oracle/ref_driver.c)."""
    sz = 4 * N
    q = np.tile(np.array([1.0, U, V, 0.0]), N)
    u, st = splitmix_uniforms(seed, sz)
    q = q + 0.05 * (2.0 * u - 1.0)
    u, st = splitmix_uniforms(st, sz)
    x = 2.0 * u - 1.0
    u, st = splitmix_uniforms(st, sz)
    q2 = q + 0.01 * (2.0 * u - 1.0)
    return q, x, q2


@pytest.fixture(scope="module", params=MESHES)
def case(request, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[request.param]
    m = K.Mesh(path, device=False)
    yield request.param, m, golden(request.param)
    m.close()


def test_generator_is_pinned(mesh_c1):
    g = golden("c1")
    assert hashlib.sha256(open(mesh_c1, "rb").read()).hexdigest() == g["msh_sha256"]


def test_host_setup_matches_reference(case):
    name, m, g = case
    h = m.host_arrays()
    for k, digest in g["setup"].items():
        assert sha(h[k]) == digest, "%s: setup array %s differs from the reference" % (name, k)
    U, V = np.cos(3.0 / (180 / np.pi)), np.sin(3.0 / (180 / np.pi))
    import math
    assert [math.cos(3.0 / (180 / math.pi)), math.sin(3.0 / (180 / math.pi))] == g["phy_uv"]


def test_oracle_kernels_match_reference(case):
    name, m, g = case
    U, V = g["phy_uv"]
    q, x, q2 = kernel_inputs(m.N, U, V)
    k = g["kernels"]
    orc = O.Oracle(m)
    r, gx, gy, gz = orc.flux(q)
    out = {"r": r, "grad_x0": gx, "grad_x1": gy, "grad_x2": gz}
    orc.set_state(q)
    out["cdt"] = orc.timestep(50.0)
    orc.fill()
    out["a_fill"] = orc.blocks().ravel()
    orc.ilu()
    out["a_ilu"] = orc.blocks().ravel()
    out["trsv_x"] = orc.sptrsv(r)
    qnorm = np.sqrt(1.0 + O.norm(q))
    assert qnorm == k["jv_qnorm"]
    w, y = orc.ijacvec(x, r, qnorm)
    out["jv_w"], out["jv_y"] = w, y
    out["forces"] = orc.forces()
    orc.set_state(q2)
    out["cdt2"] = orc.timestep(500.0)
    orc.fill()
    out["a_fill2"] = orc.blocks().ravel()
    orc.ilu()
    out["a_ilu2"] = orc.blocks().ravel()
    out["trsv_x2"] = orc.sptrsv(r)
    orc.close()
    for key, arr in out.items():
        assert sha(arr) == k[key]["sha256"], "%s: %s (|.|=%r, reference %r)" % (name, key, np.linalg.norm(arr),
                                                                              k[key]["norm"])


def test_oracle_first_order_matches_reference(case):
    """Zero LSQ terms make the reference's gradient exactly zero: its residual is
    the first-order one, which the oracle computes with first_order=1."""
    name, m, g = case
    U, V = g["phy_uv"]
    q, _, _ = kernel_inputs(m.N, U, V)
    orc = O.Oracle(m, first_order=1)
    r, gx, _, _ = orc.flux(q)
    orc.close()
    assert not gx.any()
    assert sha(r) == g["kernels"]["r_first_order"]["sha256"]


def test_oracle_solve_matches_reference(case):
    name, m, g = case
    orc = O.Oracle(m)
    hist = orc.kernel()
    orc.close()
    ref = g["solve"]
    assert len(hist) == len(ref)
    for a, b in zip(hist, ref):
        assert a["krylov"] == b["krylov"]
        assert a["cfl"] == b["cfl"]
        assert a["fnorm"] == b["fnorm"]
        assert a["forces"] == b["forces"]
