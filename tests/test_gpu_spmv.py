"""BSR SpMV y = A x with the assembled Jacobian (kf_bsr_spmv) and the
first-order-operator GMRES mode (kf_options.bsr_operator), against the
oracle.

  * y bit for bit against or_spmv (the same per-row order) on mesh A and the
    48x16x16 lattice, after a fill at a seeded state; and again after the
    ILU has factored the fill's array in place (the product uses the kept
    copy);
  * with the FP32 preconditioner the kept copy is the FP32 Jacobian: y within
    1e-6 relative (2-norm) of the oracle's FP64 product;
  * a solve to a 1e-10 drop with the BSR operator converges to the oracle's
    steady state (the residual stays second order): final forces within 1e-8
    relative of the oracle's matrix-free run.
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import perturbed_state

pytestmark = pytest.mark.gpu


def _bitwise(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_spmv_bitwise(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    m = K.Mesh(path, K.options(bsr_operator=1))
    orc = O.Oracle(m)
    try:
        q = perturbed_state(m.N, 3.0, 41)
        x = np.random.default_rng(42).standard_normal(m.sz)
        m.set_state(q)
        orc.set_state(q)
        m.timestep(30.0)
        orc.timestep(30.0)
        m.fill()
        orc.fill()
        xd, yd = K.DeviceVec(m.sz, x), K.DeviceVec(m.sz)
        m.spmv(xd, yd)
        want = orc.spmv(x)
        assert _bitwise(yd.numpy(), want)
        m.ilu()                        # factors the fill's array; the product keeps A
        m.spmv(xd, yd)
        assert _bitwise(yd.numpy(), want)
        m.set_options(K.options(bsr_operator=1, prec="fp32"))
        m.fill()
        m.spmv(xd, yd)
        y32 = yd.numpy()
        assert np.linalg.norm(y32 - want) <= 1e-6 * np.linalg.norm(want)
        assert m.last_error() == 0
    finally:
        orc.close()
        m.close()


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_bsr_operator_solve_reaches_the_same_state(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    hm = K.Mesh(path, device=False)
    orc = O.Oracle(hm, **O.config_options())
    ref = orc.kernel()
    orc.close()
    hm.close()
    m = K.Mesh(path, K.config_options(bsr_operator=1))
    try:
        h = m.kernel()
        assert m.last_error() == 0
    finally:
        m.close()
    assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10
    f, f0 = np.array(h[-1]["forces"]), np.array(ref[-1]["forces"])
    assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (f, f0)
    print("%s: BSR operator %d pseudo-steps, %d Krylov; matrix-free %d, %d"
          % (which, len(h), sum(s["krylov"] for s in h), len(ref), sum(s["krylov"] for s in ref)))
