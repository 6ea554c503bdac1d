"""Jacobi-sweep triangular solves (kf_options.jacobi_sweeps; the reference's
fun3d::Matrix::sptrsv_Jacobi, matrix/sptrsv.cpp:75-132) against the oracle.

  * FP64: k sweeps each way over the ILU(0) factors, bit for bit with the
    oracle's or_sptrsv_jacobi (the CPU tests pin it to its dense closed form
    and to the exact solve once k reaches the DAG depth), k = 1, 2, 3, 5;
  * FP32 factors and arithmetic: within 3e-6 relative of the oracle's FP64
    sweeps; BF16 factors: within 2e-2;
  * a whole solve with 3 sweeps in FP32 and BF16 reaches the 1e-10 drop with
    the final forces within 1e-8 of the oracle's FP64 exact-solve run, on
    mesh A and the 48x16x16 lattice.
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import perturbed_state

pytestmark = pytest.mark.gpu


def _bitwise(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _gpu_jacobi(m, prec, k, q, b):
    m.set_options(K.options(prec=prec, jacobi_sweeps=k))
    m.set_state(q)
    m.timestep(50.0)
    m.fill()
    m.ilu()
    bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
    m.sptrsv(bd, xd)
    assert m.last_error() == 0
    x = xd.numpy()
    # the input is left alone
    assert _bitwise(bd.numpy(), b)
    bd.free()
    xd.free()
    return x


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_jacobi_sweeps_against_oracle(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    m = K.Mesh(path, K.options())
    orc = O.Oracle(m)
    try:
        q = perturbed_state(m.N, 3.0, 51)
        b = np.random.default_rng(52).standard_normal(m.sz)
        orc.set_state(q)
        orc.timestep(50.0)
        orc.fill()
        orc.ilu()
        exact = orc.sptrsv(b)
        for k in (1, 2, 3, 5):
            xo = orc.sptrsv_jacobi(b, k)
            x = _gpu_jacobi(m, "fp64", k, q, b)
            assert _bitwise(x, xo), (which, k, np.abs(x - xo).max())
            for prec, bar in (("fp32", 3e-6), ("bf16", 2e-2)):
                xs = _gpu_jacobi(m, prec, k, q, b)
                e = np.linalg.norm(xs - xo) / np.linalg.norm(xo)
                assert e <= bar, (which, prec, k, e)
        # more sweeps, closer to the exact solve
        e1 = np.linalg.norm(orc.sptrsv_jacobi(b, 1) - exact)
        e5 = np.linalg.norm(orc.sptrsv_jacobi(b, 5) - exact)
        assert e5 < e1
    finally:
        orc.close()
        m.close()


def test_exact_solve_unchanged_after_jacobi(mesh_c1):
    """switching jacobi_sweeps back to 0 gives the level-scheduled solve bit
    for bit (the forward buffer's sentinel state is rebuilt)"""
    m = K.Mesh(mesh_c1, K.options())
    orc = O.Oracle(m)
    try:
        q = perturbed_state(m.N, 3.0, 53)
        b = np.random.default_rng(54).standard_normal(m.sz)
        _gpu_jacobi(m, "fp64", 2, q, b)
        m.set_options(K.options())
        m.set_state(q)
        m.timestep(50.0)
        m.fill()
        m.ilu()
        bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
        m.sptrsv(bd, xd)
        x = xd.numpy()
        bd.free()
        xd.free()
        orc.set_state(q)
        orc.timestep(50.0)
        orc.fill()
        orc.ilu()
        assert _bitwise(x, orc.sptrsv(b))
        assert m.last_error() == 0
    finally:
        orc.close()
        m.close()


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_jacobi_solve_converges_to_same_forces(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    first = 1 if which == "mesh_c1" else 0
    mh = K.Mesh(path, device=False)
    orc = O.Oracle(mh, **O.config_options(first_order=first))
    ref = orc.kernel()
    orc.close()
    mh.close()
    f0 = np.array(ref[-1]["forces"])
    for prec in ("fp32", "bf16"):
        m = K.Mesh(path, K.config_options(prec=prec, first_order=first, jacobi_sweeps=3))
        try:
            h = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
        assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10, (which, prec)
        f = np.array(h[-1]["forces"])
        assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (which, prec, f, f0)
        print("%s %s jacobi 3: %d steps, %d Krylov (exact FP64: %d steps, %d Krylov)"
              % (which, prec, len(h), sum(s["krylov"] for s in h), len(ref), sum(s["krylov"] for s in ref)))
