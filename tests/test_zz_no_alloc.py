"""kf_sptrsv_jacobi_noalloc, the triangular solve that allocates nothing,
against ComputeSparseTriangularSolve in the same Jacobi mode, both through
the C-ABI:
  * bit for bit with FP64, FP32 and BF16 factors, 3 sweeps each way, in two
    calls in a row: the 48x16x16 lattice (12,288 vertices), then the smaller
    reference mesh uns3dA (2,800 vertices), whose call runs over what the
    first call left in the fixed array;
  * a mesh beyond the array's limit with FP64 factors (the 60x60x50 lattice,
    180,000 vertices > 174,762) is refused with KF_NOALLOC_ESIZE and x is
    left alone; the same mesh with FP32 factors fits and matches;
  * a context with exact triangular solves (jacobi_sweeps 0) is refused with
    KF_NOALLOC_EMODE.
The calls share one array per GPU, so each is followed by a synchronisation
of its context before another context's call (the caller orders them).
"""
import numpy as np
import pytest

import kfun3d_amd as K
from conftest import lattice, perturbed_state

pytestmark = pytest.mark.gpu

ESIZE, EMODE = 1, 2
SWEEPS = 3


def _bitwise(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _factored(path, prec, seed):
    m = K.Mesh(path, K.options(prec=prec, jacobi_sweeps=SWEEPS))
    m.set_state(perturbed_state(m.N, 3.0, seed))
    m.timestep(50.0)
    m.fill()
    m.ilu()
    return m


def _both(m, seed):
    """(x from ComputeSparseTriangularSolve, b device vector) for a seeded b"""
    b = np.random.default_rng(seed).standard_normal(m.sz)
    bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
    m.sptrsv(bd, xd)
    x = xd.numpy()
    xd.free()
    assert m.last_error() == 0
    return x, bd, b


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_noalloc_matches_allocating_solve(prec, mesh_a, mesh_c1):
    big = _factored(mesh_c1, prec, 61)
    small = _factored(mesh_a, prec, 62)
    try:
        assert small.N < big.N
        ref_big, b_big, hb_big = _both(big, 63)
        ref_small, b_small, hb_small = _both(small, 64)
        x_big, x_small = K.DeviceVec(big.sz), K.DeviceVec(small.sz)
        # first call: the larger mesh
        assert big.sptrsv_noalloc(b_big, x_big) == 0
        big.synchronize()
        # second call: the smaller mesh, over the first call's leftovers
        assert small.sptrsv_noalloc(b_small, x_small) == 0
        small.synchronize()
        got_big, got_small = x_big.numpy(), x_small.numpy()
        assert _bitwise(got_big, ref_big), (prec, "48x16x16", np.abs(got_big - ref_big).max())
        assert _bitwise(got_small, ref_small), (prec, "uns3dA", np.abs(got_small - ref_small).max())
        # the inputs are left alone
        assert _bitwise(b_big.numpy(), hb_big) and _bitwise(b_small.numpy(), hb_small)
        # and again on the larger mesh after the smaller one: the same bits
        assert big.sptrsv_noalloc(b_big, x_big) == 0
        big.synchronize()
        assert _bitwise(x_big.numpy(), ref_big), prec
        assert big.last_error() == 0 and small.last_error() == 0
        for v in (b_big, b_small, x_big, x_small):
            v.free()
    finally:
        small.close()
        big.close()


def test_noalloc_limits(mesh_dir, mesh_a):
    path = lattice(mesh_dir, 60, 60, 50, 5)
    m = _factored(path, "fp64", 71)
    try:
        assert m.N == 180000
        b = np.random.default_rng(72).standard_normal(m.sz)
        bd = K.DeviceVec(m.sz, b)
        marker = np.full(m.sz, 7.0)
        xd = K.DeviceVec(m.sz, marker)
        assert m.sptrsv_noalloc(bd, xd) == ESIZE
        m.synchronize()
        assert _bitwise(xd.numpy(), marker)
        # with FP32 factors the work vectors take half the bytes and fit
        m.set_options(K.options(prec="fp32", jacobi_sweeps=SWEEPS))
        m.set_state(perturbed_state(m.N, 3.0, 71))
        m.timestep(50.0)
        m.fill()
        m.ilu()
        ref = K.DeviceVec(m.sz)
        m.sptrsv(bd, ref)
        assert m.sptrsv_noalloc(bd, xd) == 0
        m.synchronize()
        assert _bitwise(xd.numpy(), ref.numpy())
        assert m.last_error() == 0
        for v in (bd, xd, ref):
            v.free()
    finally:
        m.close()
    # exact triangular solves: no Jacobi sweeps to run
    e = K.Mesh(mesh_a, K.options())
    try:
        e.fill()
        e.ilu()
        bd, xd = K.DeviceVec(e.sz, np.ones(e.sz)), K.DeviceVec(e.sz)
        assert e.sptrsv_noalloc(bd, xd) == EMODE
        bd.free()
        xd.free()
    finally:
        e.close()
