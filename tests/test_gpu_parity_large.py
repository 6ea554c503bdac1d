"""HIP path against the oracle at the production lattice sizes.

The 144x48x48 lattice (seed 1, 331,776 vertices, 2,258,399 edges) and the
288x96x96 lattice (seed 4, 2,654,208 vertices, 18,322,367 edges), on a
seeded state (free stream plus 0.05 U(-1, 1) per entry, second order):
  * ComputeFlux (LSQ gradient and residual, src/flux.c:60-779),
    ComputeTimeStep (src/timestep.c:21-188), ComputeA (src/fill.c:115-917),
    the BSR product y = A x of the assembled Jacobian (kf_bsr_spmv), the
    reference's one-sweep ILU (matrix/ilu.cpp:49-138), one
    ComputeSparseTriangularSolve (matrix/sptrsv.cpp:34-73) and
    ComputeForces (src/forces.c:8-120): bit for bit;
  * on the 144x48x48 lattice, the first two pseudo-time steps of the
    configuration profile (alpha 3.06 deg, CFL ramp from 10): every Krylov
    count, CFL, residual norm and force bit for bit against the oracle run
    with the device's reduction tree.
The oracle runs one thread, so each lattice costs a few tens of seconds of
host time (the largest about a minute and 15 GB of host memory).
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import lattice, perturbed_state

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

LATTICES = {"c2": ((144, 48, 48), 1), "c5": ((288, 96, 96), 4)}


def _bitwise(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _describe(a, b):
    d = np.abs(a - b)
    i = int(np.argmax(d))
    return "max |diff| %.3e at %d (%r vs %r), %d of %d differ" % (d[i], i, a[i], b[i], int(np.sum(a != b)), a.size)


@pytest.fixture(scope="module", params=sorted(LATTICES))
def big(request, mesh_dir):
    (nx, ny, nz), seed = LATTICES[request.param]
    print("lattice %s %dx%dx%d seed %d: setup" % (request.param, nx, ny, nz, seed), flush=True)
    path = lattice(mesh_dir, nx, ny, nz, seed)
    m = K.Mesh(path, K.options(bsr_operator=1))   # the fill also keeps A for the BSR product
    orc = O.Oracle(m)
    yield request.param, m, orc
    orc.close()
    m.close()


@pytest.mark.timeout(900)
def test_kernels_bitwise_at_size(big):
    name, m, orc = big
    q = perturbed_state(m.N, 3.0, 31)
    qd, rd = K.DeviceVec(m.sz, q), K.DeviceVec(m.sz)
    g = [K.DeviceVec(m.sz) for _ in range(3)]
    print("%s: residual" % name, flush=True)
    m.flux(qd, g, rd)
    r = rd.numpy()
    ro, gx, gy, gz = orc.flux(q)
    for a, b, what in [(g[0].numpy(), gx, "grad x"), (g[1].numpy(), gy, "grad y"), (g[2].numpy(), gz, "grad z"),
                       (r, ro, "residual")]:
        assert _bitwise(a, b), "%s %s: %s" % (name, what, _describe(a, b))
    for v in g + [qd]:
        v.free()

    print("%s: time step, fill" % name, flush=True)
    m.set_state(q)
    orc.set_state(q)
    m.timestep(50.0)
    cdt = orc.timestep(50.0)
    assert _bitwise(m.cdt(), cdt), "%s cdt: %s" % (name, _describe(m.cdt(), cdt))
    m.fill()
    orc.fill()
    A, Ao = m.blocks().ravel(), orc.blocks().ravel()
    assert _bitwise(A, Ao), "%s fill: %s" % (name, _describe(A, Ao))
    del A, Ao
    print("%s: BSR product" % name, flush=True)
    xs = np.random.default_rng(32).standard_normal(m.sz)
    xsd, ysd = K.DeviceVec(m.sz, xs), K.DeviceVec(m.sz)
    m.spmv(xsd, ysd)
    y, yo = ysd.numpy(), orc.spmv(xs)
    assert _bitwise(y, yo), "%s spmv: %s" % (name, _describe(y, yo))
    xsd.free()
    ysd.free()
    del y, yo, xs

    print("%s: ILU(0)" % name, flush=True)
    m.ilu()
    orc.ilu()
    F, Fo = m.blocks().ravel(), orc.blocks().ravel()
    assert _bitwise(F, Fo), "%s ILU: %s" % (name, _describe(F, Fo))
    del F, Fo

    print("%s: triangular solve, forces" % name, flush=True)
    xd = K.DeviceVec(m.sz)
    m.sptrsv(rd, xd)
    x, xo = xd.numpy(), orc.sptrsv(ro)
    assert _bitwise(x, xo), "%s sptrsv: %s" % (name, _describe(x, xo))
    f, fo = m.forces(), orc.forces()
    assert _bitwise(f, fo), (name, f, fo)
    assert m.last_error() == 0
    xd.free()
    rd.free()


@pytest.mark.timeout(900)
def test_first_pseudo_steps_bitwise_at_size(big):
    """The configuration profile's first two pseudo-time steps (residual,
    CFL, time step, fill, ILU, GMRES(30) with finite-difference products and
    modified Gram-Schmidt, update, forces) against the oracle with the
    device's reduction tree, bit for bit.  Only on the 144x48x48 lattice:
    at the largest size the oracle's Krylov iterations would take minutes."""
    name, big_m, _ = big
    if name != "c2":
        pytest.skip("history only on the 144x48x48 lattice (oracle time)")
    print("%s: two pseudo-time steps" % name, flush=True)
    # a fresh context and a fresh oracle: the reference's ILU sweep starts
    # from the previous factorisation's L blocks, so both start from none
    m = K.Mesh(big_m.path, K.config_options(ts_max=2))
    try:
        ot = O.Oracle(m, **O.config_options(ts_max=2, reduce_order=1))
        try:
            ho = ot.kernel()
        finally:
            ot.close()
        hg = m.kernel()
        assert m.last_error() == 0
    finally:
        m.close()
    assert len(hg) == len(ho) == 2, (len(hg), len(ho))
    for a, b in zip(hg, ho):
        assert a["krylov"] == b["krylov"], (a, b)
        assert a["cfl"] == b["cfl"], (a, b)
        assert a["fnorm"] == b["fnorm"], (a, b)
        assert a["forces"] == b["forces"], (a, b)
    print("%s: steps %r" % (name, [(h["krylov"], h["fnorm"]) for h in hg]), flush=True)
