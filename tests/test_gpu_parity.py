"""HIP path against the oracle (the one-thread restatement that the CPU tests
pin bit for bit to the reference), on mesh A and the 48x16x16 lattice.

Bars: bit-identical for the FP64 residual, gradient, time step, Jacobian
fill, textbook ILU(0) factors, triangular solve and forces (each expression
rounds as the reference's; the gathers keep its accumulation order).  The
reference's own ILU (ilu_Edmond) adds the (a - d) + d rounding of its copy
side effects, so the HIP factors are held to it at 1e-10 relative.  iJacVec
and the full solve contain tree reductions (norms, dots) whose order differs
from the one-thread sums: tolerances are stated in each test.
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import perturbed_state

pytestmark = pytest.mark.gpu


def _bitwise(a, b):
    return np.array_equal(np.asarray(a).view(np.uint64), np.asarray(b).view(np.uint64))


def _describe(a, b):
    d = np.abs(a - b)
    i = int(np.argmax(d))
    return "max |diff| %.3e at %d (%r vs %r), %d of %d differ" % (d[i], i, a[i], b[i], int(np.sum(a != b)), a.size)


@pytest.fixture(scope="module", params=["mesh_a", "mesh_c1"])
def pair(request):
    path = request.getfixturevalue(request.param)
    m = K.Mesh(path, K.options())
    orc = O.Oracle(m)
    yield request.param, m, orc
    orc.close()
    m.close()


def _flux_gpu(m, q):
    qd = K.DeviceVec(m.sz, q)
    g = [K.DeviceVec(m.sz) for _ in range(3)]
    r = K.DeviceVec(m.sz)
    m.flux(qd, g, r)
    return r.numpy(), [x.numpy() for x in g]


def test_flux_bitwise(pair):
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 1)
    r, g = _flux_gpu(m, q)
    ro, gx, gy, gz = orc.flux(q)
    for a, b, what in [(g[0], gx, "grad x"), (g[1], gy, "grad y"), (g[2], gz, "grad z"), (r, ro, "residual")]:
        assert _bitwise(a, b), "%s %s: %s" % (name, what, _describe(a, b))


def test_flux_deterministic(pair):
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 2)
    r1, _ = _flux_gpu(m, q)
    r2, _ = _flux_gpu(m, q)
    assert _bitwise(r1, r2)


def test_flux_first_order(pair):
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 3)
    m.set_options(K.options(first_order=1))
    try:
        r, g = _flux_gpu(m, q)
    finally:
        m.set_options(K.options())
    o1 = O.Oracle(m, first_order=1)
    ro, gx, _, _ = o1.flux(q)
    o1.close()
    assert not np.any(g[0])
    assert _bitwise(r, ro), _describe(r, ro)


def test_timestep_fill_forces_bitwise(pair):
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 4)
    m.set_state(q)
    orc.set_state(q)
    m.timestep(73.0)
    cdt = m.cdt()
    cdt_o = orc.timestep(73.0)
    assert _bitwise(cdt, cdt_o), _describe(cdt, cdt_o)
    m.fill()
    orc.fill()
    A, Ao = m.blocks().ravel(), orc.blocks().ravel()
    assert _bitwise(A, Ao), _describe(A, Ao)
    f, fo = m.forces(), orc.forces()
    assert _bitwise(f, fo), (f, fo)


def test_ilu_and_sptrsv(pair):
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 5)
    m.set_state(q)
    m.timestep(50.0)
    m.fill()
    m.ilu()
    assert m.last_error() == 0
    F = m.blocks().ravel()
    # textbook ILU(0) of the oracle: bit-identical
    ot = O.Oracle(m, ilu_textbook=1)
    ot.set_state(q)
    ot.timestep(50.0)
    ot.fill()
    ot.ilu()
    Ft = ot.blocks().ravel()
    assert _bitwise(F, Ft), _describe(F, Ft)
    # the reference's sweep (one thread, with its copy side effects): 1e-10
    orc.set_state(q)
    orc.timestep(50.0)
    orc.fill()
    orc.ilu()
    Fe = orc.blocks().ravel()
    scale = np.abs(Fe).max()
    assert np.max(np.abs(F - Fe)) <= 1e-10 * scale
    # triangular solve with identical factors: bit-identical
    b = np.random.default_rng(6).standard_normal(m.sz)
    bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
    m.sptrsv(bd, xd)
    x = xd.numpy()
    xt = ot.sptrsv(b)
    ot.close()
    assert _bitwise(x, xt), _describe(x, xt)
    # repeated solves reuse the completion epochs: same bits again
    m.sptrsv(bd, xd)
    assert _bitwise(xd.numpy(), xt)


def test_ijacvec(pair):
    """The FD product divides R(w) - R(q) by h ~ 1e-8 |q|: the tree-ordered
    |x| changes h/|x| in the last bit, which can move single entries of w by
    one ulp of q.  Bar: 1e-7 relative in the 2-norm."""
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 7)
    m.set_state(q)
    orc.set_state(q)
    m.timestep(50.0)
    orc.timestep(50.0)
    r, _ = _flux_gpu(m, q)
    x = np.random.default_rng(8).standard_normal(m.sz)
    qnorm = np.sqrt(1.0 + O.norm(q))
    xd, rd, wd, yd = K.DeviceVec(m.sz, x), K.DeviceVec(m.sz, r), K.DeviceVec(m.sz), K.DeviceVec(m.sz)
    g = [K.DeviceVec(m.sz) for _ in range(3)]
    m.ijacvec(xd, rd, qnorm, g, wd, yd)
    w, y = wd.numpy(), yd.numpy()
    wo, yo = orc.ijacvec(x, r, qnorm)
    assert np.max(np.abs(w - wo)) <= 4 * np.finfo(float).eps * np.abs(wo).max()
    assert np.linalg.norm(y - yo) <= 1e-7 * np.linalg.norm(yo)


def test_kernel_history_mesh_a(mesh_a):
    """Full reference-profile solve (alpha 3, CFL 50, 15 steps, 1e7 drop):
    same Krylov counts per step, residual norms and forces within 1e-8."""
    m = K.Mesh(mesh_a, K.options())
    orc = O.Oracle(m)
    ho = orc.kernel()
    hg = m.kernel()
    orc.close()
    m.close()
    assert len(hg) == len(ho)
    for a, b in zip(hg, ho):
        assert a["krylov"] == b["krylov"], (a, b)
        assert abs(a["fnorm"] - b["fnorm"]) <= 1e-8 * abs(b["fnorm"]), (a, b)
        for fa, fb in zip(a["forces"], b["forces"]):
            assert abs(fa - fb) <= 1e-8 * max(abs(fb), 1e-12), (a, b)
