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


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_gradient_both_lsq_forms(which, mesh_a, mesh_c1, monkeypatch):
    """The gradient forms each incidence's LSQ terms from the end's seven
    moments once setup has checked them against the GEOMETRY's per-edge terms
    (kf_grad_moments = 1 on these meshes), or reads the per-edge terms
    (KF_GRAD_MOMENTS=0); both give the oracle's gradient and residual bit for
    bit."""
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    hm = K.Mesh(path, device=False)
    q = perturbed_state(hm.N, 3.0, 21)
    hm.close()
    out = {}
    for form in (1, 0):
        if form == 0:
            monkeypatch.setenv("KF_GRAD_MOMENTS", "0")
        m = K.Mesh(path, K.options())
        try:
            assert K.lib().kf_grad_moments(m.ref) == form
            out[form] = _flux_gpu(m, q)
            if form == 1:
                orc = O.Oracle(m)
                ro, gx, gy, gz = orc.flux(q)
                orc.close()
        finally:
            m.close()
    for form in (1, 0):
        r, g = out[form]
        for a, b, what in [(g[0], gx, "grad x"), (g[1], gy, "grad y"), (g[2], gz, "grad z"), (r, ro, "residual")]:
            assert _bitwise(a, b), "%s form %d %s: %s" % (which, form, what, _describe(a, b))


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


def test_timestep_folded_into_residual(pair):
    """The driver's residual that also leaves the time-step sums: the same
    residual, and cdt bit-identical to ComputeTimeStep on that state."""
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 31)
    m.set_state(q)
    m.timestep(37.0)
    cdt = m.cdt()
    qd = K.DeviceVec(m.sz, q)
    g = [K.DeviceVec(m.sz) for _ in range(3)]
    r = K.DeviceVec(m.sz)
    m.flux_spectral(qd, g, r)
    m.timestep_spectral(37.0)
    assert _bitwise(m.cdt(), cdt), _describe(m.cdt(), cdt)
    r0, _ = _flux_gpu(m, q)
    assert _bitwise(r.numpy(), r0)


def test_ilu_and_sptrsv(pair):
    """Default FP64 ILU = the reference's one-thread sweep (ilu_Edmond) bit for
    bit, on a first call and on a second call that starts from the L blocks
    the first one left behind; the triangular solve with those factors too."""
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 5)
    q2 = perturbed_state(m.N, 3.0, 55, amp=0.01) + q - perturbed_state(m.N, 3.0, 55, amp=0.0)
    b = np.random.default_rng(6).standard_normal(m.sz)
    bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
    fresh = O.Oracle(m)
    for state, cfl in ((q, 50.0), (q2, 500.0)):
        m.set_state(state)
        m.timestep(cfl)
        m.fill()
        m.ilu()
        assert m.last_error() == 0
        F = m.blocks().ravel()
        fresh.set_state(state)
        fresh.timestep(cfl)
        fresh.fill()
        fresh.ilu()
        Fe = fresh.blocks().ravel()
        assert _bitwise(F, Fe), "%s: %s" % (name, _describe(F, Fe))
        m.sptrsv(bd, xd)
        x, xe = xd.numpy(), fresh.sptrsv(b)
        assert _bitwise(x, xe), _describe(x, xe)
        m.sptrsv(bd, xd)    # repeated solves reuse the completion epochs
        assert _bitwise(xd.numpy(), xe)
    fresh.close()


@pytest.mark.parametrize("rows", [1, 2, 8, 16])
def test_sptrsv_other_schedules(mesh_a, rows, monkeypatch):
    """The triangular solve is bit-identical to the oracle for every range
    size of its schedule (the default is 4 rows per range)."""
    monkeypatch.setenv("KF_TRSV_ROWS", str(rows))
    m = K.Mesh(mesh_a, K.options())
    try:
        q = perturbed_state(m.N, 3.0, 15)
        b = np.random.default_rng(16).standard_normal(m.sz)
        m.set_state(q)
        m.timestep(50.0)
        m.fill()
        m.ilu()
        bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
        m.sptrsv(bd, xd)
        assert m.last_error() == 0
        orc = O.Oracle(m)
        orc.set_state(q)
        orc.timestep(50.0)
        orc.fill()
        orc.ilu()
        xe = orc.sptrsv(b)
        orc.close()
        assert _bitwise(xd.numpy(), xe), _describe(xd.numpy(), xe)
    finally:
        m.close()


@pytest.mark.parametrize("rows", [4, 16])
def test_sptrsv_held_entries(mesh_c1, rows, monkeypatch):
    """On a lattice (at most 7 entries on either side of the diagonal) the
    sweeps hold 7 entries per row in registers; with 8 forced they give the
    same bits, and both match the oracle."""
    monkeypatch.setenv("KF_TRSV_ROWS", str(rows))
    q = perturbed_state(12288, 3.0, 21)
    b = np.random.default_rng(22).standard_normal(4 * 12288)
    out = []
    for held in (None, "8"):
        if held:
            monkeypatch.setenv("KF_TRSV_HELD", held)
        m = K.Mesh(mesh_c1, K.options())
        try:
            assert m.N == 12288
            m.set_state(q)
            m.timestep(50.0)
            m.fill()
            m.ilu()
            bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
            m.sptrsv(bd, xd)
            m.sptrsv(bd, xd)
            assert m.last_error() == 0
            out.append(xd.numpy())
            if held is None:
                orc = O.Oracle(m)
                orc.set_state(q)
                orc.timestep(50.0)
                orc.fill()
                orc.ilu()
                xe = orc.sptrsv(b)
                orc.close()
        finally:
            m.close()
    assert _bitwise(out[0], xe), _describe(out[0], xe)
    assert _bitwise(out[1], out[0]), _describe(out[1], out[0])


def test_ilu_textbook(pair):
    """ilu_exact=0: textbook ILU(0), bit for bit against the oracle's textbook
    sweep, and within 1e-10 of the reference's sweep."""
    name, m, orc = pair
    q = perturbed_state(m.N, 3.0, 9)
    m.set_options(K.options(ilu_exact=0))
    try:
        m.set_state(q)
        m.timestep(50.0)
        m.fill()
        m.ilu()
        F = m.blocks().ravel()
    finally:
        m.set_options(K.options())
    ot = O.Oracle(m, ilu_textbook=1)
    ot.set_state(q)
    ot.timestep(50.0)
    ot.fill()
    ot.ilu()
    Ft = ot.blocks().ravel()
    ot.close()
    assert _bitwise(F, Ft), _describe(F, Ft)
    oe = O.Oracle(m)
    oe.set_state(q)
    oe.timestep(50.0)
    oe.fill()
    oe.ilu()
    Fe = oe.blocks().ravel()
    oe.close()
    assert np.max(np.abs(F - Fe)) <= 1e-10 * np.abs(Fe).max()


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


def _same_history(hg, ho):
    assert len(hg) == len(ho), (len(hg), len(ho))
    for a, b in zip(hg, ho):
        assert a["krylov"] == b["krylov"], (a, b)
        assert a["cfl"] == b["cfl"], (a, b)
        assert a["fnorm"] == b["fnorm"], (a, b)
        assert a["forces"] == b["forces"], (a, b)


@pytest.mark.parametrize("which", ["mesh_a", "c1"])
def test_kernel_reproduces_reference_history(which, mesh_a, mesh_c1):
    """Reference profile (alpha 3, CFL 50, 15 steps, 1e7 drop) with norms and
    dot products summed in index order: every step's Krylov count, CFL,
    residual norm and forces equal the reference's own run (golden fixture
    from the reference built from its sources) bit for bit."""
    import json, os
    from conftest import GOLDEN
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[which]
    ref = json.load(open(os.path.join(GOLDEN, "ref_%s.json" % which)))["solve"]
    m = K.Mesh(path, K.options(reduce_seq=1))
    hg = m.kernel()
    m.close()
    _same_history(hg, ref)


@pytest.mark.parametrize("which", ["mesh_a", "c1"])
def test_kernel_tree_reductions_match_oracle(which, mesh_a, mesh_c1):
    """Default (fast) reductions: the whole history equals, bit for bit, the
    oracle run with the device's reduction tree -- every other operation of
    the solve is the reference's."""
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[which]
    m = K.Mesh(path, K.options())
    orc = O.Oracle(m, reduce_order=1)
    ho = orc.kernel()
    orc.close()
    hg = m.kernel()
    m.close()
    _same_history(hg, ho)


@pytest.mark.parametrize("which", ["mesh_a", "c1"])
def test_kernel_default_history_close_to_reference(which, mesh_a, mesh_c1):
    """The default (benchmarked) path -- tree-ordered norms and dot products,
    resident orthogonalisation -- held directly to the reference's own
    one-thread history (golden fixture): the same pseudo-time steps and the
    same Krylov count at every step, per-step residual norms within the
    measured bounds (test_oracle_reduction.TREE_BARS: 3e-3 relative, 2e-6 of
    the initial norm), forces within 3e-7 at every step and 1e-8 at the
    end.  The reference's order sums the reductions sequentially; the
    differences are those ulps, amplified by Newton-Krylov."""
    import json, os
    from conftest import GOLDEN
    from test_oracle_reduction import check_tree_history
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[which]
    ref = json.load(open(os.path.join(GOLDEN, "ref_%s.json" % which)))["solve"]
    m = K.Mesh(path, K.options())
    hg = m.kernel()
    assert m.last_error() == 0
    m.close()
    d = check_tree_history(hg, ref)
    print("%s: default history vs reference: %r" % (which, d))


def test_norm_tree_any_length_and_alignment(mesh_a):
    """Compute2ndNorm in the fixed tree (element pairs per thread, 16-byte
    loads when the vector is 16-byte aligned) against the oracle's statement
    of the same tree, bit for bit: odd and even lengths, and pointers 8 bytes
    off a 16-byte boundary (the two-load path)."""
    m = K.Mesh(mesh_a, K.options(reduce_seq=0))
    try:
        rng = np.random.default_rng(21)
        for n in (1, 2, 7, 1000, 4 * m.N - 1, 4 * m.N, 600001):
            x = rng.standard_normal(n + 1)
            dv = K.DeviceVec(n + 1, x)
            got_a = K.lib().Compute2ndNorm(n, dv.ptr)
            got_u = K.lib().Compute2ndNorm(n, dv.ptr + 8)
            want_a = np.sqrt(O.dot_tree(x[:n], x[:n]))
            want_u = np.sqrt(O.dot_tree(x[1:], x[1:]))
            assert m.last_error() == 0
            assert _bitwise(np.array([got_a]), np.array([want_a])), (n, got_a, want_a)
            assert _bitwise(np.array([got_u]), np.array([want_u])), (n, got_u, want_u)
            dv.free()
    finally:
        m.close()
