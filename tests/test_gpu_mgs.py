"""Modified Gram-Schmidt in one launch with w held in registers (an ordinary
launch whose workgroups check their co-residency themselves and publish a
ready flag to a pinned host mirror, which the host polls), against the
per-step kernels (src/kernel.c:129-141 and the closing norm):
every h_k, |w| and the orthogonalised w bit for bit, for vector lengths that
give every register-size variant, odd lengths (the half pair) and lengths
shorter than one pass of the reduction grid."""
import ctypes as C

import numpy as np
import pytest

import kfun3d_amd as K

pytestmark = pytest.mark.gpu

S = 1024 * 256   # element pairs one pass of the reduction grid walks
SLOT = 100       # device scalar slots used here (h_0 .. h_nv)


def _fns():
    L = K.lib()
    mgs = L.kfs_mgs
    mgs.restype = None
    mgs.argtypes = [K.GP, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_int]
    fetch = L.kfs_fetch
    fetch.restype = None
    fetch.argtypes = [K.GP, C.c_int, C.c_int, C.c_void_p]
    launches = L.kfs_mgs_launches
    launches.restype = C.c_ulonglong
    launches.argtypes = [K.GP]
    return mgs, fetch, launches


@pytest.fixture(scope="module")
def contexts(mesh_a):
    """Two contexts on mesh A: per-step kernels, and the resident launch."""
    import os
    old = os.environ.get("KF_MGS_RESIDENT")
    out = {}
    try:
        for name, val in (("step", "0"), ("resident", "2")):   # 2: at every size
            os.environ["KF_MGS_RESIDENT"] = val
            out[name] = K.Mesh(mesh_a, K.options())
    finally:
        if old is None:
            os.environ.pop("KF_MGS_RESIDENT", None)
        else:
            os.environ["KF_MGS_RESIDENT"] = old
    yield out
    for m in out.values():
        m.close()


CASES = [
    (11200, 5),                      # mesh A's 4N: one pair per thread at most, most threads idle
    (1327104, 30),                   # the 144x48x48 lattice's 4N, a full GMRES(30) cycle's last step
    (2 * (2 * S + 777) + 1, 3),      # odd length: the half pair
    (2 * 4 * S, 7),                  # exactly four pairs per thread
    (2 * (7 * S + 100), 4),          # eight pair slots
    (2 * (11 * S + 4) + 1, 2),       # twelve, odd
    (10616832, 6),                   # the 288x96x96 lattice's 4N: 21 pair slots
]


@pytest.mark.parametrize("n,nv", CASES)
def test_resident_mgs_bitwise(contexts, n, nv):
    mgs, fetch, launches = _fns()
    rng = np.random.default_rng(n % 1000 + nv)
    stride = n + (n & 1) + 2   # even, so every basis vector stays 16-byte aligned
    Q = np.linalg.qr(rng.standard_normal((n, nv)))[0]   # an orthonormal basis, as GMRES holds
    V = np.zeros((nv, stride))
    V[:, :n] = Q.T
    V = V.ravel()
    w0 = rng.standard_normal(n)
    vb = K.DeviceVec(nv * stride, V)
    res = {}
    for name in ("step", "resident"):
        m = contexts[name]
        before = launches(m.ref)
        w = K.DeviceVec(n, w0)
        mgs(m.ref, n, vb.ptr, stride, nv, w.ptr, SLOT)
        h = np.empty(nv + 1)
        fetch(m.ref, SLOT, nv + 1, h.ctypes.data)
        assert m.last_error() == 0
        res[name] = (h, w.numpy(), launches(m.ref) - before)
        w.free()
    vb.free()
    (hs, ws, ls), (hr, wr, lr) = res["step"], res["resident"]
    assert ls == 0 and lr == 1, "the resident launch did not run (%d, %d)" % (ls, lr)
    assert np.array_equal(hs.view(np.uint64), hr.view(np.uint64)), (hs, hr)
    assert np.array_equal(ws.view(np.uint64), wr.view(np.uint64))
    # and the result is an orthogonalisation: w is orthogonal to the basis to
    # rounding, and |w| is the stored norm
    Vm = V.reshape(nv, stride)[:, :n]
    assert np.max(np.abs(Vm @ wr)) < 1e-12 * np.linalg.norm(w0)
    assert abs(np.linalg.norm(wr) - hr[nv]) <= 1e-12 * hr[nv]


def test_resident_mgs_repeated(contexts):
    """Consecutive launches on one context use fresh sentinels: the same
    input gives the same bits every time."""
    mgs, fetch, launches = _fns()
    n, nv = 1327104, 4
    rng = np.random.default_rng(5)
    V = rng.standard_normal(nv * n)
    w0 = rng.standard_normal(n)
    vb = K.DeviceVec(nv * n, V)
    m = contexts["resident"]
    outs = []
    for _ in range(3):
        w = K.DeviceVec(n, w0)
        mgs(m.ref, n, vb.ptr, n, nv, w.ptr, SLOT)
        h = np.empty(nv + 1)
        fetch(m.ref, SLOT, nv + 1, h.ctypes.data)
        outs.append((h, w.numpy()))
        w.free()
    vb.free()
    assert m.last_error() == 0
    for h, w in outs[1:]:
        assert np.array_equal(h.view(np.uint64), outs[0][0].view(np.uint64))
        assert np.array_equal(w.view(np.uint64), outs[0][1].view(np.uint64))


def test_resident_mgs_default_sizes(mesh_a, monkeypatch):
    """By default the resident launch is taken wherever it fits: for the
    144x48x48 and the 288x96x96 lattices' vectors alike."""
    monkeypatch.delenv("KF_MGS_RESIDENT", raising=False)
    mgs, fetch, launches = _fns()
    m = K.Mesh(mesh_a, K.options())
    try:
        for n, taken in ((1327104, 1), (10616832, 1)):
            nv = 2
            rng = np.random.default_rng(7)
            vb = K.DeviceVec(nv * n, rng.standard_normal(nv * n))
            w = K.DeviceVec(n, rng.standard_normal(n))
            before = launches(m.ref)
            mgs(m.ref, n, vb.ptr, n, nv, w.ptr, SLOT)
            h = np.empty(nv + 1)
            fetch(m.ref, SLOT, nv + 1, h.ctypes.data)
            assert launches(m.ref) - before == taken, n
            w.free()
            vb.free()
        assert m.last_error() == 0
    finally:
        m.close()


@pytest.mark.parametrize("n,nv", [(1327104, 3), (2 * (2 * S + 777) + 1, 2), (10616832, 4)])
def test_resident_mgs_unit_bitwise(contexts, n, nv):
    """The GMRES form (kfs_mgs_unit: then w *= 1.0 / |w|): scaled inside the
    resident launch, by k_scale_recip after the per-step kernels; the same
    bits either way, and |w| = 1 to rounding."""
    L = K.lib()
    unit = L.kfs_mgs_unit
    unit.restype = C.c_int
    unit.argtypes = [K.GP, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                     C.c_double, C.c_void_p]
    _, fetch, launches = _fns()
    rng = np.random.default_rng(n % 997 + nv)
    stride = n + (n & 1) + 2
    Q = np.linalg.qr(rng.standard_normal((n, nv)))[0]
    V = np.zeros((nv, stride))
    V[:, :n] = Q.T
    vb = K.DeviceVec(nv * stride, V.ravel())
    w0 = rng.standard_normal(n)
    res = {}
    for name in ("step", "resident"):
        m = contexts[name]
        before = launches(m.ref)
        w = K.DeviceVec(n, w0)
        h = np.empty(nv + 1)
        unit(m.ref, n, vb.ptr, stride, nv, w.ptr, SLOT, h.ctypes.data, -1, 0.0, None)   # the values come back with the call
        assert m.last_error() == 0
        res[name] = (h, w.numpy(), launches(m.ref) - before)
        w.free()
    vb.free()
    (hs, ws, ls), (hr, wr, lr) = res["step"], res["resident"]
    assert ls == 0 and lr == 1
    assert np.array_equal(hs.view(np.uint64), hr.view(np.uint64))
    assert np.array_equal(ws.view(np.uint64), wr.view(np.uint64))
    assert abs(np.linalg.norm(wr) - 1.0) < 1e-12


@pytest.mark.parametrize("dims", [(48, 16, 16, 0), (144, 48, 48, 1)])
def test_solve_history_with_resident_mgs(mesh_dir, dims, monkeypatch):
    """Whole solves (the 48x16x16 and 144x48x48 lattices, to a 1e-10 drop)
    with the resident launch taken for every orthogonalisation have the
    history of the per-step kernels bit for bit: every step's CFL, residual
    norm, Krylov count and forces."""
    from conftest import lattice
    path = lattice(mesh_dir, *dims)
    hist = {}
    for val in ("0", "2"):
        monkeypatch.setenv("KF_MGS_RESIDENT", val)
        m = K.Mesh(path, K.config_options())
        try:
            hist[val] = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
    a, b = hist["0"], hist["2"]
    assert len(a) == len(b) and len(a) > 5
    for ra, rb in zip(a, b):
        assert ra["krylov"] == rb["krylov"], (ra, rb)
        fa = np.array([ra["cfl"], ra["fnorm"]] + ra["forces"])
        fb = np.array([rb["cfl"], rb["fnorm"]] + rb["forces"])
        assert np.array_equal(fa.view(np.uint64), fb.view(np.uint64)), (ra, rb)


def test_resident_mgs_abort_falls_back(mesh_a, monkeypatch):
    """A resident launch whose workgroups are not all co-resident aborts
    before writing anything (forced here from the host, KF_MGS_RESIDENT=3);
    the call is then repeated with the per-step kernels -- the same bits as
    a context that never tried -- and the context keeps those kernels."""
    L = K.lib()
    unit = L.kfs_mgs_unit
    unit.restype = C.c_int
    unit.argtypes = [K.GP, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                     C.c_double, C.c_void_p]
    aborts = L.kfs_mgs_aborts
    aborts.restype = C.c_ulonglong
    aborts.argtypes = [K.GP]
    _, fetch, launches = _fns()
    n, nv = 1327104, 5
    rng = np.random.default_rng(11)
    Q = np.linalg.qr(rng.standard_normal((n, nv)))[0]
    vb = K.DeviceVec(nv * n, np.ascontiguousarray(Q.T).ravel())
    w0 = rng.standard_normal(n)
    res = {}
    for val in ("0", "3"):
        monkeypatch.setenv("KF_MGS_RESIDENT", val)
        m = K.Mesh(mesh_a, K.options())
        try:
            out = []
            for _ in range(2):   # the second call goes straight to the per-step kernels
                w = K.DeviceVec(n, w0)
                h = np.empty(nv + 1)
                unit(m.ref, n, vb.ptr, n, nv, w.ptr, SLOT, h.ctypes.data, -1, 0.0, None)
                out.append((h, w.numpy()))
                w.free()
            assert m.last_error() == 0
            res[val] = (out, launches(m.ref), aborts(m.ref))
        finally:
            m.close()
    vb.free()
    (o0, l0, a0), (o3, l3, a3) = res["0"], res["3"]
    assert (l0, a0) == (0, 0) and (l3, a3) == (1, 1), (l0, a0, l3, a3)
    for (h0, w0_), (h3, w3) in zip(o0, o3):
        assert np.array_equal(h0.view(np.uint64), h3.view(np.uint64))
        assert np.array_equal(w0_.view(np.uint64), w3.view(np.uint64))


@pytest.mark.parametrize("n", [1327104, 2 * (2 * S + 777) + 1])
def test_resident_mgs_leaves_next_norm(contexts, n):
    """With xslot, the resident launch also leaves |v| of the stored unit
    vector v in scal[xslot] -- the bits kfs_nrm2 forms for v afterwards --
    so that the next iJacVec need not compute it."""
    L = K.lib()
    unit = L.kfs_mgs_unit
    unit.restype = C.c_int
    unit.argtypes = [K.GP, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                     C.c_double, C.c_void_p]
    nrm2 = L.kfs_nrm2
    nrm2.restype = None
    nrm2.argtypes = [K.GP, C.c_size_t, C.c_void_p, C.c_int]
    _, fetch, _ = _fns()
    nv, XS = 3, SLOT + 60
    rng = np.random.default_rng(n % 991)
    stride = n + (n & 1) + 2
    Q = np.linalg.qr(rng.standard_normal((n, nv)))[0]
    V = np.zeros((nv, stride))
    V[:, :n] = Q.T
    vb = K.DeviceVec(nv * stride, V.ravel())
    m = contexts["resident"]
    w = K.DeviceVec(n, rng.standard_normal(n))
    h = np.empty(nv + 1)
    assert unit(m.ref, n, vb.ptr, stride, nv, w.ptr, SLOT, h.ctypes.data, XS, 0.0, None) == 1
    left = np.empty(1)
    fetch(m.ref, XS, 1, left.ctypes.data)
    nrm2(m.ref, n, w.ptr, XS)
    again = np.empty(1)
    fetch(m.ref, XS, 1, again.ctypes.data)
    assert m.last_error() == 0
    assert left.view(np.uint64)[0] == again.view(np.uint64)[0], (left, again)
    assert abs(again[0] - 1.0) < 1e-14
    w.free()
    vb.free()


def test_resident_mgs_leaves_perturbed_state(contexts):
    """With xslot and wpert (vectors of the state's length), the launch also
    writes the next iJacVec's perturbed state (h / |v|) v + q -- the bits
    kfs_fd_perturb forms from v, |v| and the context's state q."""
    L = K.lib()
    unit = L.kfs_mgs_unit
    unit.restype = C.c_int
    unit.argtypes = [K.GP, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                     C.c_double, C.c_void_p]
    perturb = L.kfs_fd_perturb
    perturb.restype = None
    perturb.argtypes = [K.GP, C.c_void_p, C.c_double, C.c_int, C.c_void_p]
    m = contexts["resident"]
    n, nv, XS = m.sz, 4, SLOT + 60
    fd_h = 1.490116119384766e-08 * 3.7
    rng = np.random.default_rng(29)
    Q = np.linalg.qr(rng.standard_normal((n, nv)))[0]
    vb = K.DeviceVec(nv * n, np.ascontiguousarray(Q.T).ravel())
    w = K.DeviceVec(n, rng.standard_normal(n))
    wp = K.DeviceVec(n)
    h = np.empty(nv + 1)
    assert unit(m.ref, n, vb.ptr, n, nv, w.ptr, SLOT, h.ctypes.data, XS, fd_h, wp.ptr) == 2
    ref = K.DeviceVec(n)
    perturb(m.ref, w.ptr, fd_h, XS, ref.ptr)
    a, b = wp.numpy(), ref.numpy()
    assert m.last_error() == 0
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    for x in (vb, w, wp, ref):
        x.free()


def _same_history(a, b):
    assert len(a) == len(b) and len(a) > 5
    for ra, rb in zip(a, b):
        assert ra["krylov"] == rb["krylov"], (ra, rb)
        fa = np.array([ra["cfl"], ra["fnorm"]] + ra["forces"])
        fb = np.array([rb["cfl"], rb["fnorm"]] + rb["forces"])
        assert np.array_equal(fa.view(np.uint64), fb.view(np.uint64)), (ra, rb)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("which", ["mesh_a", "c1"])
def test_solve_history_with_folded_combination(which, prec, mesh_a, mesh_c1, monkeypatch):
    """GMRES's finite-difference combination formed by the forward sweep as
    it loads its right-hand side (KF_TRSV_FOLD=1; by default only where the
    sweeps are throughput-bound, which includes the single-precision
    preconditioners on the largest lattice) gives the history of the
    separate combination kernel bit for bit, in every precision."""
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[which]
    hist = {}
    for val in ("0", "1"):
        monkeypatch.setenv("KF_TRSV_FOLD", val)
        m = K.Mesh(path, K.options(prec=prec))   # the reference profile: 15 steps
        try:
            hist[val] = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
    _same_history(hist["0"], hist["1"])


@pytest.mark.parametrize("which", ["mesh_a", "c1"])
def test_solve_history_without_flag_polling(which, mesh_a, mesh_c1, monkeypatch):
    """The host waits for the resident orthogonalisation by polling its
    ready flag in the pinned host mirror; KF_MGS_NOPOLL=1 forces the
    fallback (synchronise the stream), which must give the same history bit
    for bit."""
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[which]
    hist = {}
    for val in ("0", "1"):
        monkeypatch.setenv("KF_MGS_NOPOLL", val)
        m = K.Mesh(path, K.options())
        try:
            hist[val] = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
    _same_history(hist["0"], hist["1"])
