"""The host's parallel incidence builder (kf_incidence: atomic counts and
slots, then each vertex's short list sorted) against a stable numpy
construction of the same lists: for every vertex, its incident edges in
ascending edge id, each entry edge << 1 | side (side 1: the vertex is the
edge's n1).  These lists fix the order of every per-vertex sum of the
residual, the fill and the LSQ weights, so they must not depend on thread
timing: the builder runs several times with 8 threads and must agree each
time."""
import ctypes as C

import numpy as np

import kfun3d_amd as K
from conftest import lattice


def _incidence(N, n0, n1):
    f = K.lib().kf_incidence
    f.restype = None
    f.argtypes = [C.c_size_t, C.c_size_t, C.c_void_p, C.c_void_p, C.POINTER(C.POINTER(C.c_uint)),
                  C.POINTER(C.POINTER(C.c_uint))]
    vp, ve = C.POINTER(C.c_uint)(), C.POINTER(C.c_uint)()
    n0 = np.ascontiguousarray(n0, dtype=np.uint32)
    n1 = np.ascontiguousarray(n1, dtype=np.uint32)
    f(N, len(n0), n0.ctypes.data, n1.ctypes.data, C.byref(vp), C.byref(ve))
    ptr = np.ctypeslib.as_array(vp, shape=(N + 1,)).copy()
    ent = np.ctypeslib.as_array(ve, shape=(2 * len(n0),)).copy()
    libc = C.CDLL(None)
    libc.free(C.cast(vp, C.c_void_p))
    libc.free(C.cast(ve, C.c_void_p))
    return ptr, ent


def _expected(N, n0, n1):
    E = len(n0)
    owner = np.empty(2 * E, dtype=np.int64)
    owner[0::2], owner[1::2] = n0, n1
    ent = np.empty(2 * E, dtype=np.uint64)
    ent[0::2] = np.arange(E, dtype=np.uint64) << 1
    ent[1::2] = (np.arange(E, dtype=np.uint64) << 1) | 1
    order = np.argsort(owner, kind="stable")          # by vertex, then edge order
    ptr = np.zeros(N + 1, dtype=np.int64)
    np.add.at(ptr, owner + 1, 1)
    return np.cumsum(ptr), ent[order]


def test_incidence_matches_stable_construction(mesh_dir, mesh_a):
    for path in (mesh_a, lattice(mesh_dir, 24, 10, 9, 5)):
        m = K.Mesh(path, device=False)
        h = m.host_arrays()
        N = m.N
        m.close()
        want_p, want_e = _expected(N, h["e_n0"], h["e_n1"])
        for _ in range(4):
            p, e = _incidence(N, h["e_n0"], h["e_n1"])
            assert np.array_equal(p.astype(np.int64), want_p)
            assert np.array_equal(e.astype(np.uint64), want_e)


def test_incidence_on_unsorted_edges():
    """edges in no particular order, n0 > n1 allowed: the lists still ascend
    in edge id"""
    rng = np.random.default_rng(9)
    N, E = 500, 4000
    n0 = rng.integers(0, N, E)
    n1 = (n0 + 1 + rng.integers(0, N - 1, E)) % N
    want_p, want_e = _expected(N, n0, n1)
    p, e = _incidence(N, n0, n1)
    assert np.array_equal(p.astype(np.int64), want_p)
    assert np.array_equal(e.astype(np.uint64), want_e)
