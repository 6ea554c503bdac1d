"""The coarse level of the two-level preconditioner as the host builds it
(kfun3d_amd/host/aggregate.c), on the CPU:

  * on a Kuhn lattice with even sides, the greedy aggregation in vertex
    order gives exactly the 2 x 2 x 2 cubes of lattice points: vertex
    (i, j, k), v = i + nx (j + ny k), lies in cube (i // 2, j // 2, k // 2),
    and there are nx ny nz / 8 aggregates (the closed form of the greedy
    pass on this graph: a vertex whose i, j, k are all even is reached by no
    earlier root's ring, and its forward ring holds the rest of its cube);
  * on any mesh (mesh A, an odd lattice) every vertex has one aggregate,
    every aggregate is connected in the BSR pattern's graph, and the coarse
    pattern is exactly the set of aggregate pairs the fine pattern's blocks
    map to, rows sorted, with the diagonal present.
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import lattice


def _pattern(m):
    orc = O.Oracle(m)
    rp, col = orc.pattern()
    orc.close()
    return rp.astype(np.int64), col.astype(np.int64)


def _check_generic(m):
    rp, col = _pattern(m)
    agg, crp, ccol = m.coarse_plan()
    N, nc = m.N, len(crp) - 1
    agg = agg.astype(np.int64)
    assert agg.min() == 0 and agg.max() == nc - 1
    assert len(np.unique(agg)) == nc
    # each aggregate connected: a union-find over the fine edges inside aggregates
    parent = np.arange(N)

    def find(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    rows = np.repeat(np.arange(N), np.diff(rp))
    inside = agg[rows] == agg[col]
    for a, b in zip(rows[inside], col[inside]):
        ra, rb = find(a), find(b)
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    roots = np.array([find(v) for v in range(N)])
    # one component per aggregate
    comp_of_agg = {}
    for v in range(N):
        comp_of_agg.setdefault(agg[v], set()).add(roots[v])
    assert all(len(s) == 1 for s in comp_of_agg.values())
    # coarse pattern = the aggregate pairs of the fine blocks
    pairs = np.unique(agg[rows] * nc + agg[col])
    got = np.repeat(np.arange(nc), np.diff(crp.astype(np.int64))) * nc + ccol.astype(np.int64)
    assert np.array_equal(got, pairs)          # sorted rows, no duplicates, nothing else
    assert np.all(np.isin(np.arange(nc) * (nc + 1), got))
    return agg, nc


def test_lattice_cubes(mesh_c1):
    m = K.Mesh(mesh_c1, device=False)
    try:
        agg, nc = _check_generic(m)
        nx, ny, nz = 48, 16, 16
        assert m.N == nx * ny * nz
        v = np.arange(m.N)
        i, j, k = v % nx, (v // nx) % ny, v // (nx * ny)
        cube = (i // 2) + (nx // 2) * ((j // 2) + (ny // 2) * (k // 2))
        assert nc == (nx // 2) * (ny // 2) * (nz // 2)
        # the same partition: aggregate -> cube is a bijection
        key = np.unique(agg * (nc + 1) + cube)
        assert len(key) == nc
        assert len(np.unique(key // (nc + 1))) == nc and len(np.unique(key % (nc + 1))) == nc
    finally:
        m.close()


@pytest.mark.parametrize("which", ["mesh_a", "odd"])
def test_aggregates_connected_and_pattern(which, mesh_a, mesh_dir):
    path = mesh_a if which == "mesh_a" else lattice(mesh_dir, 9, 7, 5, 3)
    m = K.Mesh(path, device=False)
    try:
        agg, nc = _check_generic(m)
        assert nc < m.N
    finally:
        m.close()


def test_cpu_definitions_pinned(mesh_a):
    """The CPU pieces the GPU tests compare against: galerkin() equals the
    sparse product P^T A P, and block_ilu0() is the oracle's textbook
    ILU(0) (its factors' solve agrees to 1e-13) with L U = A on the
    pattern."""
    import scipy.sparse as sp
    from conftest import perturbed_state
    from coarse_ref import block_ilu0, block_ilu0_solve, galerkin

    m = K.Mesh(mesh_a, device=False)
    orc = O.Oracle(m, ilu_textbook=1)
    try:
        orc.set_state(perturbed_state(m.N, 3.0, 17))
        orc.timestep(30.0)
        orc.fill()
        rp, col = orc.pattern()
        rp, col = rp.astype(np.int64), col.astype(np.int64)
        A = orc.blocks().reshape(-1, 4, 4).copy()
        N = m.N
        Af = sp.bsr_matrix((A, col, rp), shape=(4 * N, 4 * N)).tocsr()
        agg, crp, ccol = (a.astype(np.int64) for a in m.coarse_plan())
        nc = len(crp) - 1
        P = sp.csr_matrix((np.ones(4 * N), 4 * np.repeat(agg, 4) + np.tile(np.arange(4), N), np.arange(4 * N + 1)),
                          shape=(4 * N, 4 * nc))
        Ac = galerkin(A, rp, col, agg, crp, ccol)
        G = (P.T @ Af @ P).toarray()
        Acd = sp.bsr_matrix((Ac, ccol, crp), shape=(4 * nc, 4 * nc)).toarray()
        assert np.abs(G - Acd).max() <= 1e-14 * np.abs(G).max()
        L, U = block_ilu0(rp, col, A)
        Ad = Af.toarray()
        on = Ad != 0
        assert np.abs((L @ U).toarray() - Ad)[on].max() <= 1e-13 * np.abs(Ad).max()
        orc.ilu()
        b = np.random.default_rng(1).standard_normal(4 * N)
        x, xr = block_ilu0_solve((L, U), b), orc.sptrsv(b)
        assert np.linalg.norm(x - xr) <= 1e-13 * np.linalg.norm(xr)
        # a planted mistake fails: the factors of the transposed blocks
        xt = block_ilu0_solve(block_ilu0(rp, col, A.transpose(0, 2, 1).copy()), b)
        assert np.linalg.norm(xt - xr) > 1e-2 * np.linalg.norm(xr)
    finally:
        orc.close()
        m.close()
