"""The host-side schedule of the level-scheduled triangular solves
(kfun3d_amd/host/plan.c), checked on the CPU for every range size: each
range appears in exactly one task, every slot holds a valid range or the
empty marker, the ranges of a task never read one another, every range a
task reads from sits in an earlier task (so the ticket order is a
topological order and no wait can deadlock), and the local levels and
depths are those of the in-range dependence chains."""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O


def _lower_upper(m):
    orc = O.Oracle(m)
    rp, col = orc.pattern()
    orc.close()
    return rp.astype(np.int64), col.astype(np.int64)


def _check(m, rp, col, upper, R):
    N = m.N
    task, tdep, lvl, rowtab = m.trsv_plan(upper, R)
    S = 16 // R
    ng = (N + R - 1) // R
    flat = task.ravel()
    valid = flat != 0xFFFFFFFF
    assert np.all(flat[valid] < ng)
    # every range exactly once
    assert np.array_equal(np.sort(flat[valid]), np.arange(ng))
    # no task is empty
    assert np.all(valid.reshape(-1, S).any(axis=1))
    # position of each range in the ticket order
    where = np.empty(ng, dtype=np.int64)
    where[flat[valid]] = np.nonzero(valid)[0] // S
    # range and in-range position of every row
    rows = np.arange(N)
    if upper:
        rng = (N - 1 - rows) // R
    else:
        rng = rows // R
    # dependences: forward row r reads the columns left of its diagonal,
    # backward row r the columns right of it
    depth = np.zeros(ng, dtype=np.int64)
    ll = np.zeros(N, dtype=np.int64)
    order = rows[::-1] if upper else rows
    for r in order:
        cols = col[rp[r]:rp[r + 1]]
        cols = cols[cols > r] if upper else cols[cols < r]
        inside = cols[rng[cols] == rng[r]]
        outside = cols[rng[cols] != rng[r]]
        ll[r] = (ll[inside].max() + 1) if len(inside) else 0
        depth[rng[r]] = max(depth[rng[r]], ll[r] + 1)
        if len(outside):
            assert np.all(where[rng[outside]] < where[rng[r]]), (r, outside)
    assert np.array_equal(lvl.astype(np.int64), ll)
    tmax = np.zeros(len(task), dtype=np.int64)
    np.maximum.at(tmax, np.nonzero(valid)[0] // S, depth[flat[valid]])
    assert np.array_equal(tdep.astype(np.int64), tmax)
    # per-task row table: slot q holds row q % R of range task[t, q // R]
    diag = np.array([rp[r] + np.nonzero(col[rp[r]:rp[r + 1]] == r)[0][0] for r in range(N)])
    for t in range(len(task)):
        for q in range(16):
            g = int(task[t, q // R]); kr = q % R
            r, ja, jb = (int(v) for v in rowtab[t, q])
            if g == 0xFFFFFFFF or g * R + kr >= N:
                assert (r, ja, jb) == (0xFFFFFFFF, 0, 0)
                continue
            assert r == (N - 1 - (g * R + kr) if upper else g * R + kr)
            assert (ja, jb) == ((diag[r], rp[r + 1]) if upper else (rp[r], diag[r]))
    return len(task)


@pytest.mark.parametrize("R", [1, 2, 4, 8, 16])
def test_plan_mesh_a(mesh_a, R):
    m = K.Mesh(mesh_a, device=False)
    rp, col = _lower_upper(m)
    for upper in (0, 1):
        _check(m, rp, col, upper, R)
    m.close()


@pytest.mark.parametrize("R", [4, 16])
def test_plan_lattice(mesh_c1, R):
    m = K.Mesh(mesh_c1, device=False)
    rp, col = _lower_upper(m)
    for upper in (0, 1):
        nt = _check(m, rp, col, upper, R)
        # 48x16x16 lattice: the range DAG has (48 / R) + 15 + 15 levels
        assert nt >= (48 // R + 30) * 1
    m.close()


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_ilu_plan(which, mesh_a, mesh_c1):
    """The ILU(0) schedule: rows in level order (a row's lower columns come
    at earlier tickets), and for every lower block (i, m) exactly the pairs
    (U(m,k), position of k in row i) a merge of rows m and i finds, in
    ascending order of the block of row m."""
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    m = K.Mesh(path, device=False)
    rp, col = _lower_upper(m)
    N = m.N
    diag = np.array([rp[r] + np.nonzero(col[rp[r]:rp[r + 1]] == r)[0][0] for r in range(N)])
    rec, ldiag, moff, mjm, mlo = m.ilu_plan(len(col))
    m.close()
    rows = rec[:, 0].astype(np.int64)
    assert np.array_equal(np.sort(rows), np.arange(N))
    pos = np.empty(N, dtype=np.int64)
    pos[rows] = np.arange(N)
    nxt = 0
    for t in range(N):
        i, j0, L, nlow, mbeg, mcnt = (int(v) for v in rec[t, :6])
        assert (j0, L, nlow) == (rp[i], rp[i + 1] - rp[i], diag[i] - rp[i])
        assert mbeg == nxt
        cols_i = col[rp[i]:rp[i + 1]]
        k = 0
        for p in range(nlow):
            mm = col[j0 + p]
            assert pos[mm] < t                      # row m is factored before row i
            assert ldiag[j0 + p] == diag[mm]
            assert moff[j0 + p] == k
            for jm in range(diag[mm] + 1, rp[mm + 1]):
                hit = np.nonzero(cols_i == col[jm])[0]
                if len(hit):
                    assert mjm[mbeg + k] == jm and mlo[mbeg + k] == hit[0]
                    k += 1
        assert k == mcnt
        nxt = mbeg + mcnt
    assert nxt == len(mjm)
