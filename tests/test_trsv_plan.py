"""The host-side schedule of the level-scheduled triangular solves
(kfun3d_amd/host/plan.c), checked on the CPU for every range size: each
range appears in exactly one task, every slot holds a valid range or the
empty marker, a row reads only rows of its own task or of tasks at earlier
tickets (so the ticket order is a topological order and no wait can
deadlock), the local levels and depths are those of the dependence chains
inside each task, and the per-position slot table names the slot of every
same-task column."""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O


def _lower_upper(m):
    orc = O.Oracle(m)
    rp, col = orc.pattern()
    orc.close()
    return rp.astype(np.int64), col.astype(np.int64)


def _check(m, rp, col, upper, R, cluster=True):
    N = m.N
    task, tdep, lvl, rowtab, eslot = m.trsv_plan(upper, R, cluster)
    S = 16 // R
    ng = (N + R - 1) // R
    flat = task.ravel()
    valid = flat != 0xFFFFFFFF
    assert np.all(flat[valid] < ng)
    # every range exactly once
    assert np.array_equal(np.sort(flat[valid]), np.arange(ng))
    # no task is empty, members fill the leading slots
    v2 = valid.reshape(-1, S)
    assert np.all(v2[:, 0])
    assert np.all(v2[:, 1:] <= v2[:, :-1])
    # ticket and member position of each range
    where = np.empty(ng, dtype=np.int64)
    where[flat[valid]] = np.nonzero(valid)[0] // S
    member = np.empty(ng, dtype=np.int64)
    member[flat[valid]] = np.nonzero(valid)[0] % S
    rows = np.arange(N)
    pos = (N - 1 - rows) if upper else rows
    rng = pos // R
    rtask = where[rng]
    slot = member[rng] * R + pos % R
    # dependences: forward row r reads the columns left of its diagonal,
    # backward row r the columns right of it; rows are visited in sweep
    # order, so a same-task column's level is known before the row's
    depth = np.zeros(len(task), dtype=np.int64)
    ll = np.zeros(N, dtype=np.int64)
    order = rows[::-1] if upper else rows
    for r in order:
        js = np.arange(rp[r], rp[r + 1])
        cols = col[js]
        keep = (cols > r) if upper else (cols < r)
        js, cols = js[keep], cols[keep]
        same = rtask[cols] == rtask[r]
        inside, outside = cols[same], cols[~same]
        ll[r] = (ll[inside].max() + 1) if len(inside) else 0
        depth[rtask[r]] = max(depth[rtask[r]], ll[r] + 1)
        if len(outside):
            assert np.all(rtask[outside] < rtask[r]), (r, outside)
        assert np.array_equal(eslot[js[same]].astype(np.int64), slot[inside])
        assert np.all(eslot[js[~same]] == 0xFF)
    assert np.all(ll < 255)
    assert np.array_equal(lvl.astype(np.int64), ll)
    assert np.array_equal(tdep.astype(np.int64), depth)
    # per-task row table: slot q holds row q % R of range task[t, q // R]
    diag = np.array([rp[r] + np.nonzero(col[rp[r]:rp[r + 1]] == r)[0][0] for r in range(N)])
    for t in range(len(task)):
        for q in range(16):
            g = int(task[t, q // R]) if q // R < S else 0xFFFFFFFF
            kr = q % R
            r, ja, jb = (int(v) for v in rowtab[t, q])
            if g == 0xFFFFFFFF or g * R + kr >= N:
                assert (r, ja, jb) == (0xFFFFFFFF, 0, 0)
                continue
            assert r == (N - 1 - (g * R + kr) if upper else g * R + kr)
            assert (ja, jb) == ((diag[r], rp[r + 1]) if upper else (rp[r], diag[r]))
    return len(task), int(depth.max())


@pytest.mark.parametrize("R", [1, 2, 4, 8, 16])
def test_plan_mesh_a(mesh_a, R):
    m = K.Mesh(mesh_a, device=False)
    rp, col = _lower_upper(m)
    for upper in (0, 1):
        for cluster in (True, False):
            _check(m, rp, col, upper, R, cluster)
    m.close()


def test_plan_one_level_tasks(mesh_c1):
    """cluster = 0: the ranges of a task never read one another, and a
    task's ranges are consecutive ranges of one level of the range DAG."""
    m = K.Mesh(mesh_c1, device=False)
    rp, col = _lower_upper(m)
    R = 4
    for upper in (0, 1):
        _check(m, rp, col, upper, R, cluster=False)
        task, tdep, lvl, rowtab, eslot = m.trsv_plan(upper, R, False)
        N = m.N
        rows = np.arange(N)
        pos = (N - 1 - rows) if upper else rows
        rng = pos // R
        for r in range(N):
            cs = col[rp[r]:rp[r + 1]]
            cs = cs[cs > r] if upper else cs[cs < r]
            js = np.arange(rp[r], rp[r + 1])[np.isin(col[rp[r]:rp[r + 1]], cs)]
            inside = rng[cs] == rng[r]
            # only columns of the row's own range are same-task
            assert np.all((eslot[js] != 0xFF) == inside)
        assert tdep.max() <= R
    m.close()


@pytest.mark.parametrize("R", [4, 16])
def test_plan_lattice(mesh_c1, R):
    m = K.Mesh(mesh_c1, device=False)
    rp, col = _lower_upper(m)
    for upper in (0, 1):
        nt, dmax = _check(m, rp, col, upper, R)
        # 48x16x16 lattice: most tasks are full, and clustering ranges
        # across y and z keeps the chain inside a task short
        assert nt <= 1.5 * m.N / 16
        assert dmax <= 2 * R + 4
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
