"""Pins of the oracle that need no reference: closed forms and identities the
mathematics fixes.  Each is chosen so that a plausible slip (a dropped term,
a wrong sign or index, a transposed block) fails it."""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import lattice


@pytest.fixture(scope="module")
def small(mesh_dir):
    m = K.Mesh(lattice(mesh_dir, 7, 6, 5, 3), device=False)
    yield m
    m.close()


def freestream(N, alpha=3.0):
    a = alpha / (180 / np.pi)
    return np.tile(np.array([1.0, np.cos(a), np.sin(a), 0.0]), N)


def interior_rows(m):
    """vertices none of whose neighbours is a boundary vertex"""
    h = m.host_arrays()
    bnd = np.zeros(m.N, bool)
    bnd[h["s_nptr"]] = True
    bnd[h["f_nptr"]] = True
    near = bnd.copy()
    np.logical_or.at(near, h["e_n0"], bnd[h["e_n1"]])
    np.logical_or.at(near, h["e_n1"], bnd[h["e_n0"]])
    return np.where(~near)[0], np.where(~bnd)[0]


def test_geometric_closure(small):
    h = small.host_arrays()
    c = np.zeros((small.N, 3))
    nA = np.stack([h["e_x0"] * h["e_x3"], h["e_x1"] * h["e_x3"], h["e_x2"] * h["e_x3"]], 1)
    np.add.at(c, h["e_n0"], nA)
    np.add.at(c, h["e_n1"], -nA)
    np.add.at(c, h["s_nptr"], np.stack([h["s_x0"], h["s_x1"], h["s_x2"]], 1))
    np.add.at(c, h["f_nptr"], np.stack([h["f_x0"], h["f_x1"], h["f_x2"]], 1))
    assert np.abs(c).max() < 1e-14


def test_freestream_preserved_in_interior(small):
    orc = O.Oracle(small)
    r, _, _, _ = orc.flux(freestream(small.N))
    orc.close()
    _, interior = interior_rows(small)
    assert np.abs(r.reshape(-1, 4)[interior]).max() < 1e-14


def test_lsq_gradient_exact_on_linear_fields(small):
    h = small.host_arrays()
    x, y, z = h["n_x0"], h["n_x1"], h["n_x2"]
    slopes = np.array([[0.3, -1.2, 0.7], [2.0, 0.5, -0.25], [-0.4, 0.1, 1.5], [1.0, -2.0, 0.125]])
    q = np.stack([0.5 + s[0] * x + s[1] * y + s[2] * z for s in slopes], 1).ravel()
    orc = O.Oracle(small)
    _, gx, gy, gz = orc.flux(q)
    orc.close()
    for c in range(4):
        assert np.allclose(gx.reshape(-1, 4)[:, c], slopes[c, 0], atol=1e-11)
        assert np.allclose(gy.reshape(-1, 4)[:, c], slopes[c, 1], atol=1e-11)
        assert np.allclose(gz.reshape(-1, 4)[:, c], slopes[c, 2], atol=1e-11)


def _dense(rowptr, col, blocks, N):
    A = np.zeros((4 * N, 4 * N))
    for i in range(N):
        for j in range(rowptr[i], rowptr[i + 1]):
            A[4 * i:4 * i + 4, 4 * col[j]:4 * col[j] + 4] = blocks[j]
    return A


def test_jacobian_matches_central_differences_at_uniform_state(small):
    """At a uniform state the Roe matrix |A| multiplies q_R - q_L = 0, so the
    frozen-|A| first-order Jacobian is the exact derivative of the first-order
    residual on rows whose stencil avoids the boundary terms."""
    q = freestream(small.N)
    orc = O.Oracle(small, first_order=1)
    orc.set_state(q)
    orc.timestep(1e300)          # cdt ~ 0: A is the flux Jacobian alone
    orc.fill()
    rp, col = orc.pattern()
    A = _dense(rp, col, orc.blocks(), small.N)
    d = np.random.default_rng(0).standard_normal(small.sz)
    eps = 1e-6
    rp_, _, _, _ = orc.flux(q + eps * d)
    rm_, _, _, _ = orc.flux(q - eps * d)
    fd = (rp_ - rm_) / (2 * eps)
    orc.close()
    rows, _ = interior_rows(small)
    sel = (4 * rows[:, None] + np.arange(4)).ravel()
    Ad = A @ d
    assert np.abs(Ad[sel] - fd[sel]).max() <= 1e-7 * np.abs(fd[sel]).max()


def _filled(small, seed):
    orc = O.Oracle(small)
    q = freestream(small.N) + 0.05 * np.random.default_rng(seed).standard_normal(small.sz)
    orc.set_state(q)
    orc.timestep(20.0)
    orc.fill()
    return orc


def test_bsr_product_matches_scipy(small):
    """or_spmv (the device product's order) against scipy's BSR product of
    the same blocks, pattern and vector: the sums differ only in order."""
    import scipy.sparse as sp
    orc = _filled(small, 4)
    rp, col = orc.pattern()
    A = sp.bsr_matrix((orc.blocks(), col.astype(np.int64), rp.astype(np.int64)), shape=(small.sz, small.sz))
    x = np.random.default_rng(5).standard_normal(small.sz)
    y = orc.spmv(x)
    orc.close()
    want = A @ x
    assert np.abs(y - want).max() <= 1e-13 * np.abs(want).max()


def test_bsr_product_extracts_columns_exactly(small):
    """With x a unit vector every row sums exact zeros and one product a * 1,
    so y is column k of A bit for bit: a transposed block or a wrong column
    index would move the entries."""
    orc = _filled(small, 6)
    rp, col = orc.pattern()
    B = orc.blocks()
    rng = np.random.default_rng(7)
    for k in rng.choice(small.sz, 12, replace=False):
        x = np.zeros(small.sz)
        x[k] = 1.0
        y = orc.spmv(x)
        want = np.zeros(small.sz)
        v, cc = divmod(int(k), 4)
        for i in range(small.N):
            for j in range(rp[i], rp[i + 1]):
                if col[j] == v:
                    want[4 * i:4 * i + 4] = B[j][:, cc]
        assert np.array_equal(y, want), k
    orc.close()


def test_ilu0_defining_property_and_triangular_solve(small):
    orc = O.Oracle(small, ilu_textbook=1)
    q = freestream(small.N) + 0.05 * np.random.default_rng(1).standard_normal(small.sz)
    orc.set_state(q)
    orc.timestep(20.0)
    orc.fill()
    rp, col = orc.pattern()
    A = _dense(rp, col, orc.blocks(), small.N)
    orc.ilu()
    F = orc.blocks()
    N = small.N
    L = np.eye(4 * N)
    U = np.zeros((4 * N, 4 * N))
    for i in range(N):
        for j in range(rp[i], rp[i + 1]):
            c = col[j]
            blk = F[j]
            if c < i:
                L[4 * i:4 * i + 4, 4 * c:4 * c + 4] = blk
            elif c == i:
                U[4 * i:4 * i + 4, 4 * c:4 * c + 4] = np.linalg.inv(blk)
            else:
                U[4 * i:4 * i + 4, 4 * c:4 * c + 4] = blk
    LU = L @ U
    mask = _dense(rp, col, np.ones((len(col), 4, 4)), N) > 0
    assert np.abs((LU - A)[mask]).max() <= 1e-12 * np.abs(A).max()
    b = np.random.default_rng(2).standard_normal(small.sz)
    x = orc.sptrsv(b)
    orc.close()
    assert np.abs(LU @ x - b).max() <= 1e-10 * np.abs(b).max()


def test_reference_ilu_close_to_textbook(small):
    """ilu_Edmond with one thread is ILU(0) up to the (a - d) + d rounding of
    its copy side effects."""
    q = freestream(small.N) + 0.05 * np.random.default_rng(3).standard_normal(small.sz)
    out = []
    for tb in (0, 1):
        orc = O.Oracle(small, ilu_textbook=tb)
        orc.set_state(q)
        orc.timestep(20.0)
        orc.fill()
        orc.ilu()
        out.append(orc.blocks())
        orc.close()
    assert np.abs(out[0] - out[1]).max() <= 1e-12 * np.abs(out[1]).max()


def test_inverse_matches_numpy():
    rng = np.random.default_rng(4)
    for _ in range(50):
        a = rng.standard_normal((4, 4)) + 4 * np.eye(4)
        assert np.allclose(O.inv4(a), np.linalg.inv(a), rtol=1e-12, atol=1e-14)


def test_wall_forces_vanish_at_unit_pressure(small):
    orc = O.Oracle(small)
    q = freestream(small.N)
    q[0::4] = 1.0
    orc.set_state(q)
    assert np.array_equal(orc.forces(), np.zeros(3))
    q[0::4] = 1.5
    orc.set_state(q)
    assert np.abs(orc.forces()).max() > 0
    orc.close()


def _factor_parts(orc, N):
    """dense strictly-lower L, strictly-upper U and block-diagonal D^-1 of the
    in-place factors (the diagonal blocks hold D^-1 already)"""
    rp, col = orc.pattern()
    F = orc.blocks()
    L = np.zeros((4 * N, 4 * N))
    U = np.zeros((4 * N, 4 * N))
    Dinv = np.zeros((4 * N, 4 * N))
    for i in range(N):
        for j in range(rp[i], rp[i + 1]):
            c = col[j]
            dst = L if c < i else (Dinv if c == i else U)
            dst[4 * i:4 * i + 4, 4 * c:4 * c + 4] = F[j]
    return L, U, Dinv


@pytest.fixture(scope="module")
def factored(small):
    orc = O.Oracle(small)
    q = freestream(small.N) + 0.05 * np.random.default_rng(5).standard_normal(small.sz)
    orc.set_state(q)
    orc.timestep(20.0)
    orc.fill()
    orc.ilu()
    yield orc, _factor_parts(orc, small.N)
    orc.close()


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_jacobi_sweeps_closed_form(small, factored, sweeps):
    """k Jacobi sweeps each way (matrix/sptrsv.cpp:75-132) written with dense
    matrices: y_s = b - L y_{s-1} from y_0 = b; x_0 = D^-1 y_k;
    x_s = D^-1 (y_k - U x_{s-1})."""
    orc, (L, U, Dinv) = factored
    b = np.random.default_rng(6).standard_normal(small.sz)
    y = b.copy()
    for _ in range(sweeps):
        y = b - L @ y
    x = Dinv @ y
    for _ in range(sweeps):
        x = Dinv @ (y - U @ x)
    got = orc.sptrsv_jacobi(b, sweeps)
    assert np.abs(got - x).max() <= 1e-12 * np.abs(x).max()


def test_jacobi_sweeps_reach_the_exact_solve(small, factored):
    """Sweep s fixes every row whose dependency chain is at most s long, so
    once s reaches the depth of both triangular DAGs the sweeps give the
    level-scheduled solve (up to the order of the sums); one sweep short of
    that they do not."""
    orc, (L, U, Dinv) = factored
    N = small.N
    rp, col = orc.pattern()
    depth_l = np.zeros(N, int)
    for i in range(N):
        lo = [depth_l[c] for c in col[rp[i]:rp[i + 1]] if c < i]
        depth_l[i] = 1 + max(lo, default=0)
    depth_u = np.zeros(N, int)
    for i in range(N - 1, -1, -1):
        hi = [depth_u[c] for c in col[rp[i]:rp[i + 1]] if c > i]
        depth_u[i] = 1 + max(hi, default=0)
    k = int(max(depth_l.max(), depth_u.max()))
    b = np.random.default_rng(7).standard_normal(small.sz)
    exact = orc.sptrsv(b)
    scale = np.abs(exact).max()
    got = orc.sptrsv_jacobi(b, k)
    assert np.abs(got - exact).max() <= 1e-14 * scale
    # the error falls with every pair of sweeps and is gone only at the depth
    errs = [np.abs(orc.sptrsv_jacobi(b, s) - exact).max() / scale for s in (1, 3, k - 2)]
    assert errs[0] > 1e-3 and errs[0] > errs[1] > errs[2] > 1e-12, errs
