"""CPU definitions of the two-level preconditioner's pieces for the tests
(no counterpart in the reference): the Galerkin sum of fine blocks over
aggregate pairs and a textbook block ILU(0) with its solve.  block_ilu0 is
pinned against the oracle's textbook ILU(0) solve by
tests/test_coarse_plan.py."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def galerkin(A, rp, col, agg, crp, ccol):
    """sum of the fine blocks over each aggregate pair, as a dict-free scatter"""
    nc = len(crp) - 1
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    I, J = agg[rows], agg[col]
    key = I * nc + J
    ckey = np.repeat(np.arange(nc), np.diff(crp)) * nc + ccol
    pos = np.searchsorted(ckey, key)
    assert np.array_equal(ckey[pos], key)
    out = np.zeros((len(ckey), 4, 4))
    np.add.at(out, pos, A)
    return out


def block_ilu0(crp, ccol, blocks):
    """textbook block ILU(0) (IKJ order) on a BSR pattern: returns L (unit
    lower) and U as scipy CSR matrices"""
    n = len(crp) - 1
    a = blocks.copy()
    where = [dict((int(ccol[j]), j) for j in range(crp[i], crp[i + 1])) for i in range(n)]
    diag = [where[i][i] for i in range(n)]
    for i in range(n):
        for j in range(crp[i], crp[i + 1]):
            k = int(ccol[j])
            if k >= i:
                break
            a[j] = a[j] @ np.linalg.inv(a[diag[k]])
            for jk in range(diag[k] + 1, crp[k + 1]):
                t = where[i].get(int(ccol[jk]))
                if t is not None:
                    a[t] = a[t] - a[j] @ a[jk]
    rows = np.repeat(np.arange(n), np.diff(crp))

    def part(keep):
        ptr = np.concatenate([[0], np.cumsum(np.bincount(rows[keep], minlength=n))])
        return sp.bsr_matrix((a[keep], ccol[keep], ptr), shape=(4 * n, 4 * n)).tocsr()

    low = ccol < rows
    return part(low) + sp.identity(4 * n, format="csr"), part(~low)


def block_ilu0_solve(lu, b):
    """x = U^-1 L^-1 b; U's diagonal blocks are full 4x4 blocks, so the
    backward solve is a general sparse solve of the block upper triangle"""
    L, U = lu
    return spla.spsolve(U.tocsc(), spla.spsolve_triangular(L, b, lower=True))
