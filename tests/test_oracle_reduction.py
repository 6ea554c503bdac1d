"""Pins of the oracle's fixed-order tree reduction (or_dot_tree, the order of
the HIP path's norms and dot products) and of the solve history that order
produces against the reference's own one-thread history.

or_dot_tree has no counterpart in the reference: it restates the device's
reduction tree so that the GPU's default (fast) path can be checked bit for
bit.  Nothing here retypes that tree.  Instead:
  * integer-valued operands make every partial sum exact, so any summation
    order returns the exact dot product; a dropped, doubled or misindexed
    term (the half pair of an odd length, a stride that skips a pass, a
    partial left out of the final tree) changes the integer result;
  * on random data the tree must stay within the forward-error bound of
    its depth against a correctly rounded sum (math.fsum);
  * 0, 1 and 2 elements, odd lengths, lengths around one grid pass
    (2 x 1024 x 256 elements) and several passes are all covered.
Then the history: with tree-ordered reductions the oracle's Newton-Krylov
run (which the GPU reproduces bit for bit) is held to the reference's own
one-thread run (tests/golden) within the measured bounds stated below.
"""
import json
import math
import os

import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import GOLDEN

PASS = 2 * 1024 * 256          # elements one pass of the 1024 x 256 grid covers
U = 2.0 ** -53                  # unit roundoff of IEEE double


LENGTHS = [0, 1, 2, 3, 5, 255, 256, 257, 1023, 2047, 2049, PASS - 1, PASS, PASS + 1, PASS + 2, 3 * PASS - 7,
           4 * 331776, 4 * 331776 + 1]


@pytest.mark.parametrize("n", LENGTHS)
def test_tree_exact_on_integer_data(n):
    """|x_i|, |y_i| <= 64: every product and partial sum is an integer below
    2^53, so the tree must return the exact integer dot product."""
    rng = np.random.default_rng(1000 + n)
    x = rng.integers(-64, 65, n).astype(np.float64)
    y = rng.integers(-64, 65, n).astype(np.float64)
    exact = int(np.dot(x.astype(np.int64), y.astype(np.int64)))
    assert O.dot_tree(x, y) == float(exact)
    assert O.dot_tree(x, x) == float(int(np.dot(x.astype(np.int64), x.astype(np.int64))))


@pytest.mark.parametrize("n", [PASS - 1, PASS + 1, 3 * PASS - 7])
def test_tree_counts_every_term_once(n):
    """A single unit entry at each position class (first, last, the half pair
    of an odd length, the first element of the second pass, a block
    boundary) contributes exactly once."""
    for i in sorted({0, n - 1, n // 2, min(PASS, n - 1), 256 * 2 - 1, 256 * 2, 1023 * 512}):
        if i >= n:
            continue
        x = np.zeros(n)
        y = np.zeros(n)
        x[i] = 3.0
        y[i] = 5.0
        assert O.dot_tree(x, y) == 15.0, i


def _depth(n):
    """additions any term passes through: the strided pair chain of one
    thread (two adds per pair and pass), the 256-thread block tree (8), the
    1024-partial strided sum (4) and the final tree (8), plus the product"""
    passes = max(1, -(-n // PASS))
    return 2 * passes + 8 + 4 + 8 + 1


@pytest.mark.parametrize("n", [7, 1000, PASS - 3, PASS + 5, 2 * PASS + 11, 4 * 331776])
def test_tree_within_forward_error_bound(n):
    """|tree - fsum| <= gamma_k sum |x_i y_i| with k the depth of the tree
    (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed., 4.2);
    the sequential sum's bound would be gamma_n, far looser."""
    rng = np.random.default_rng(77 + n)
    x = rng.standard_normal(n) * np.exp(rng.uniform(-8, 8, n))
    y = rng.standard_normal(n)
    got = O.dot_tree(x, y)
    want = math.fsum((x * y).tolist())   # products are exact to 1 ulp: included in k
    k = _depth(n)
    gamma = k * U / (1 - k * U)
    bound = gamma * float(np.sum(np.abs(x * y)))
    assert abs(got - want) <= bound, (got, want, bound)


def test_tree_norm_matches_sqrt_of_tree_dot():
    """or_norm is the reference's sequential norm; sqrt(dot_tree(x, x)) is the
    device's.  Both must be within their bounds of the exact norm."""
    rng = np.random.default_rng(5)
    x = rng.standard_normal(PASS + 3)
    exact = math.sqrt(math.fsum((x * x).tolist()))
    assert abs(math.sqrt(O.dot_tree(x, x)) - exact) <= 64 * U * exact
    assert abs(O.norm(x) - exact) <= x.size * U * exact


def _golden_solve(name):
    with open(os.path.join(GOLDEN, "ref_%s.json" % name)) as f:
        return json.load(f)["solve"]


# Measured deviation of the tree-ordered history from the reference's
# one-thread history (same Krylov count at every step on both meshes):
#   mesh A: |d fnorm| / fnorm <= 1.2e-3 (step 7, drop 9e7),
#           |d fnorm| / fnorm_0 <= 8.4e-7 (step 2), forces <= 9.8e-8 mid-run,
#           final forces 2.9e-14;
#   C1:     |d fnorm| / fnorm <= 3.2e-5, |d fnorm| / fnorm_0 <= 4.2e-8,
#           final forces 3.4e-13.
# Newton-Krylov amplifies a one-ulp change of a norm into these; the bars
# hold them with about a 2x margin.
TREE_BARS = dict(rel=3e-3, rel0=2e-6, forces_mid=3e-7, forces_final=1e-8)


def tree_history_deviation(hist, ref):
    """per-history maxima of the deviations the bars above limit"""
    f0 = ref[0]["fnorm"]
    out = dict(rel=0.0, rel0=0.0, forces_mid=0.0, forces_final=0.0, krylov_equal=True, steps_equal=len(hist) == len(ref))
    for a, b in zip(hist, ref):
        d = abs(a["fnorm"] - b["fnorm"])
        out["rel"] = max(out["rel"], d / b["fnorm"])
        out["rel0"] = max(out["rel0"], d / f0)
        out["forces_mid"] = max(out["forces_mid"], max(abs(u - v) / abs(v) for u, v in zip(a["forces"], b["forces"])))
        out["krylov_equal"] &= a["krylov"] == b["krylov"]
    out["forces_final"] = max(abs(u - v) / abs(v) for u, v in zip(hist[-1]["forces"], ref[-1]["forces"]))
    return out


def check_tree_history(hist, ref):
    d = tree_history_deviation(hist, ref)
    assert d["steps_equal"], (len(hist), len(ref))
    assert d["krylov_equal"], [(a["krylov"], b["krylov"]) for a, b in zip(hist, ref)]
    for k, bar in TREE_BARS.items():
        assert d[k] <= bar, (k, d[k], bar)
    return d


@pytest.mark.parametrize("name", ["mesh_a", "c1"])
def test_tree_history_close_to_reference(name, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "c1": mesh_c1}[name]
    m = K.Mesh(path, device=False)
    orc = O.Oracle(m, reduce_order=1)
    hist = orc.kernel()
    orc.close()
    m.close()
    check_tree_history(hist, _golden_solve(name))
