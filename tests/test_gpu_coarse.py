"""The two-level preconditioner (kf_options.coarse = 1) on the GPU.  It has
no counterpart in the reference; the fine level is the reference's block
ILU(0), already bit-identical to the oracle, so these tests pin what the
coarse level adds:

  * the device's aggregates and coarse pattern are the host plan's, and its
    Galerkin blocks P^T A P equal the sums of the oracle's Jacobian blocks
    over aggregate pairs (FP64 to 1e-14 relative per block; FP32 to 1e-6);
  * one application x = M b equals its definition assembled on the CPU:
    zc = (Lc Uc)^-1 P^T b with a textbook block ILU(0) of the Galerkin
    matrix (numpy), z = P zc, x = z + (LU)^-1 (b - A z) with the oracle's
    fine factors and solve -- within 1e-9 relative (2-norm) in FP64;
  * Kernel() with the two- and three-level preconditioners reaches the
    oracle's steady state (forces within 1e-8 relative, FP64, FP32 and BF16
    preconditioners) in no more Krylov iterations than block ILU(0) alone,
    and so does a four-level one whose fine and coarse ILU(0) solves are
    Jacobi sweeps (2 fine, 4 coarse);
  * on the 144x48x48 lattice, exact FP64 levels and the BF16 path with
    Jacobi sweeps (9 levels) reach the plain FP64 path's forces within 1e-8
    in under half its Krylov iterations;
  * the small levels' Jacobi solves in one workgroup (KF_JAC_SMALL, at
    most 128 block rows) give the same solve, bit for bit, as one launch
    per sweep: the whole history of a 5-level BF16 solve with sweeps, in
    two child processes (the switch is read once per process);
  * switching the number of levels on one context rebuilds or drops the
    deeper levels: a three-level, then a two-level, then a three-level
    solve on one mesh each repeat their own fresh-context history (with the
    textbook fine ILU(0): the reference-exact one carries the previous
    factorisation's L blocks from call to call, as the reference's does, so
    there a second solve on one context differs from the first).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import kfun3d_amd as K
import oracle as O
from conftest import perturbed_state
from coarse_ref import block_ilu0, block_ilu0_solve, galerkin

pytestmark = pytest.mark.gpu


def _seeded(which, mesh_a, mesh_c1, opts):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    m = K.Mesh(path, opts)
    orc = O.Oracle(m)
    q = perturbed_state(m.N, 3.0, 17)
    m.set_state(q)
    orc.set_state(q)
    m.timestep(30.0)
    orc.timestep(30.0)
    m.fill()
    orc.fill()
    return m, orc


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_galerkin_blocks(which, mesh_a, mesh_c1):
    m, orc = _seeded(which, mesh_a, mesh_c1, K.options(coarse=1))
    try:
        m.ilu()
        agg, crp, ccol, blocks = m.coarse_matrix()
        hagg, hcrp, hccol = m.coarse_plan()
        assert np.array_equal(agg, hagg) and np.array_equal(crp, hcrp) and np.array_equal(ccol, hccol)
        rp, col = orc.pattern()
        A = orc.blocks().reshape(-1, 4, 4)
        want = galerkin(A, rp.astype(np.int64), col.astype(np.int64), agg.astype(np.int64),
                         crp.astype(np.int64), ccol.astype(np.int64))
        scale = np.abs(want).max(axis=(1, 2), keepdims=True)
        assert np.all(np.abs(blocks - want) <= 1e-14 * scale + 1e-300), np.abs(blocks - want).max()
        m.set_options(K.options(coarse=1, prec="fp32"))
        m.fill()
        m.ilu()
        b32 = m.coarse_matrix()[3]
        assert np.all(np.abs(b32 - want) <= 1e-6 * scale), np.abs(b32 - want).max()
        assert m.last_error() == 0
    finally:
        orc.close()
        m.close()


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_apply_matches_definition(which, mesh_a, mesh_c1):
    m, orc = _seeded(which, mesh_a, mesh_c1, K.options(coarse=1))
    try:
        rp, col = orc.pattern()
        rp, col = rp.astype(np.int64), col.astype(np.int64)
        A = orc.blocks().reshape(-1, 4, 4).copy()
        m.ilu()
        orc.ilu()
        agg, crp, ccol, Ac = m.coarse_matrix()
        agg, crp, ccol = agg.astype(np.int64), crp.astype(np.int64), ccol.astype(np.int64)
        nc, N = len(crp) - 1, m.N
        b = np.random.default_rng(5).standard_normal(m.sz)
        bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
        m.sptrsv(bd, xd)
        got = xd.numpy()
        # the definition, on the CPU
        P = sp.csr_matrix((np.ones(4 * N), 4 * np.repeat(agg, 4) + np.tile(np.arange(4), N), np.arange(4 * N + 1)),
                          shape=(4 * N, 4 * nc))
        zc = block_ilu0_solve(block_ilu0(crp, ccol, Ac), P.T @ b)
        z = P @ zc
        Afine = sp.bsr_matrix((A, col, rp), shape=(4 * N, 4 * N)).tocsr()
        want = z + orc.sptrsv(b - Afine @ z)
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err <= 1e-9, err
        # the coarse correction is not zero: it changes the ILU(0) solve
        plain = orc.sptrsv(b)
        assert np.linalg.norm(got - plain) > 1e-3 * np.linalg.norm(plain)
        assert m.last_error() == 0
    finally:
        orc.close()
        m.close()


@pytest.mark.parametrize("cfg", [(1, 0, 0), (2, 0, 0), (3, 2, 4)], ids=["2lv", "3lv", "4lv-j2-cj4"])
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_kernel_multilevel(which, prec, cfg, mesh_a, mesh_c1):
    levels, jac, cj = cfg
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    hm = K.Mesh(path, device=False)
    orc = O.Oracle(hm, **O.config_options())
    ref = orc.kernel()
    orc.close()
    hm.close()
    m = K.Mesh(path, K.config_options(coarse=levels, prec=prec, jacobi_sweeps=jac, coarse_sweeps=cj))
    try:
        h = m.kernel()
        assert m.last_error() == 0
    finally:
        m.close()
    assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10
    f, f0 = np.array(h[-1]["forces"]), np.array(ref[-1]["forces"])
    assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (f, f0)
    k2, k1 = sum(s["krylov"] for s in h), sum(s["krylov"] for s in ref)
    print("%s %s: %d-level (sweeps %d fine, %d coarse) %d pseudo-steps, %d Krylov; ILU(0) %d, %d"
          % (which, prec, levels + 1, jac, cj, len(h), k2, len(ref), k1))
    if not jac:
        assert k2 <= k1


def test_levels_switch(mesh_c1):
    runs = {}
    for levels in (2, 1):
        m = K.Mesh(mesh_c1, K.config_options(coarse=levels, ilu_exact=0))
        try:
            runs[levels] = m.kernel()
        finally:
            m.close()
    m = K.Mesh(mesh_c1, K.config_options(coarse=2, ilu_exact=0))
    try:
        q0 = m.state()
        for levels in (2, 1, 2):
            m.set_options(K.config_options(coarse=levels, ilu_exact=0))
            m.set_state(q0)
            h = m.kernel()
            assert [s["krylov"] for s in h] == [s["krylov"] for s in runs[levels]]
            assert h[-1]["forces"] == runs[levels][-1]["forces"]
        assert m.last_error() == 0
    finally:
        m.close()


@pytest.mark.slow
def test_multilevel_at_144x48x48(mesh_dir):
    """The 0.33M-vertex lattice (seed 1): the multilevel paths -- exact FP64
    levels, and BF16 factors with 2 fine and 3 coarse Jacobi sweeps over 9
    levels -- reach the plain FP64 path's forces within 1e-8 in under half
    its Krylov iterations (the FP64 path is the reference's algorithm, held
    to the oracle by the parity tests)."""
    from conftest import lattice
    path = lattice(mesh_dir, 144, 48, 48, 1)
    m = K.Mesh(path, K.config_options())
    try:
        q0 = m.state()
        ref = m.kernel()
        f0 = np.array(ref[-1]["forces"])
        k1 = sum(s["krylov"] for s in ref)
        for opts in (dict(coarse=8), dict(prec="bf16", coarse=8, jacobi_sweeps=2, coarse_sweeps=3)):
            m.set_options(K.config_options(**opts))
            m.set_state(q0)
            h = m.kernel()
            assert m.last_error() == 0
            assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10
            f = np.array(h[-1]["forces"])
            assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (opts, f, f0)
            k2 = sum(s["krylov"] for s in h)
            print("144x48x48 %s: %d Krylov against %d" % (opts, k2, k1))
            assert 2 * k2 < k1
    finally:
        m.close()


_SMALL_CHILD = """
import sys
sys.path.insert(0, sys.argv[2])
import kfun3d_amd as K
m = K.Mesh(sys.argv[1], K.config_options(prec="bf16", coarse=4, jacobi_sweeps=2, coarse_sweeps=3))
h = m.kernel()
assert m.last_error() == 0
m.close()
for s in h:
    print(s["krylov"], s["fnorm"].hex(), " ".join(f.hex() for f in s["forces"]))
"""


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_small_levels_in_one_workgroup(which, mesh_a, mesh_c1):
    import os
    import subprocess
    import sys
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    out = {}
    for flag in ("1", "0"):
        env = dict(os.environ, KF_JAC_SMALL=flag)
        r = subprocess.run([sys.executable, "-c", _SMALL_CHILD, path, root], env=env, stdout=subprocess.PIPE,
                           stderr=subprocess.PIPE, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        out[flag] = r.stdout.split("\n")
    assert len(out["1"]) > 3
    assert out["1"] == out["0"]

