"""Mixed-precision preconditioner (FP32 fill, ILU(0) factors and triangular
solves; the BF16 variant stores the factors in BF16 and computes in FP32)
against the plain FP64 path, whose results the parity tests hold bit for bit
to the oracle.  The residual, the Arnoldi process, the Hessenberg least
squares and the update stay FP64, so only the preconditioner changes.

Bars (measured errors are about ten times smaller):
  * FP32 Jacobian fill and ILU(0) factors: max |difference| <= 3e-6 of the
    largest FP64 entry; one preconditioner application M^-1 b: 2-norm error
    <= 3e-6 relative;
  * BF16 factors (8-bit mantissa): M^-1 b within 2e-2 relative;
  * whole solve to a 1e-10 residual drop (config profile): the same number
    of pseudo-time steps, Krylov counts within 10 percent, final forces
    within 1e-8 relative of the FP64 path.
"""
import numpy as np
import pytest

import kfun3d_amd as K
from conftest import perturbed_state

pytestmark = pytest.mark.gpu


def _precond(m, prec, q, b):
    m.set_options(K.options(prec=prec))
    m.set_state(q)
    m.timestep(50.0)
    m.fill()
    A = m.blocks().copy()
    m.ilu()
    F = m.blocks().copy()
    bd, xd = K.DeviceVec(m.sz, b), K.DeviceVec(m.sz)
    m.sptrsv(bd, xd)
    assert m.last_error() == 0
    return A, F, xd.numpy()


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_fp32_bf16_preconditioner(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    m = K.Mesh(path, K.options())
    try:
        q = perturbed_state(m.N, 3.0, 21)
        b = np.random.default_rng(22).standard_normal(m.sz)
        A64, F64, x64 = _precond(m, "fp64", q, b)
        A32, F32, x32 = _precond(m, "fp32", q, b)
        _, _, xbf = _precond(m, "bf16", q, b)
    finally:
        m.close()
    assert np.abs(A32 - A64).max() <= 3e-6 * np.abs(A64).max()
    assert np.abs(F32 - F64).max() <= 3e-6 * np.abs(F64).max()
    assert np.linalg.norm(x32 - x64) <= 3e-6 * np.linalg.norm(x64)
    e_bf = np.linalg.norm(xbf - x64) / np.linalg.norm(x64)
    assert 1e-5 < e_bf <= 2e-2, e_bf   # BF16 rounding is visible, and bounded


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_mixed_solve_matches_fp64(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    first = 1 if which == "mesh_c1" else 0     # configuration 1 is first order
    hist = {}
    for prec in ("fp64", "fp32", "bf16"):
        m = K.Mesh(path, K.config_options(prec=prec, first_order=first))
        try:
            hist[prec] = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
    h64 = hist["fp64"]
    assert h64[0]["fnorm"] / h64[-1]["fnorm"] >= 1e10
    f64 = np.array(h64[-1]["forces"])
    k64 = sum(s["krylov"] for s in h64)
    for prec in ("fp32", "bf16"):
        h = hist[prec]
        assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10, prec
        assert len(h) == len(h64), (prec, len(h), len(h64))
        k = sum(s["krylov"] for s in h)
        assert abs(k - k64) <= 0.1 * k64, (prec, k, k64)
        f = np.array(h[-1]["forces"])
        assert np.all(np.abs(f - f64) <= 1e-8 * np.abs(f64)), (prec, f, f64)


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_cgs2_matches_mgs(which, mesh_a, mesh_c1):
    """Batched classical Gram-Schmidt with reorthogonalisation (two passes,
    one reduction each) in place of the reference's modified Gram-Schmidt:
    the same pseudo-time steps, Krylov counts within 5 percent, forces
    within 1e-8 relative, at a 1e-10 residual drop."""
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    hist = {}
    for ortho in ("mgs", "cgs2"):
        m = K.Mesh(path, K.config_options(ortho=ortho))
        try:
            hist[ortho] = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
    hm, hc = hist["mgs"], hist["cgs2"]
    assert hc[0]["fnorm"] / hc[-1]["fnorm"] >= 1e10
    assert len(hc) == len(hm)
    km, kc = sum(s["krylov"] for s in hm), sum(s["krylov"] for s in hc)
    assert abs(kc - km) <= 0.05 * km, (kc, km)
    fm, fc = np.array(hm[-1]["forces"]), np.array(hc[-1]["forces"])
    assert np.all(np.abs(fc - fm) <= 1e-8 * np.abs(fm)), (fc, fm)


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_fgmres_ilu_matches_gmres(which, mesh_a, mesh_c1):
    """Right-preconditioned flexible GMRES with the ILU(0) preconditioner
    (the frame the learned correction plugs into) reaches the 1e-10 drop and
    the same converged solution as the reference's left-preconditioned GMRES:
    forces within 1e-8 relative.  The two stop their inner solves on
    different residuals (unpreconditioned / preconditioned), so the inexact
    Newton steps differ: pseudo-time steps within 3, Krylov totals within a
    factor 2."""
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    hist = {}
    for kry in ("gmres", "fgmres"):
        m = K.Mesh(path, K.config_options(krylov=kry))
        try:
            hist[kry] = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
    hg, hf = hist["gmres"], hist["fgmres"]
    assert hf[0]["fnorm"] / hf[-1]["fnorm"] >= 1e10
    assert abs(len(hf) - len(hg)) <= 3, (len(hf), len(hg))
    kg, kf = sum(s["krylov"] for s in hg), sum(s["krylov"] for s in hf)
    assert kf <= 2 * kg and kg <= 2 * kf, (kf, kg)
    fg, ff = np.array(hg[-1]["forces"]), np.array(hf[-1]["forces"])
    assert np.all(np.abs(ff - fg) <= 1e-8 * np.abs(fg)), (ff, fg)
