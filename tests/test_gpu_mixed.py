"""Mixed-precision preconditioner (FP32 fill, ILU(0) factors and triangular
solves; the BF16 variant stores the factors in BF16 and computes in FP32),
batched classical Gram-Schmidt and flexible GMRES, each held to the oracle's
FP64 values (the one-thread restatement the CPU tests pin bit for bit to the
reference).  The residual, the Arnoldi process, the Hessenberg least squares
and the update stay FP64, so only the preconditioner changes.

Bars (measured errors are about ten times smaller):
  * FP32 Jacobian fill and ILU(0) factors: max |difference| <= 3e-6 of the
    largest FP64 entry; one preconditioner application M^-1 b: 2-norm error
    <= 3e-6 relative;
  * BF16 factors (8-bit mantissa): M^-1 b within 2e-2 relative;
  * whole solve to a 1e-10 residual drop (configuration profile): the same
    number of pseudo-time steps as the oracle's FP64 run, Krylov counts
    within 10 percent, final forces within 1e-8 relative of it.
The 216x72x72 lattice (seed 2, 1,119,744 vertices, second order) is run the
same way: FP64, FP32 and BF16 solves to a 1e-10 drop, and one
preconditioner application against the oracle's FP64 fill, ILU and solve.
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O
from conftest import lattice, perturbed_state

pytestmark = pytest.mark.gpu

_ORACLE_RUNS = {}


def oracle_solve(path, first_order):
    """the oracle's FP64 configuration-profile solve to a 1e-10 drop, in the
    reference's own (sequential) reduction order; cached per mesh"""
    key = (path, first_order)
    if key not in _ORACLE_RUNS:
        m = K.Mesh(path, device=False)
        orc = O.Oracle(m, **O.config_options(first_order=first_order))
        _ORACLE_RUNS[key] = orc.kernel()
        orc.close()
        m.close()
    return _ORACLE_RUNS[key]


def oracle_precond(m, q, b):
    """the oracle's FP64 Jacobian, ILU(0) factors and M^-1 b at state q"""
    orc = O.Oracle(m)
    try:
        orc.set_state(q)
        orc.timestep(50.0)
        orc.fill()
        A = orc.blocks()
        orc.ilu()
        F = orc.blocks()
        x = orc.sptrsv(b)
    finally:
        orc.close()
    return A, F, x


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
    x = xd.numpy()
    bd.free()
    xd.free()
    return A, F, x


def _check_precond(A64, F64, x64, A32, F32, x32, xbf):
    assert np.abs(A32 - A64).max() <= 3e-6 * np.abs(A64).max()
    assert np.abs(F32 - F64).max() <= 3e-6 * np.abs(F64).max()
    e32 = np.linalg.norm(x32 - x64) / np.linalg.norm(x64)
    assert e32 <= 3e-6, e32
    e_bf = np.linalg.norm(xbf - x64) / np.linalg.norm(x64)
    assert 1e-5 < e_bf <= 2e-2, e_bf   # BF16 rounding is visible, and bounded
    return e32, e_bf


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_fp32_bf16_preconditioner(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    m = K.Mesh(path, K.options())
    try:
        q = perturbed_state(m.N, 3.0, 21)
        b = np.random.default_rng(22).standard_normal(m.sz)
        A64, F64, x64 = oracle_precond(m, q, b)
        A32, F32, x32 = _precond(m, "fp32", q, b)
        _, _, xbf = _precond(m, "bf16", q, b)
    finally:
        m.close()
    _check_precond(A64.ravel(), F64.ravel(), x64, A32.ravel(), F32.ravel(), x32, xbf)


def _check_solve_against(h, ref, what, krylov_tol=0.1, steps_tol=0):
    assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10, what
    assert abs(len(h) - len(ref)) <= steps_tol, (what, len(h), len(ref))
    k, k0 = sum(s["krylov"] for s in h), sum(s["krylov"] for s in ref)
    assert abs(k - k0) <= krylov_tol * k0, (what, k, k0)
    f, f0 = np.array(h[-1]["forces"]), np.array(ref[-1]["forces"])
    assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (what, f, f0)


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_mixed_solve_matches_fp64(which, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    first = 1 if which == "mesh_c1" else 0     # configuration 1 is first order
    ref = oracle_solve(path, first)
    for prec in ("fp64", "fp32", "bf16"):
        m = K.Mesh(path, K.config_options(prec=prec, first_order=first))
        try:
            h = m.kernel()
            assert m.last_error() == 0
        finally:
            m.close()
        _check_solve_against(h, ref, prec)


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_cgs2_matches_mgs(which, mesh_a, mesh_c1):
    """Batched classical Gram-Schmidt with reorthogonalisation (two passes,
    one reduction each) in place of the reference's modified Gram-Schmidt:
    the same pseudo-time steps as the oracle's MGS run, Krylov counts within
    5 percent, forces within 1e-8 relative, at a 1e-10 residual drop."""
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    ref = oracle_solve(path, 0)
    m = K.Mesh(path, K.config_options(ortho="cgs2"))
    try:
        h = m.kernel()
        assert m.last_error() == 0
    finally:
        m.close()
    _check_solve_against(h, ref, "cgs2", krylov_tol=0.05)


@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_fgmres_ilu_matches_gmres(which, mesh_a, mesh_c1):
    """Right-preconditioned flexible GMRES with the ILU(0) preconditioner
    (the frame the learned correction plugs into) reaches the 1e-10 drop and
    the same converged solution as the oracle's left-preconditioned GMRES:
    forces within 1e-8 relative.  The two stop their inner solves on
    different residuals (unpreconditioned / preconditioned), so the inexact
    Newton steps differ: pseudo-time steps within 3, Krylov totals within a
    factor 2."""
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    ref = oracle_solve(path, 0)
    m = K.Mesh(path, K.config_options(krylov="fgmres"))
    try:
        h = m.kernel()
        assert m.last_error() == 0
    finally:
        m.close()
    _check_solve_against(h, ref, "fgmres", krylov_tol=1.0, steps_tol=3)


@pytest.fixture(scope="module")
def mesh_c3(mesh_dir):
    """Configuration 3: 216x72x72 Kuhn lattice, seed 2 (1,119,744 vertices)."""
    return lattice(mesh_dir, 216, 72, 72, 2)


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_c3_preconditioner_against_oracle(mesh_c3):
    """One FP32 and one BF16 preconditioner build and application on the
    216x72x72 lattice against the oracle's FP64 fill, ILU(0) and solve."""
    m = K.Mesh(mesh_c3, K.options())
    try:
        q = perturbed_state(m.N, 3.0, 23)
        b = np.random.default_rng(24).standard_normal(m.sz)
        print("c3: oracle preconditioner", flush=True)
        A64, F64, x64 = oracle_precond(m, q, b)
        print("c3: FP32 and BF16 preconditioners", flush=True)
        A32, F32, x32 = _precond(m, "fp32", q, b)
        _, _, xbf = _precond(m, "bf16", q, b)
    finally:
        m.close()
    e32, ebf = _check_precond(A64.ravel(), F64.ravel(), x64, A32.ravel(), F32.ravel(), x32, xbf)
    print("c3: M^-1 b relative error FP32 %.2e, BF16 %.2e" % (e32, ebf))


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_c3_mixed_solves(mesh_c3):
    """Solves to a 1e-10 drop on the 216x72x72 lattice (second order): the
    FP32 and BF16-factor preconditioners against the plain FP64 path (which
    the other tests hold to the oracle): the same pseudo-time steps, Krylov
    totals within 10 percent, final forces within 1e-8 relative."""
    m = K.Mesh(mesh_c3, K.config_options())
    hist = {}
    try:
        q0 = m.state()
        for prec in ("fp64", "fp32", "bf16"):
            print("c3: %s solve" % prec, flush=True)
            m.set_options(K.config_options(prec=prec))
            m.set_state(q0)
            hist[prec] = m.kernel()
            assert m.last_error() == 0
    finally:
        m.close()
    for prec in ("fp32", "bf16"):
        _check_solve_against(hist[prec], hist["fp64"], prec)
    print("c3: Krylov %r" % {p: sum(s["krylov"] for s in h) for p, h in hist.items()})
