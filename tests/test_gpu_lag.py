"""kf_options.lag_factors = k keeps the last factorisation (Jacobian fill
and factorisation skipped) while the step's CFL is at least the CFL it was
formed at and below k times it -- k = 1: while unchanged, once the ramp has
reached cfl_max; the Krylov operator stays the exact matrix-free product
(src/kernel.c:24-45).  No counterpart in the reference, which refactors at
every step (src/kernel.c:91-94).

  * with plain ILU(0) and with the multilevel preconditioner, k = 1 and 4,
    the solve reaches the oracle's steady state (forces within 1e-8
    relative);
  * fewer factorisations run than pseudo-steps (counted by the profiler),
    and exactly one per step without the option.
"""
import numpy as np
import pytest

import kfun3d_amd as K
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lagk", [1, 4])
@pytest.mark.parametrize("opts", [dict(), dict(prec="bf16", coarse=3, jacobi_sweeps=2, coarse_sweeps=3)],
                         ids=["ilu0", "multilevel"])
@pytest.mark.parametrize("which", ["mesh_a", "mesh_c1"])
def test_lagged_preconditioner(which, opts, lagk, mesh_a, mesh_c1):
    path = {"mesh_a": mesh_a, "mesh_c1": mesh_c1}[which]
    hm = K.Mesh(path, device=False)
    orc = O.Oracle(hm, **O.config_options())
    ref = orc.kernel()
    orc.close()
    hm.close()
    counts = {}
    for lag in (0, lagk):
        m = K.Mesh(path, K.config_options(lag_factors=lag, **opts))
        try:
            m.profile(True)
            h = m.kernel()
            counts[lag] = (len(h), m.profile_get("ilu")[1])
            m.profile(False)
            assert m.last_error() == 0
        finally:
            m.close()
        assert h[0]["fnorm"] / h[-1]["fnorm"] >= 1e10
        f, f0 = np.array(h[-1]["forces"]), np.array(ref[-1]["forces"])
        assert np.all(np.abs(f - f0) <= 1e-8 * np.abs(f0)), (lag, f, f0)
        print("%s %s lag %d: %d pseudo-steps, %d factorisations, %d Krylov"
              % (which, opts, lag, counts[lag][0], counts[lag][1], sum(s["krylov"] for s in h)))
    steps0, ilu0 = counts[0]
    steps1, ilu1 = counts[lagk]
    assert ilu0 == steps0
    assert ilu1 < steps1
