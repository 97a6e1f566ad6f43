"""The drop-in GeneralizedAlphaProcess (xkep-cae_amd/time_integration.py, state
on the GPU) with the consistent mass and a Rayleigh damping matrix against two
steps of the reference's (tests/golden/dynamics_7wire.npz), and the fused
assembly refusing a mass it cannot apply as a diagonal."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import golden, pkg

pytestmark = pytest.mark.gpu


def _csr(g, p):
    return sp.csr_matrix((g[f"{p}_data"], g[f"{p}_indices"], g[f"{p}_indptr"]), shape=tuple(g[f"{p}_shape"]))


def test_generalized_alpha_with_consistent_mass_and_damping(gpu):
    T = pkg("time_integration")
    g = golden("dynamics_7wire.npz")
    rho, rho_inf, dt, _, _ = g["params"]
    M, C, K0 = _csr(g, "M"), _csr(g, "C"), _csr(g, "K0")
    ga = T.GeneralizedAlphaProcess(M, damping_matrix=C, rho_inf=rho_inf)
    assert not ga.lumped
    u = np.zeros(M.shape[0])
    for k in range(2):
        up = ga.predict(u, dt)
        np.testing.assert_allclose(up, g[f"s{k}_u_pred"], rtol=0, atol=1e-15 * max(1.0, np.abs(up).max()))
        u = g[f"s{k}_u"]
        ga.correct(u, None, dt)
        np.testing.assert_allclose(ga.vel, g[f"s{k}_vel"], rtol=1e-13, atol=1e-12 * np.abs(g[f"s{k}_vel"]).max())
        Re = ga.effective_residual(g[f"s{k}_R"], dt)
        np.testing.assert_allclose(Re, g[f"s{k}_R_eff"], rtol=1e-12, atol=1e-12 * np.abs(Re).max())
        ref = _csr(g, f"s{k}_K_eff")
        assert abs(ga.effective_stiffness(K0, dt) - ref).max() <= 1e-13 * abs(ref).max()
    with pytest.raises(ValueError, match="lumped"):
        ga.dyn_struct(dt)
    lumped = T.GeneralizedAlphaProcess(sp.diags(M.diagonal()), rho_inf=rho_inf)
    assert lumped.lumped
    lumped.dyn_struct(dt)
