"""The oracle against the reference's own outputs (tests/golden/*.npz, written
by tests/golden/make_golden.py from the reference's classes)."""

import numpy as np
import scipy.sparse as sp

import pytest

from conftest import golden, pkg
from oracle import assembly, beam, contact


def csr(g, p):
    return sp.csr_matrix((g[p + "_data"], g[p + "_indices"], g[p + "_indptr"]), shape=tuple(g[p + "_shape"]))


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def srel(A, B):
    return sp.linalg.norm(A - B) / sp.linalg.norm(B)


def test_element_force_and_tangent_match_reference():
    g = golden("elem_small.npz")
    sec = beam.circle_section(float(g["radius"]), 200e3, 0.3)
    for k in range(3):
        K, f = beam.assemble_cr_beam3d(g["X"], g["conn"], g[f"u{k}"], sec)
        assert rel(f, g[f"f{k}"]) < 1e-13
        assert srel(K, csr(g, f"K{k}")) < 1e-13


def test_element_batch_matches_reference_scalar_routines():
    g = golden("elem_small.npz")
    sec = beam.circle_section(1.0, 200e3, 0.3)
    fe, Ke = beam.cr_beam_element_batch(g["X"], g["conn"][:6], g["u2"], sec)
    assert np.abs(fe - g["scalar_f"]).max() <= 1e-12 * np.abs(g["scalar_f"]).max()
    assert np.abs(Ke - g["scalar_K"]).max() <= 1e-12 * np.abs(g["scalar_K"]).max()


def test_lumped_mass_matches_reference():
    g = golden("elem_small.npz")
    sec = beam.circle_section(1.0, 200e3, 0.3)
    np.testing.assert_allclose(beam.lumped_mass_diag(g["X"], g["conn"], 7.85e-9, sec), g["M_lumped"],
                               rtol=1e-14, atol=0)


def pair_errors(p: contact.Pairs, a, pre):
    key = np.stack([np.minimum(p.elem_a, p.elem_b), np.maximum(p.elem_a, p.elem_b)], 1)
    o = np.lexsort((key[:, 1], key[:, 0]))
    assert np.array_equal(key[o], a[pre + "_key"])
    errs = {}
    for f in ["s", "t", "gap", "normal", "tangent1", "tangent2", "p_n", "z_t", "k_pen", "k_t", "q_trial_norm",
              "dissipation"]:
        v, r = getattr(p, f)[o], a[pre + "_" + f]
        errs[f] = np.abs(v - r).max() / max(np.abs(r).max(), 1e-300)
    assert np.array_equal(p.status[o], a[pre + "_status"])
    assert np.array_equal(p.stick[o], a[pre + "_stick"])
    return errs


def test_two_increments_of_newton_assembly_match_reference():
    a = golden("assembly_7wire.npz")
    k_pen, sd, mu, rho_inf, dt, rho, ktr = a["params"]
    X, conn = a["X"], a["conn"]
    sec = beam.circle_section(float(a["radius"]), 200e3, 0.3)
    radii = np.full(len(conn), float(a["radius"]))
    cfg = contact.ContactConfig(mu=mu)
    prob = assembly.Problem(X, conn, sec, k_pen, sd, ktr, mu, cfg)
    ga = assembly.GeneralizedAlpha(beam.lumped_mass_diag(X, conn, rho, sec), rho_inf)
    pairs, c0 = contact.detect_candidates(contact.Pairs(), X, conn, radii, cfg)
    assert np.array_equal(c0, a["cand0"])
    u = np.zeros(6 * len(X))
    u_ref = np.zeros_like(u)
    for inc in range(2):
        np.testing.assert_allclose(ga.predict(u, dt), a[f"inc{inc}_u_pred"], rtol=1e-14, atol=1e-300)
        base = a[f"inc{inc}_it0_u"]
        coords = X + base.reshape(-1, 6)[:, :3]
        pairs, cand = contact.detect_candidates(pairs, coords, conn, radii, cfg)
        assert np.array_equal(cand, a[f"inc{inc}_cand"])
        pairs = contact.update_geometry(pairs, coords, cfg)
        errs = pair_errors(pairs, a, f"inc{inc}_start")
        assert max(errs.values()) < 1e-13, errs
        for it in range(2):
            u = a[f"inc{inc}_it{it}_u"]
            res = assembly.assemble_iteration(prob, pairs, u, u_ref, a["f_ext"], a["fixed"], ga, dt)
            pairs = res.pairs
            pre = f"inc{inc}_it{it}"
            assert rel(res.R, a[pre + "_R"]) < 1e-13
            assert srel(res.K, csr(a, pre + "_K")) < 1e-13
            assert res.n_active == int(a[pre + "_n_active"])
            errs = pair_errors(pairs, a, pre + "_pairs")
            assert max(errs.values()) < 1e-13, errs
        u_ref = u.copy()


def _csr(g, p):
    import scipy.sparse as sp

    return sp.csr_matrix((g[f"{p}_data"], g[f"{p}_indices"], g[f"{p}_indptr"]), shape=tuple(g[f"{p}_shape"]))


def test_consistent_mass_matches_reference_and_closed_forms():
    """Oracle and product consistent mass against the reference's
    assemble_mass(lumped=False), automatic and explicit reference axis; total
    translational mass rho A sum(L) in every direction."""
    g = golden("dynamics_7wire.npz")
    rho = float(g["params"][0])
    sec = beam.circle_section(1.0, 200e3, 0.3)
    E = pkg("elements")
    for name, vref in (("M", None), ("Mv", np.array([0.0, 1.0, 0.0]))):
        ref = _csr(g, name)
        Mo = beam.consistent_mass(g["X"], g["conn"], rho, sec, v_ref=vref)
        assert abs(Mo - ref).max() <= 1e-15 * abs(ref).max()
        asm = E.ULCRBeamAssembler(g["X"], g["conn"], sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y,
                                  sec.kappa_z, v_ref=vref)
        Mp = asm.assemble_mass(rho, lumped=False)
        assert abs(Mp - ref).max() <= 1e-14 * abs(ref).max()
        X, conn = np.asarray(g["X"]), np.asarray(g["conn"])
        m_tot = rho * sec.A * np.linalg.norm(X[conn[:, 1]] - X[conn[:, 0]], axis=1).sum()
        for d in range(3):
            e = np.zeros(ref.shape[0])
            e[d::6] = 1.0
            assert e @ (Mp @ e) == pytest.approx(m_tot, rel=1e-12)


def test_generalized_alpha_with_damping_matches_reference():
    """The oracle's Generalized-alpha with the consistent mass and a Rayleigh
    damping matrix reproduces two steps of the reference's (predict, correct,
    effective residual and stiffness)."""
    g = golden("dynamics_7wire.npz")
    rho, rho_inf, dt, _, _ = g["params"]
    M, C, K0 = _csr(g, "M"), _csr(g, "C"), _csr(g, "K0")
    ga = assembly.GeneralizedAlpha(M, rho_inf, damping_matrix=C)
    u = np.zeros(M.shape[0])
    for k in range(2):
        up = ga.predict(u, dt)
        np.testing.assert_allclose(up, g[f"s{k}_u_pred"], rtol=0, atol=1e-15 * max(1.0, np.abs(up).max()))
        u = g[f"s{k}_u"]
        ga.correct(u, None, dt)
        np.testing.assert_allclose(ga.vel, g[f"s{k}_vel"], rtol=1e-13, atol=1e-12 * np.abs(g[f"s{k}_vel"]).max())
        Re = ga.effective_residual(g[f"s{k}_R"], dt)
        np.testing.assert_allclose(Re, g[f"s{k}_R_eff"], rtol=1e-12, atol=1e-12 * np.abs(Re).max())
        Ke = ga.effective_stiffness(K0, dt)
        ref = _csr(g, f"s{k}_K_eff")
        assert abs(Ke - ref).max() <= 1e-13 * abs(ref).max()
