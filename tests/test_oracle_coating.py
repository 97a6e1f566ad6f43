"""The oracle's Kelvin-Voigt coating terms (oracle/coating.py) pinned by
  * the known answers of the reference's own coating tests
    (xkep_cae/contact/coating/tests/test_physics.py: one pair A0-A1 / B0-B1 at
    y = 0 / 0.5, s = t = 0.5, n = (0, 1, 0), t1 = (1, 0, 0), t2 = (0, 0, 1),
    k = 1e6): linearity in the compression, action-reaction, the signs of the
    A and B forces along n, the dashpot's rate proportionality and its
    absence at constant compression, k_eff = k + c/dt, symmetry and positive
    semi-definiteness of the stiffness, the Coulomb limit and history reset of
    the coating friction, and the elastic-only normal force of that friction;
  * closed forms the formulas fix: node forces w_k f_n n with w = [-(1-s),
    -s, 1-t, t] at arbitrary s, t; the stiffness is the derivative of the
    coating force (material part) along a normal approach where the frame
    does not turn (finite differences through update_geometry);
  * the geometry branch: core gap, compression max(0, coat_total - gap_core)
    and activation/deactivation by the compression."""

import numpy as np
import pytest

from oracle import coating as oc
from oracle import contact as ocon

COORDS = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.5, 0.0], [1.0, 0.5, 0.0]])
NDOF = 24


def cfg(k=1e6, c=0.0, mu=0.0, ktr=0.5):
    return ocon.ContactConfig(coating_stiffness=k, coating_damping=c, coating_mu=mu, coating_k_t_ratio=ktr)


def one_pair(cc=0.0, ccp=0.0, s=0.5, t=0.5, n=(0.0, 1.0, 0.0), t1=(1.0, 0.0, 0.0), t2=(0.0, 0.0, 1.0),
             z=(0.0, 0.0), stick=True, qtn=0.0):
    p = ocon.Pairs.fresh([0], [1], np.array([[0, 1], [2, 3]]), np.array([0.3, 0.3]), np.array([0.3, 0.3]))
    p.s[:] = s
    p.t[:] = t
    p.normal[:] = n
    p.tangent1[:] = t1
    p.tangent2[:] = t2
    p.coating_compression[:] = cc
    p.coating_compression_prev[:] = ccp
    p.coating_z_t[:] = z
    p.coating_stick[:] = stick
    p.coating_q_trial_norm[:] = qtn
    return p


def K_of(p, c, dt=0.01):
    m, b = oc.stiffness_blocks(p, c, dt)
    return ocon.blocks_to_csr(p, m, b, NDOF)


# ---------------------------------------------------------------- the reference tests' known answers
def test_no_compression_no_force():
    assert np.allclose(oc.forces(one_pair(0.0), cfg(), 0.01, NDOF), 0.0)


def test_force_proportional_to_compression():
    f1 = oc.forces(one_pair(0.01), cfg(), 0.01, NDOF)
    f2 = oc.forces(one_pair(0.02), cfg(), 0.01, NDOF)
    nz = np.abs(f1) > 1e-30
    assert nz.any()
    np.testing.assert_allclose(f2[nz] / f1[nz], 2.0, rtol=1e-10)


def test_action_reaction_and_direction():
    f = oc.forces(one_pair(0.01), cfg(), 0.01, NDOF).reshape(4, 6)
    fA, fB = f[0, :3] + f[1, :3], f[2, :3] + f[3, :3]
    np.testing.assert_allclose(fA + fB, 0.0, atol=1e-12)
    n = np.array([0.0, 1.0, 0.0])
    assert fB @ n > 0 and fA @ n < 0


def test_dashpot():
    dt = 0.01
    p = one_pair(0.02, 0.01)
    assert np.linalg.norm(oc.forces(p, cfg(c=1e4), dt, NDOF)) > np.linalg.norm(oc.forces(p, cfg(), dt, NDOF))
    fast, slow = one_pair(0.01, 0.0), one_pair(0.01, 0.005)
    vf = np.linalg.norm(oc.forces(fast, cfg(c=1e4), dt, NDOF) - oc.forces(fast, cfg(), dt, NDOF))
    vs = np.linalg.norm(oc.forces(slow, cfg(c=1e4), dt, NDOF) - oc.forces(slow, cfg(), dt, NDOF))
    assert vf > 0
    np.testing.assert_allclose(vf / vs, 2.0, rtol=0.01)
    eq = one_pair(0.01, 0.01)
    np.testing.assert_allclose(oc.forces(eq, cfg(), dt, NDOF), oc.forces(eq, cfg(c=1e4), dt, NDOF), atol=1e-12)


def test_stiffness_known_answers():
    k, c, dt = 1e6, 1e4, 0.01
    p = one_pair(0.01)
    Ke, Kv = K_of(p, cfg(k), dt).toarray(), K_of(p, cfg(k, c), dt).toarray()
    np.testing.assert_allclose(np.abs(Kv).sum() / np.abs(Ke).sum(), 1.0 + c / (k * dt), rtol=1e-10)
    np.testing.assert_allclose(Kv, Kv.T, atol=1e-10)
    assert np.linalg.eigvalsh(Kv).min() >= -1e-8
    assert K_of(one_pair(0.0), cfg(), dt).nnz == 0


def test_friction_known_answers():
    u0 = np.random.RandomState(42).randn(NDOF) * 0.001
    assert np.allclose(oc.friction_forces(one_pair(0.0), cfg(mu=0.3), u0, np.zeros(NDOF), NDOF)[0], 0.0)
    assert np.allclose(oc.friction_forces(one_pair(0.01), cfg(mu=0.0), u0, np.zeros(NDOF), NDOF)[0], 0.0)
    u = np.zeros(NDOF)
    u[0] = u[6] = -0.1
    f1 = oc.friction_forces(one_pair(0.01), cfg(mu=0.3), u, np.zeros(NDOF), NDOF)[0]
    f2 = oc.friction_forces(one_pair(0.01), cfg(c=1e6, mu=0.3), u, np.zeros(NDOF), NDOF)[0]
    np.testing.assert_allclose(f1, f2, atol=1e-10)
    u[0] = u[6] = -1.0
    f = oc.friction_forces(one_pair(0.01), cfg(mu=0.3), u, np.zeros(NDOF), NDOF)[0].reshape(4, 6)
    assert np.linalg.norm(f[0, :3] + f[1, :3]) <= 0.3 * 1e6 * 0.01 * (1.0 + 1e-8)
    sep = one_pair(0.0, z=(100.0, 200.0), stick=False)
    _, p = oc.friction_forces(sep, cfg(mu=0.3), np.zeros(NDOF), np.zeros(NDOF), NDOF)
    np.testing.assert_allclose(p.coating_z_t, 0.0, atol=1e-15)
    assert bool(p.coating_stick[0])


def test_friction_stiffness_known_answers():
    p = one_pair(0.01, z=(10.0, 20.0), stick=True, qtn=22.36)
    m, b = oc.friction_stiffness_blocks(p, cfg(mu=0.3))
    K = ocon.blocks_to_csr(p, m, b, NDOF).toarray()
    np.testing.assert_allclose(K, K.T, atol=1e-10)
    m, b = oc.friction_stiffness_blocks(one_pair(0.0), cfg(mu=0.3))
    assert ocon.blocks_to_csr(one_pair(0.0), m, b, NDOF).nnz == 0


# ---------------------------------------------------------------- closed forms
def test_node_forces_closed_form():
    s, t, k, c, dt, cc, ccp = 0.3, 0.8, 2e5, 3e3, 0.02, 0.004, 0.001
    n = np.array([0.6, 0.0, 0.8])
    f = oc.forces(one_pair(cc, ccp, s=s, t=t, n=n), cfg(k, c), dt, NDOF).reshape(4, 6)[:, :3]
    f_n = k * cc + c * (cc - ccp) / dt
    w = np.array([-(1 - s), -s, 1 - t, t])
    np.testing.assert_allclose(f, w[:, None] * f_n * n[None, :], rtol=1e-14, atol=1e-12)
    # stick return mapping: q = k_t du_t, node forces w_k q_a t_a
    u = np.zeros(NDOF)
    u[12:15] = [1e-6, 0.0, 2e-6]     # B0 moves in the tangent plane of n = (0, 1, 0)
    p = one_pair(cc, s=s, t=t)
    ff, p2 = oc.friction_forces(p, cfg(k, mu=0.5, ktr=0.5), u, np.zeros(NDOF), NDOF)
    kt = 0.5 * k
    q = kt * (1 - t) * np.array([1e-6, 2e-6])
    np.testing.assert_allclose(p2.coating_z_t[0], q, rtol=1e-14)
    assert bool(p2.coating_stick[0])
    ft = q[0] * np.array([1.0, 0, 0]) + q[1] * np.array([0, 0, 1.0])
    np.testing.assert_allclose(ff.reshape(4, 6)[:, :3], w[:, None] * ft[None, :], rtol=1e-14, atol=1e-20)
    # slip: |q| = mu k cc along the trial direction
    u[12:15] = [1.0, 0.0, 0.0]
    _, p3 = oc.friction_forces(p, cfg(k, mu=0.5, ktr=0.5), u, np.zeros(NDOF), NDOF)
    np.testing.assert_allclose(np.linalg.norm(p3.coating_z_t[0]), 0.5 * k * cc, rtol=1e-14)
    assert not bool(p3.coating_stick[0])


@pytest.mark.parametrize("core", [0.25, 0.3])
def test_stiffness_is_the_force_derivative_along_the_normal(core):
    """Two parallel segments along x (A at y = 0, B at y = h), core radii
    `core`, contact radii 0.3: moving B along y keeps n, s, t fixed, so the
    coating force changes only through the compression and K_coat must equal
    its derivative (the dashpot enters as k_eff = k + c/dt with cc_prev fixed)."""
    k, c, dt = 1e5, 50.0, 0.01
    cf = cfg(k, c)
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.5, 0.0], [1.0, 0.5, 0.0]])
    conn = np.array([[0, 1], [2, 3]])

    def state(h):
        x = X.copy()
        x[2:, 1] = h
        p = ocon.Pairs.fresh([0], [1], conn, np.array([0.3, 0.3]), np.array([core, core]))
        p = ocon.update_geometry(p, x, cf)
        p.coating_compression_prev[:] = 0.0
        return p

    h = 0.5    # core gap 0.5 - 2 core: compressed for core = 0.3 (penetration) and 0.25 (coating)
    p = state(h)
    assert p.coating_compression[0] > 0.0
    K = K_of(p, cf, dt).toarray()
    e = 1e-7
    fp = oc.forces(state(h + e), cf, dt, NDOF)
    fm = oc.forces(state(h - e), cf, dt, NDOF)
    dfdB = (fp - fm) / (2 * e)          # d f / d(y of B0 and B1 together)
    col = K[:, 6 * 2 + 1] + K[:, 6 * 3 + 1]
    np.testing.assert_allclose(col, dfdB, rtol=1e-6, atol=1e-6 * np.abs(col).max())


def test_geometry_branch_and_activation():
    X = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.55, 0.0], [1.0, 0.55, 0.0]])
    conn = np.array([[0, 1], [2, 3]])
    cf = cfg(1e5)
    cf.g_on, cf.g_off = 0.0, 1e-6
    p = ocon.Pairs.fresh([0], [1], conn, np.array([0.3, 0.3]), np.array([0.25, 0.25]))
    p = ocon.update_geometry(p, X, cf)
    np.testing.assert_allclose(p.gap, 0.55 - 0.5, rtol=1e-14)                       # core gap
    np.testing.assert_allclose(p.coating_compression, 0.1 - 0.05, rtol=1e-12)        # 2 x 0.05 coat - 0.05
    assert p.status[0] == ocon.ACTIVE                                                 # activated by the coating
    X2 = X.copy()
    X2[2:, 1] = 0.7
    p2 = ocon.update_geometry(p, X2, cf)
    assert p2.coating_compression[0] == 0.0 and p2.status[0] == ocon.INACTIVE      # released and gap >= g_off
    p3 = ocon.update_geometry(p, X, ocon.ContactConfig(g_on=0.0, g_off=1e-6))     # coating off: plain gap
    np.testing.assert_allclose(p3.gap, 0.55 - 0.6, rtol=1e-13)


def test_no_coating_strategy_returns_zeros():
    """_create_coating_strategy picks NoCoatingProcess at zero stiffness; its
    terms are zeros of the reference's shapes (coating/strategy.py:19-77)."""
    from conftest import pkg
    K = pkg("contact")
    nc = K._create_coating_strategy(coating_stiffness=0.0)
    assert isinstance(nc, K.NoCoatingProcess)
    assert isinstance(K._create_coating_strategy(coating_stiffness=1.0), K.KelvinVoigtCoatingProcess)
    cfg = K.ContactConfig()
    x = np.zeros((5, 3))
    assert np.array_equal(nc.forces([], x, cfg, 0.1), np.zeros(30))
    assert nc.stiffness([], x, cfg, 30, 0.1).shape == (30, 30) and nc.stiffness([], x, cfg, 30, 0.1).nnz == 0
    u = np.ones(30)
    assert np.array_equal(nc.friction_forces([], x, cfg, u, u), np.zeros(30))
    assert nc.friction_stiffness([], x, cfg, 30).nnz == 0
    assert nc.process(None) is None


@pytest.mark.parametrize("regime", ["stick", "slip"])
def test_friction_stiffness_is_the_derivative_of_the_friction_force(regime):
    """At a fixed frame (s, t, t1, t2) and a fixed history before the
    evaluation, the coating friction force is a function of u alone; its
    Jacobian over the 12 translational DOFs of the pair's nodes, by central
    differences, equals the coating friction stiffness (w_k w_l) T D T^T with D
    from the evaluated state -- k_t I in stick, (mu p_n / |q_trial|) k_t
    (I - q^ q^T) in slip.  A wrong sign, frame or weight fails this."""
    s, t = 0.3, 0.7
    z = (100.0, -50.0) if regime == "stick" else (3600.0, 900.0)
    p0 = one_pair(0.01, s=s, t=t, z=z, stick=True)
    c = cfg(mu=0.3)                                   # k_t = 5e5, mu p_n = 3000
    rng = np.random.default_rng(5 if regime == "stick" else 6)
    u = np.zeros(NDOF)
    dofs = (np.arange(4)[:, None] * 6 + np.arange(3)[None, :]).ravel()
    u[dofs] = rng.normal(0.0, 4e-4, 12)
    f, p = oc.friction_forces(p0, c, u, np.zeros(NDOF), NDOF)
    assert bool(p.coating_stick[0]) == (regime == "stick")
    m, b = oc.friction_stiffness_blocks(p, c)
    K = ocon.blocks_to_csr(p, m, b, NDOF).toarray()[np.ix_(dofs, dofs)]
    h = 1e-8
    J = np.zeros((12, 12))
    for j, d in enumerate(dofs):
        up, um = u.copy(), u.copy()
        up[d] += h
        um[d] -= h
        fp = oc.friction_forces(p0, c, up, np.zeros(NDOF), NDOF)[0]
        fm = oc.friction_forces(p0, c, um, np.zeros(NDOF), NDOF)[0]
        J[:, j] = (fp - fm)[dofs] / (2 * h)
    np.testing.assert_allclose(K, J, rtol=1e-6, atol=1e-6 * np.abs(K).max())
    assert np.abs(K).max() > 1e3
