"""Pins of the oracle by mathematics (closed forms, invariants, brute force,
finite differences, an independent library routine), chosen so that a dropped
term, a wrong sign or index, or a transposed operand fails one of them."""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import assembly, beam, contact

SEC = beam.circle_section(0.5, 200e3, 0.3)


def one_element(u12, X2=None):
    X2 = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, 0.05]]) if X2 is None else X2
    f, K = beam.cr_beam_element_batch(X2, np.array([[0, 1]]), u12, SEC)
    return f[0], K[0]


# ------------------------------------------------------------------ rotations


def test_exp_map_matches_scipy_rotation():
    rng = np.random.default_rng(0)
    th = rng.normal(0, 1.2, (200, 3))
    th[:3] = [[0, 0, 0], [1e-13, 0, 0], [0, 3.1, 0]]
    np.testing.assert_allclose(beam.rotvec_to_rotmat(th), Rotation.from_rotvec(th).as_matrix(), atol=1e-14)


def test_log_map_inverts_exp_map_across_shepperd_branches():
    rng = np.random.default_rng(1)
    ax = rng.normal(size=(400, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ang = rng.uniform(0, np.pi - 1e-6, 400)
    th = ax * ang[:, None]
    R = beam.rotvec_to_rotmat(th)
    tr = np.trace(R, axis1=1, axis2=2)
    assert (tr < np.max(np.stack([R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]], 1), axis=1)).any()  # non-trace branches used
    np.testing.assert_allclose(beam.rotmat_to_rotvec(R), th, atol=1e-9)


def test_minimal_rotation_maps_axis_and_fixes_normal():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(50, 3))
    b = rng.normal(size=(50, 3))
    a /= np.linalg.norm(a, axis=1, keepdims=True)
    b /= np.linalg.norm(b, axis=1, keepdims=True)
    R = beam.rodrigues_min_rotation(a, b)
    np.testing.assert_allclose(np.einsum("nij,nj->ni", R, a), b, atol=1e-13)
    w = np.cross(a, b)
    np.testing.assert_allclose(np.einsum("nij,nj->ni", R, w), w, atol=1e-13)
    np.testing.assert_allclose(R @ np.transpose(R, (0, 2, 1)), np.tile(np.eye(3), (50, 1, 1)), atol=1e-13)


def test_tangent_operator_and_inverse_are_inverse():
    rng = np.random.default_rng(3)
    th = rng.normal(0, 1.0, (100, 3))
    th[0] = 1e-12
    prod = beam.tangent_op(th) @ beam.tangent_op_inv(th)
    np.testing.assert_allclose(prod, np.tile(np.eye(3), (100, 1, 1)), atol=1e-12)


# ------------------------------------------------------------------ element


def test_zero_state_has_zero_force_and_linear_stiffness():
    f, K = one_element(np.zeros(12))
    assert np.abs(f).max() == 0.0
    X2 = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, 0.05]])
    Klin = beam.ke_global_linear(X2, SEC)
    np.testing.assert_allclose(K, Klin, atol=1e-9 * np.abs(Klin).max())


def test_rigid_translation_and_rotation_give_zero_force():
    X2 = np.array([[0.0, 0.0, 0.0], [0.3, 0.1, 0.05]])
    u = np.zeros(12)
    u[0:3] = u[6:9] = [0.01, -0.02, 0.03]
    f, _ = one_element(u, X2)
    assert np.abs(f).max() < 1e-6
    th = np.array([0.3, -0.5, 0.7])
    R = Rotation.from_rotvec(th).as_matrix()
    u = np.zeros(12)
    u[0:3] = R @ X2[0] - X2[0]
    u[6:9] = R @ X2[1] - X2[1]
    u[3:6] = u[9:12] = th
    f, _ = one_element(u, X2)
    # the corotational force must vanish; compare with the force of a 1e-4 bend
    ub = np.zeros(12)
    ub[9:12] = [0, 0, 1e-4]
    fb, _ = one_element(ub, X2)
    assert np.abs(f).max() < 1e-7 * np.abs(fb).max()


def test_axial_stretch_closed_form():
    X2 = np.array([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]])
    d = 1e-3
    u = np.zeros(12)
    u[6] = d
    f, _ = one_element(u, X2)
    P = SEC.E * SEC.A / 2.0 * d
    np.testing.assert_allclose(f[[0, 6]], [-P, P], rtol=1e-12)
    assert np.abs(np.delete(f, [0, 6])).max() < 1e-9 * P


def test_tangent_matches_central_difference_of_force():
    rng = np.random.default_rng(4)
    X2 = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0]])
    u = rng.normal(0, 1e-3, 12)
    _, K = one_element(u, X2)
    eps = 1e-7
    Kfd = np.zeros((12, 12))
    for j in range(12):
        up, um = u.copy(), u.copy()
        up[j] += eps
        um[j] -= eps
        Kfd[:, j] = (one_element(up, X2)[0] - one_element(um, X2)[0]) / (2 * eps)
    Kfd = 0.5 * (Kfd + Kfd.T)
    np.testing.assert_allclose(K, Kfd, rtol=1e-4, atol=1e-4 * np.abs(Kfd).max())


def test_tangent_is_symmetric():
    rng = np.random.default_rng(5)
    _, K = one_element(rng.normal(0, 0.05, 12))
    np.testing.assert_allclose(K, K.T, atol=1e-12 * np.abs(K).max())


def test_lumped_mass_closed_form():
    X = np.array([[0.0, 0, 0], [2.0, 0, 0], [5.0, 0, 0]])
    conn = np.array([[0, 1], [1, 2]])
    rho = 7.85e-9
    d = beam.lumped_mass_diag(X, conn, rho, SEC)
    m1, m2 = rho * SEC.A * 2.0, rho * SEC.A * 3.0
    np.testing.assert_allclose(d[6:9], (m1 + m2) / 2)
    np.testing.assert_allclose(d[9], rho * (SEC.Iy + SEC.Iz) * (2.0 + 3.0) / 2)
    np.testing.assert_allclose(d[10:12], m1 * 4 / 78 + m2 * 9 / 78)


# ------------------------------------------------------------------ broadphase


def random_segments(n, seed, span=10.0, length=1.0):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(0, span, (n, 3))
    d = rng.normal(size=(n, 3))
    x1 = x0 + length * d / np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.05, 0.3, n)
    return x0, x1, r


@pytest.mark.parametrize("cell", [None, 0.3, 1.0, 7.0])
def test_broadphase_equals_bruteforce_overlap_set_for_any_cell(cell):
    x0, x1, r = random_segments(300, 7)
    got = contact.broadphase_aabb(x0, x1, r, margin=0.05, cell_size=cell)
    ref = contact.broadphase_bruteforce(x0, x1, r, margin=0.05)
    assert np.array_equal(got, ref)
    assert len(ref) > 50


def test_broadphase_touching_boxes_are_inclusive():
    x0 = np.array([[0.0, 0, 0], [1.0, 0, 0]])
    x1 = np.array([[0.5, 0, 0], [2.0, 0, 0]])
    r = np.array([0.25, 0.25])  # hi_0 = 0.75 = lo_1 exactly
    assert contact.broadphase_aabb(x0, x1, r).tolist() == [[0, 1]]


# ------------------------------------------------------------------ narrowphase


def test_closest_points_match_dense_grid_minimum():
    rng = np.random.default_rng(8)
    n = 60
    A0, A1, B0, B1 = (rng.normal(size=(n, 3)) for _ in range(4))
    s, t, dist, nrm, par = contact.closest_point_segments(A0, A1, B0, B1)
    def grid_min(k, sl, tl):
        pa = A0[k] + sl[:, None] * (A1[k] - A0[k])
        pb = B0[k] + tl[:, None] * (B1[k] - B0[k])
        D = np.linalg.norm(pa[:, None] - pb[None], axis=2)
        i, j = np.unravel_index(np.argmin(D), D.shape)
        return D[i, j], sl[i], tl[j]

    for k in range(n):
        g = np.linspace(0, 1, 201)
        _, s0, t0 = grid_min(k, g, g)
        fs = np.clip(np.linspace(s0 - 0.01, s0 + 0.01, 401), 0, 1)
        ft = np.clip(np.linspace(t0 - 0.01, t0 + 0.01, 401), 0, 1)
        dmin, _, _ = grid_min(k, fs, ft)
        assert dist[k] <= dmin + 1e-12
        assert dist[k] >= dmin - 1e-6
        pk = A0[k] + s[k] * (A1[k] - A0[k]) - (B0[k] + t[k] * (B1[k] - B0[k]))
        np.testing.assert_allclose(np.linalg.norm(pk), dist[k], rtol=1e-12)
    assert 0.0 <= s.min() and s.max() <= 1.0 and 0.0 <= t.min() and t.max() <= 1.0


def test_parallel_segment_gaps_closed_form():
    # two parallel segments 2.0 apart with radii 0.5 -> gap 1.0; 0.5 apart and radii 0.5 -> gap -0.5
    for sep, r, want in [(2.0, 0.5, 1.0), (0.5, 0.5, -0.5)]:
        X = np.array([[0, 0, 0], [1, 0, 0], [0, sep, 0], [1, sep, 0]], float)
        conn = np.array([[0, 1], [2, 3]])
        p = contact.Pairs.fresh([0], [1], conn, [r, r], [r, r])
        p = contact.update_geometry(p, X, contact.ContactConfig())
        np.testing.assert_allclose(p.gap, [want], atol=1e-14)


def test_frame_orthonormal_and_transported_with_rigid_rotation():
    rng = np.random.default_rng(9)
    n0 = rng.normal(size=(40, 3))
    n0 /= np.linalg.norm(n0, axis=1, keepdims=True)
    z = np.zeros((40, 3))
    n, t1, t2 = contact.build_contact_frame(n0, z, z, np.zeros(40, bool), np.zeros(40, bool))
    for a, b in [(n, t1), (n, t2), (t1, t2)]:
        assert np.abs(np.einsum("ij,ij->i", a, b)).max() < 1e-14
    np.testing.assert_allclose(np.linalg.norm(t1, axis=1), 1.0, atol=1e-14)
    # a small rotation of the normal carries t1 by the minimal rotation
    R = Rotation.from_rotvec([0.0, 0.0, 0.02]).as_matrix()
    nr = n @ R.T
    _, t1r, _ = contact.build_contact_frame(nr, t1, n, np.ones(40, bool), np.ones(40, bool))
    Rmin = beam.rodrigues_min_rotation(n, nr)
    np.testing.assert_allclose(t1r, np.einsum("nij,nj->ni", Rmin, t1), atol=1e-12)


def test_hysteresis_truth_table():
    cfg = contact.ContactConfig(g_on=0.0, g_off=1e-6)
    X = np.array([[0, 0, 0], [1, 0, 0], [0.5, 0, -1], [0.5, 0, 1]], float)
    conn = np.array([[0, 1], [2, 3]])
    cases = [  # (separation along y, start status, expected)
        (1.0 - 1e-3, 0, 1), (1.0 + 1e-7, 0, 0), (1.0 + 5e-7, 1, 1), (1.0 + 2e-6, 1, 0), (1.0 + 2e-6, 2, 0)]
    for sep, st0, want in cases:
        Y = X.copy()
        Y[2:, 1] = sep
        p = contact.Pairs.fresh([0], [1], conn, [0.5, 0.5], [0.5, 0.5])
        p.status[:] = st0
        p = contact.update_geometry(p, Y, cfg)
        assert p.status[0] == want, (sep, st0)


# ------------------------------------------------------------------ Huber / friction


def test_huber_closed_form_and_derivative():
    d = 2.0
    x = np.array([-4.0, -2.0, 0.0, 2.0, 4.0])
    np.testing.assert_allclose(contact.huber(x, d), [0.0, 0.0, 0.5, 2.0, 4.0])
    np.testing.assert_allclose(contact.huber_deriv(x, d), [0.0, 0.0, 0.5, 1.0, 1.0])
    np.testing.assert_allclose(contact.huber(x, 0.0), np.maximum(0.0, x))
    h = 1e-6
    xs = np.linspace(-3, 3, 41)
    fd = (contact.huber(xs + h, d) - contact.huber(xs - h, d)) / (2 * h)
    np.testing.assert_allclose(fd, contact.huber_deriv(xs, d), atol=1e-6)


def crossing_pair(offset_y, rng=None):
    X = np.array([[0, 0, 0], [1, 0, 0], [0.4, offset_y, -0.5], [0.6, offset_y, 0.5]], float)
    conn = np.array([[0, 1], [2, 3]])
    p = contact.Pairs.fresh([0], [1], conn, [0.5, 0.5], [0.5, 0.5])
    return X, conn, p


def test_return_map_stick_slip_closed_forms():
    X, conn, p = crossing_pair(0.999)
    p = contact.update_geometry(p, X, contact.ContactConfig())
    k_pen, mu = 1e4, 0.3
    _, _, p = contact.huber_evaluate(p, k_pen, 0.0, 24)
    pn = p.p_n[0]
    assert pn == pytest.approx(k_pen * 1e-3, rel=1e-9)
    u_ref = np.zeros(24)
    for du, stick in [(1e-6, True), (1e-3, False)]:
        u = np.zeros(24)
        u[12:15] = du * p.tangent1[0]  # node B0 slides along t1
        u[18:21] = du * p.tangent1[0]
        fr = contact.friction_evaluate(p, u, u_ref, k_pen, 0.5, mu, 24)
        q = fr.pairs.z_t[0]
        kt = 0.5 * k_pen
        if stick:
            np.testing.assert_allclose(q, [kt * du, 0.0], atol=1e-12 * kt * du)
            np.testing.assert_allclose(fr.D_t[0], kt * np.eye(2))
            assert fr.pairs.status[0] == contact.ACTIVE
        else:
            np.testing.assert_allclose(np.linalg.norm(q), mu * pn, rtol=1e-12)
            assert q[0] > 0 and abs(q[1]) < 1e-12 * abs(q[0])
            ev = np.sort(np.linalg.eigvalsh(fr.D_t[0]))
            np.testing.assert_allclose(ev, [0.0, mu * pn / (kt * du) * kt], atol=1e-9 * kt)
            assert fr.pairs.status[0] == contact.SLIDING
        assert fr.pairs.dissipation[0] == pytest.approx(float(q @ [du, 0.0]), rel=1e-12)


def frozen_force(p, X, u, k_pen, q=None):
    """f_c (and f_fric for fixed q) at the pair's fixed (s, t), with the normal and
    frame following the closest-point chord (used for the FD checks)."""
    x = X + u.reshape(-1, 6)[:, :3]
    s, t = p.s[0], p.t[0]
    pa = (1 - s) * x[p.nodes_a[0, 0]] + s * x[p.nodes_a[0, 1]]
    pb = (1 - t) * x[p.nodes_b[0, 0]] + t * x[p.nodes_b[0, 1]]
    d = pa - pb
    dist = np.linalg.norm(d)
    n = d / dist
    gap = dist - (p.radius_a[0] + p.radius_b[0])
    c = np.array([1 - s, s, -(1 - t), -t])
    f = np.zeros(len(u))
    nodes = [p.nodes_a[0, 0], p.nodes_a[0, 1], p.nodes_b[0, 0], p.nodes_b[0, 1]]
    if q is None:
        pn = contact.huber(k_pen * (-gap), 0.0)
        vec = pn * n
    else:
        nn, t1, t2 = contact.build_contact_frame(n[None], p.tangent1, p.normal, np.ones(1, bool), np.ones(1, bool))
        vec = q[0] * t1[0] + q[1] * t2[0]
    for k in range(4):
        f[6 * nodes[k]:6 * nodes[k] + 3] += c[k] * vec
    return f


def test_huber_tangent_equals_fd_of_residual_at_fixed_closest_points():
    X, conn, p = crossing_pair(0.999)
    cfg = contact.ContactConfig()
    p = contact.update_geometry(p, X, cfg)
    k_pen = 1e4
    _, _, p = contact.huber_evaluate(p, k_pen, 0.0, 24)
    Kc = contact.huber_tangent(p, k_pen, 0.0, 24).toarray()
    u0 = np.zeros(24)
    h = 1e-8
    dofs = [0, 1, 2, 6, 7, 8, 12, 13, 14, 18, 19, 20]
    for j in dofs:
        up, um = u0.copy(), u0.copy()
        up[j] += h
        um[j] -= h
        col = -(frozen_force(p, X, up, k_pen) - frozen_force(p, X, um, k_pen)) / (2 * h)  # residual carries -f_c
        np.testing.assert_allclose(Kc[dofs, j], col[dofs], atol=1e-5 * np.abs(Kc).max())


def test_friction_geometric_tangent_equals_fd_at_fixed_q():
    X, conn, p = crossing_pair(0.999)
    cfg = contact.ContactConfig()
    p = contact.update_geometry(p, X, cfg)
    k_pen = 1e4
    _, _, p = contact.huber_evaluate(p, k_pen, 0.0, 24)
    u = np.zeros(24)
    u[12:15] = 1e-6 * p.tangent1[0]
    fr = contact.friction_evaluate(p, u, np.zeros(24), k_pen, 0.5, 0.3, 24)
    m, blocks = contact.friction_geometric_blocks(fr.pairs, fr)
    Kg = contact.blocks_to_csr(fr.pairs, m, blocks, 24).toarray()
    q = fr.q[0]
    h = 1e-8
    dofs = [0, 1, 2, 6, 7, 8, 12, 13, 14, 18, 19, 20]
    base = np.zeros(24)
    for j in dofs:
        up, um = base.copy(), base.copy()
        up[j] += h
        um[j] -= h
        col = (frozen_force(fr.pairs, X, up, k_pen, q) - frozen_force(fr.pairs, X, um, k_pen, q)) / (2 * h)
        np.testing.assert_allclose(Kg[dofs, j], col[dofs], atol=1e-5 * np.abs(Kg).max())


def test_friction_material_block_is_plus_D_g_gT():
    X, conn, p = crossing_pair(0.999)
    p = contact.update_geometry(p, X, contact.ContactConfig())
    _, _, p = contact.huber_evaluate(p, 1e4, 0.0, 24)
    u = np.zeros(24)
    u[12:15] = 1e-6 * p.tangent1[0]
    fr = contact.friction_evaluate(p, u, np.zeros(24), 1e4, 0.5, 0.3, 24)
    m, b = contact.friction_material_blocks(fr.pairs, fr)
    c = contact.shape_coeffs(fr.pairs)[0]
    g = [np.concatenate([ck * fr.pairs.tangent1[0] for ck in c]), np.concatenate([ck * fr.pairs.tangent2[0] for ck in c])]
    want = sum(fr.D_t[0][a, b_] * np.outer(g[a], g[b_]) for a in range(2) for b_ in range(2))
    got = b[0].transpose(0, 2, 1, 3).reshape(12, 12)
    np.testing.assert_allclose(got, want, atol=1e-12 * np.abs(want).max())


def test_generalized_alpha_chung_hulbert_parameters():
    ga = assembly.GeneralizedAlpha(np.ones(3), rho_inf=1.0)
    assert (ga.alpha_m, ga.alpha_f, ga.gamma, ga.beta) == (0.5, 0.5, 0.5, 0.25)
    ga = assembly.GeneralizedAlpha(np.ones(3), rho_inf=0.8)
    am, af = (2 * 0.8 - 1) / 1.8, 0.8 / 1.8
    assert ga.alpha_m == pytest.approx(am) and ga.alpha_f == pytest.approx(af)
    assert ga.gamma == pytest.approx(0.5 - am + af)
    assert ga.beta == pytest.approx(0.25 * (1 - am + af) ** 2)
