"""The oracle's ContactGeometryStrategy twin (oracle/contact.py:
strategy_detect, strategy_update_geometry, constraint_jacobian_ptp) pinned
against the reference's own Strategy classes (tests/golden/strategy_7wire.npz,
written by running contact/geometry/strategy.py) and against what the
mathematics fixes: the constraint Jacobian is the derivative of the gaps
(central finite differences of the narrowphase), so a dropped coefficient, a
wrong sign or a swapped node fails."""

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import golden
from oracle import contact as ocon

FIELDS = ["s", "t", "gap", "normal", "tangent1", "tangent2"]


def keyed(p: ocon.Pairs):
    key = np.stack([np.minimum(p.elem_a, p.elem_b), np.maximum(p.elem_a, p.elem_b)], axis=1)
    order = np.lexsort((key[:, 1], key[:, 0]))
    return key[order], order


def check_pairs(g, prefix, p: ocon.Pairs, tol=1e-12):
    key, order = keyed(p)
    np.testing.assert_array_equal(key, g[f"{prefix}_key"])
    for f in FIELDS:
        np.testing.assert_allclose(getattr(p, f)[order], g[f"{prefix}_{f}"], rtol=0, atol=tol, err_msg=f)
    np.testing.assert_array_equal(p.status[order], g[f"{prefix}_status"])


def golden_rows(g, prefix):
    G = sp.csr_matrix((g[f"{prefix}_G_data"], g[f"{prefix}_G_indices"], g[f"{prefix}_G_indptr"]),
                      shape=tuple(g[f"{prefix}_G_shape"]))
    return {tuple(k): G.getrow(r) for r, k in enumerate(g[f"{prefix}_G_keys"].tolist())}


def check_jacobian(g, prefix, p: ocon.Pairs, ndof, tol=1e-15):
    G, act = ocon.constraint_jacobian_ptp(p, ndof, 6)
    ref = golden_rows(g, prefix)
    assert G.shape == (len(ref), ndof)
    assert act == sorted(act) and all(p.status[i] != ocon.INACTIVE for i in act)
    for r, i in enumerate(act):
        k = (min(p.elem_a[i], p.elem_b[i]), max(p.elem_a[i], p.elem_b[i]))
        a, b = G.getrow(r), ref[k]
        np.testing.assert_array_equal(a.indices, b.indices)
        np.testing.assert_allclose(a.data, b.data, rtol=0, atol=tol)


@pytest.fixture(scope="module")
def g():
    return golden("strategy_7wire.npz")


def test_detect_matches_reference(g):
    p = ocon.strategy_detect(g["x0"], g["conn"], g["radii"])
    check_pairs(g, "det", p)
    assert np.all(p.core_radius_a == 0.0) and np.all(p.core_radius_b == 0.0)
    check_jacobian(g, "det", p, 6 * len(g["X"]))


def test_detect_without_shared_node_filter_matches_reference(g):
    p = ocon.strategy_detect(g["x0"], g["conn"], g["radii"], exclude_same_layer=False)
    check_pairs(g, "all", p)
    check_jacobian(g, "all", p, 6 * len(g["X"]))
    # adjacent elements share a node: their rows hold 9 summed entries, not 12
    conn = np.asarray(g["conn"])
    share = np.array([len(set(conn[a]) & set(conn[b])) > 0 for a, b in zip(p.elem_a, p.elem_b)])
    assert share.any()


def test_update_geometry_with_and_without_config_matches_reference(g):
    p = ocon.strategy_detect(g["x0"], g["conn"], g["radii"])
    p = ocon.strategy_update_geometry(p, g["x1"], ocon.ContactConfig(g_on=0.0, g_off=1e-6))
    check_pairs(g, "upd", p)
    check_jacobian(g, "upd", p, 6 * len(g["X"]))
    p = ocon.strategy_update_geometry(p, g["x0"], None)
    check_pairs(g, "upd2", p)
    check_jacobian(g, "upd2", p, 6 * len(g["X"]))
    p = ocon.strategy_update_geometry(p, g["x1"], ocon.ContactConfig(g_on=0.0, g_off=1e-6, coating_stiffness=1e3))
    check_pairs(g, "coat", p)
    key, order = keyed(p)
    np.testing.assert_allclose(p.coating_compression[order], g["coat_compression"], rtol=0, atol=1e-12)


def test_jacobian_is_the_derivative_of_the_gaps(g):
    """Row of pair i = d gap_i / d x (translations of its four nodes), by
    central differences of the narrowphase on the pair's own endpoints."""
    x = np.asarray(g["x0"])
    p = ocon.strategy_detect(x, g["conn"], g["radii"])
    ndof = 6 * len(x)
    G, act = ocon.constraint_jacobian_ptp(p, ndof, 6)
    ends = [x[p.nodes_a[:, 0]], x[p.nodes_a[:, 1]], x[p.nodes_b[:, 0]], x[p.nodes_b[:, 1]]]
    _, _, _, _, parallel = ocon.closest_point_segments(*ends)
    h = 1e-6
    rad = p.radius_a + p.radius_b
    checked = 0
    for r, i in enumerate(act):
        if parallel[i]:
            continue
        row = G.getrow(r).toarray().ravel()
        nodes = [p.nodes_a[i, 0], p.nodes_a[i, 1], p.nodes_b[i, 0], p.nodes_b[i, 1]]
        for k in range(4):
            for d in range(3):
                e = [q[i:i + 1].copy() for q in ends]
                e[k][0, d] += h
                gp = ocon.closest_point_segments(*e)[2][0] - rad[i]
                e[k][0, d] -= 2 * h
                gm = ocon.closest_point_segments(*e)[2][0] - rad[i]
                fd = (gp - gm) / (2 * h)
                assert row[6 * nodes[k] + d] == pytest.approx(fd, rel=1e-5, abs=1e-7), (i, k, d)
        checked += 1
    assert checked > 100


def test_jacobian_drops_tiny_entries_and_skips_inactive():
    """Entries with |c_k n_d| <= 1e-30 are not stored; INACTIVE pairs give no row."""
    conn = np.array([[0, 1], [2, 3], [4, 5]])
    x = np.array([[0, 0, 0], [1, 0, 0], [0.5, 1.0, 0], [0.5, 1.0, 1.0], [5, 5, 5], [6, 5, 5]], float)
    p = ocon.Pairs.fresh([0, 0], [1, 2], conn, np.full(3, 0.6), np.zeros(3))
    p = ocon.strategy_update_geometry(p, x, None)
    assert list(p.status) == [ocon.ACTIVE, ocon.INACTIVE]
    G, act = ocon.constraint_jacobian_ptp(p, 36, 6)
    assert act == [0] and G.shape == (1, 36)
    # pair (0, 1): s = 0.5, t = 0, n = (0, -1, 0): only the y DOFs of nodes 0, 1, 2 are stored
    np.testing.assert_array_equal(G.indices, [1, 7, 13])
    np.testing.assert_allclose(G.data, [-0.5, -0.5, 1.0])
