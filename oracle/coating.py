"""Oracle: Kelvin-Voigt coating pair terms, FP64 NumPy.

Test infrastructure only (see ``oracle/__init__.py``).

Restates ``KelvinVoigtCoatingProcess`` (contact/coating/strategy.py:80-384),
vectorised over the structure-of-arrays pair state of ``contact.Pairs``:

  forces               :99-144   f_n = k cc + c (cc - cc_prev) / dt, node
                                 forces -f_n n (A side, weights 1-s, s) and
                                 +f_n n (B side, weights 1-t, t); every pair
                                 with cc > 0 or cc_prev > 0, status ignored,
                                 skipped when |f_n| < 1e-30
  stiffness            :146-192  k_eff g g^T, g = [w_k n], w = [-(1-s), -s, 1-t, t],
                                 k_eff = k + c / dt, pairs with cc > 1e-15
  friction_forces      :194-286  Coulomb return mapping of the coating
                                 history (k_t = k k_t_ratio, p_n = k cc, elastic
                                 part only) with du = u - u_ref at the closest
                                 points; state reset for cc <= 0
  friction_stiffness   :288-370  w_k w_l T D T^T, T = [t1 t2], D from
                                 ``_tangent_2x2_core`` (friction/law_friction.py:126-153)

Node blocks are returned as (M, 4, 4, 3, 3) arrays over the nodes
(A0, A1, B0, B1) like ``contact.huber_tangent_blocks``; the reference drops
entries with |val| <= 1e-30 from its COO lists, which ``blocks_to_csr``
repeats.  The node forces are summed in pair order (``np.add.at``) as the
reference's loop does.
"""

from __future__ import annotations

import numpy as np

from . import contact

NODE_WEIGHT_SIGNS = np.array([-1.0, -1.0, 1.0, 1.0])


def coating_weights(p: contact.Pairs) -> np.ndarray:
    """w = [-(1-s), -s, 1-t, t] per pair (strategy.py:178)."""
    return np.stack([-(1.0 - p.s), -p.s, 1.0 - p.t, p.t], axis=1)


def _scatter(f: np.ndarray, p: contact.Pairs, m: np.ndarray, vec: np.ndarray) -> None:
    """Node forces of pairs m: A0 += (1-s)(-vec), A1 += s(-vec), B0 += (1-t) vec, B1 += t vec."""
    if not m.any():
        return
    w = np.stack([(1.0 - p.s[m]), p.s[m], (1.0 - p.t[m]), p.t[m]], axis=1)   # (M,4)
    sg = np.array([-1.0, -1.0, 1.0, 1.0])
    vals = w[:, :, None] * (sg[None, :, None] * vec[:, None, :])
    np.add.at(f, contact.contact_dofs(p)[m].ravel(), vals.ravel())


def forces(p: contact.Pairs, cfg: contact.ContactConfig, dt: float, ndof: int) -> np.ndarray:
    """``KelvinVoigtCoatingProcess.forces`` (strategy.py:99-144)."""
    f = np.zeros(ndof)
    if len(p) == 0:
        return f
    k, c = cfg.coating_stiffness, cfg.coating_damping
    cc, ccp = p.coating_compression, p.coating_compression_prev
    sel = ~((cc <= 0.0) & (ccp <= 0.0))
    f_n = k * cc
    if c > 0.0 and dt > 0.0:
        f_n = f_n + c * ((cc - ccp) / dt)
    m = sel & ~(np.abs(f_n) < 1e-30)
    _scatter(f, p, m, f_n[m, None] * p.normal[m])
    return f


def effective_stiffness(cfg: contact.ContactConfig, dt: float) -> float:
    """k_eff = k + c / dt (backward-Euler dashpot, strategy.py:156-159)."""
    k_eff = cfg.coating_stiffness
    if cfg.coating_damping > 0.0 and dt > 0.0:
        k_eff += cfg.coating_damping / dt
    return k_eff


def stiffness_blocks(p: contact.Pairs, cfg: contact.ContactConfig, dt: float):
    """``KelvinVoigtCoatingProcess.stiffness`` (strategy.py:146-192) as node
    blocks k_eff (w_k n)(w_l n)^T of the pairs with cc > 1e-15."""
    m = p.coating_compression > 1e-15
    k_eff = effective_stiffness(cfg, dt)
    w = coating_weights(p)[m]
    n = p.normal[m]
    g = w[:, :, None] * n[:, None, :]                      # (M,4,3) = g_n per node
    val = k_eff * (g[:, :, None, :, None] * g[:, None, :, None, :])
    val[np.abs(val) <= 1e-30] = 0.0
    return m, val


def friction_forces(p: contact.Pairs, cfg: contact.ContactConfig, u, u_ref, ndof: int):
    """``KelvinVoigtCoatingProcess.friction_forces`` (strategy.py:194-286).
    Returns (f, pairs with the coating history updated)."""
    p = p.copy()
    f = np.zeros(ndof)
    mu = cfg.coating_mu
    if mu <= 0.0 or len(p) == 0:
        return f, p
    k = cfg.coating_stiffness
    k_t = k * cfg.coating_k_t_ratio
    cc = p.coating_compression
    off = cc <= 0.0
    p.coating_z_t = np.where(off[:, None], 0.0, p.coating_z_t)
    p.coating_stick = np.where(off, True, p.coating_stick)
    p.coating_q_trial_norm = np.where(off, 0.0, p.coating_q_trial_norm)
    p.coating_dissipation = np.where(off, 0.0, p.coating_dissipation)
    on = ~off
    if not on.any():
        return f, p
    p_n = k * cc
    du = np.asarray(u, float) - np.asarray(u_ref, float)
    dn = lambda nodes: du[(nodes[:, None] * 6 + np.arange(3)[None, :])]
    s, t = p.s[:, None], p.t[:, None]
    duA = (1.0 - s) * dn(p.nodes_a[:, 0]) + s * dn(p.nodes_a[:, 1])
    duB = (1.0 - t) * dn(p.nodes_b[:, 0]) + t * dn(p.nodes_b[:, 1])
    dur = duB - duA
    dut = np.stack([np.einsum("ij,ij->i", dur, p.tangent1), np.einsum("ij,ij->i", dur, p.tangent2)], axis=1)
    # _return_mapping_core (law_friction.py:89-123)
    qt = p.coating_z_t + k_t * dut
    qtn = np.linalg.norm(qt, axis=1)
    fy = mu * p_n
    stick = qtn <= fy
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(stick[:, None], qt, fy[:, None] * qt / qtn[:, None])
    diss = np.einsum("ij,ij->i", q, dut)
    p.coating_z_t = np.where(on[:, None], q, p.coating_z_t)
    p.coating_stick = np.where(on, stick, p.coating_stick)
    p.coating_q_trial_norm = np.where(on, qtn, p.coating_q_trial_norm)
    p.coating_dissipation = np.where(on, diss, p.coating_dissipation)
    f_t = q[on, 0:1] * p.tangent1[on] + q[on, 1:2] * p.tangent2[on]
    _scatter(f, p, on, f_t)
    return f, p


def _tangent_2x2(k_t: float, p_n, mu: float, z_t, q_trial_norm, stick):
    """``_tangent_2x2_core`` (law_friction.py:126-153), vectorised."""
    M = len(p_n)
    D = np.zeros((M, 2, 2))
    eye = np.eye(2)[None]
    ok = (p_n > 0.0) & (mu > 0.0)
    zn = np.linalg.norm(z_t, axis=1)
    el = ok & (stick | (zn < 1e-30) | (q_trial_norm < 1e-30))
    D[el] = k_t * eye
    sl = ok & ~el
    if sl.any():
        qh = z_t[sl] / zn[sl, None]
        ratio = (mu * p_n[sl]) / q_trial_norm[sl]
        D[sl] = (ratio * k_t)[:, None, None] * (eye - qh[:, :, None] * qh[:, None, :])
    return D


def friction_stiffness_blocks(p: contact.Pairs, cfg: contact.ContactConfig):
    """``KelvinVoigtCoatingProcess.friction_stiffness`` (strategy.py:288-370) as
    node blocks (w_k w_l) T D T^T over the pairs with cc > 0 (the history as
    ``friction_forces`` left it)."""
    mu = cfg.coating_mu
    if mu <= 0.0 or len(p) == 0:
        return np.zeros(len(p), bool), np.zeros((0, 4, 4, 3, 3))
    k = cfg.coating_stiffness
    k_t = k * cfg.coating_k_t_ratio
    m = p.coating_compression > 0.0
    D = _tangent_2x2(k_t, k * p.coating_compression[m], mu, p.coating_z_t[m], p.coating_q_trial_norm[m],
                     p.coating_stick[m])
    T = np.stack([p.tangent1[m], p.tangent2[m]], axis=2)        # (M,3,2)
    K3 = np.einsum("mia,mab,mjb->mij", T, D, T)
    w = coating_weights(p)[m]
    val = (w[:, :, None] * w[:, None, :])[:, :, :, None, None] * K3[:, None, None]
    val[np.abs(val) <= 1e-30] = 0.0
    return m, val


def commit(p: contact.Pairs) -> contact.Pairs:
    """End of a converged step: coating_compression_prev <- coating_compression
    (contact/solver/process.py:539-547)."""
    p = p.copy()
    p.coating_compression_prev = p.coating_compression.copy()
    return p
