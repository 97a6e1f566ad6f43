"""Oracle: beam-beam contact path, FP64 NumPy, structure-of-arrays pair state.

Test infrastructure only (see ``oracle/__init__.py``).

Restates, vectorised over pairs:
  - ``_broadphase_aabb``                contact/_broadphase.py:13-96
  - ``DetectCandidatesProcess``         contact/_manager_process.py:172-286
  - ``_closest_point_segments_batch``   contact/geometry/_compute.py:55-151
  - ``_build_contact_frame_batch``      contact/geometry/_compute.py:154-261
  - ``UpdateGeometryProcess``           contact/_manager_process.py:327-403, :475-494
  - the ContactGeometryStrategy twin    contact/geometry/strategy.py:47-305
    (``_detect_candidates``, ``_batch_update_geometry``,
    ``_build_constraint_jacobian_ptp``)
  - ``HuberContactForceProcess``        contact/contact_force/strategy.py:104-254
  - ``_friction_return_mapping_loop``   contact/friction/_assembly.py:260-342
    with ``_return_mapping_core`` / ``_tangent_2x2_core`` (law_friction.py:89-153)
  - friction tangents                   contact/friction/_assembly.py:108-257

Pair state layout follows ``_ContactStateOutput`` / ``_ContactPairOutput``
(contact/_contact_pair.py:39-81); status codes are ``ContactStatus``
(contact/_types.py:8-11): INACTIVE=0, ACTIVE=1, SLIDING=2.

Every per-pair update in the reference loops depends only on that pair, so the
vectorised form computes the same per-pair values; global vectors are summed in
pair order with ``np.add.at``.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

INACTIVE, ACTIVE, SLIDING = 0, 1, 2


@dataclass
class ContactConfig:
    """The ``_ContactConfigInput`` fields the path reads (contact/_contact_pair.py:84-170)."""

    g_on: float = 0.0
    g_off: float = 1e-6
    k_t_ratio: float = 0.5
    mu: float = 0.3
    midpoint_prescreening: bool = True
    prescreening_margin: float = 0.0
    exclude_same_layer: bool = False
    elem_layer_map: dict | None = None
    coating_stiffness: float = 0.0
    coating_damping: float = 0.0
    coating_mu: float = 0.0
    coating_k_t_ratio: float = 0.5


def _z(n, *shape, dtype=float):
    return np.zeros((n, *shape), dtype=dtype)


@dataclass
class Pairs:
    """Structure-of-arrays pair list (one row per ``_ContactPairOutput``)."""

    elem_a: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    elem_b: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    nodes_a: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))
    nodes_b: np.ndarray = field(default_factory=lambda: np.zeros((0, 2), np.int64))
    radius_a: np.ndarray = field(default_factory=lambda: np.zeros(0))
    radius_b: np.ndarray = field(default_factory=lambda: np.zeros(0))
    core_radius_a: np.ndarray = field(default_factory=lambda: np.zeros(0))
    core_radius_b: np.ndarray = field(default_factory=lambda: np.zeros(0))
    s: np.ndarray = field(default_factory=lambda: np.zeros(0))
    t: np.ndarray = field(default_factory=lambda: np.zeros(0))
    gap: np.ndarray = field(default_factory=lambda: np.zeros(0))
    normal: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    tangent1: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    tangent2: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    k_pen: np.ndarray = field(default_factory=lambda: np.zeros(0))
    k_t: np.ndarray = field(default_factory=lambda: np.zeros(0))
    p_n: np.ndarray = field(default_factory=lambda: np.zeros(0))
    z_t: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    q_trial_norm: np.ndarray = field(default_factory=lambda: np.zeros(0))
    status: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    stick: np.ndarray = field(default_factory=lambda: np.zeros(0, bool))
    dissipation: np.ndarray = field(default_factory=lambda: np.zeros(0))
    # Kelvin-Voigt coating state (_contact_pair.py:57-62)
    coating_compression: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coating_compression_prev: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coating_z_t: np.ndarray = field(default_factory=lambda: np.zeros((0, 2)))
    coating_stick: np.ndarray = field(default_factory=lambda: np.zeros(0, bool))
    coating_q_trial_norm: np.ndarray = field(default_factory=lambda: np.zeros(0))
    coating_dissipation: np.ndarray = field(default_factory=lambda: np.zeros(0))

    def __len__(self) -> int:
        return len(self.elem_a)

    def copy(self) -> "Pairs":
        return Pairs(**{f.name: getattr(self, f.name).copy() for f in dataclasses.fields(self)})

    @staticmethod
    def fresh(ea, eb, conn, r, cr) -> "Pairs":
        """New pairs with default ``_ContactStateOutput()`` (``_make_pair``)."""
        n = len(ea)
        ea = np.asarray(ea, np.int64)
        eb = np.asarray(eb, np.int64)
        p = Pairs(
            elem_a=ea, elem_b=eb,
            nodes_a=np.asarray(conn, np.int64)[ea].reshape(n, 2),
            nodes_b=np.asarray(conn, np.int64)[eb].reshape(n, 2),
            radius_a=np.asarray(r, float)[ea], radius_b=np.asarray(r, float)[eb],
            core_radius_a=np.asarray(cr, float)[ea], core_radius_b=np.asarray(cr, float)[eb],
            s=_z(n), t=_z(n), gap=_z(n), normal=_z(n, 3), tangent1=_z(n, 3), tangent2=_z(n, 3),
            k_pen=_z(n), k_t=_z(n), p_n=_z(n), z_t=_z(n, 2), q_trial_norm=_z(n),
            status=np.zeros(n, np.int64), stick=np.ones(n, bool), dissipation=_z(n),
            coating_compression=_z(n), coating_compression_prev=_z(n), coating_z_t=_z(n, 2),
            coating_stick=np.ones(n, bool), coating_q_trial_norm=_z(n), coating_dissipation=_z(n),
        )
        return p

    def append(self, other: "Pairs") -> "Pairs":
        return Pairs(**{f.name: np.concatenate([getattr(self, f.name), getattr(other, f.name)])
                        for f in dataclasses.fields(self)})


# --------------------------------------------------------------------------
# broadphase
# --------------------------------------------------------------------------


def aabb_bounds(x0: np.ndarray, x1: np.ndarray, radii: np.ndarray, margin: float):
    """lo = min(x0,x1) - (r+margin), hi = max(x0,x1) + (r+margin)
    (_broadphase.py:45-49)."""
    ex = np.asarray(radii, float)[:, None] + margin
    return np.minimum(x0, x1) - ex, np.maximum(x0, x1) + ex


def broadphase_aabb(x0, x1, radii, *, margin: float = 0.0, cell_size: float | None = None):
    """Candidate pairs (i<j), sorted lexicographically.

    The reference bins every AABB into all uniform-grid cells it covers, takes
    every pair sharing a cell, and keeps those whose AABBs overlap inclusively
    on all three axes (_broadphase.py:52-96).  Two inclusive-overlapping AABBs
    always share the cell holding max(lo_i, lo_j) (floor(fl(x * inv)) is
    monotone), so the output is exactly the inclusive-overlap set; this
    restatement enumerates it with the same grid and returns it sorted (the
    reference returns it in Python-set order)."""
    x0 = np.asarray(x0, float)
    x1 = np.asarray(x1, float)
    n = len(x0)
    if n < 2:
        return np.zeros((0, 2), np.int64)
    lo, hi = aabb_bounds(x0, x1, radii, margin)
    if cell_size is None:
        cell_size = max(float(np.mean(np.max(hi - lo, axis=1))) * 1.5, 1e-30)
    inv = 1.0 / cell_size
    ilo = np.floor(lo * inv).astype(np.int64)
    ihi = np.floor(hi * inv).astype(np.int64)
    # enumerate (cell, segment) memberships
    ext = ihi - ilo + 1
    cnt = ext[:, 0] * ext[:, 1] * ext[:, 2]
    seg = np.repeat(np.arange(n), cnt)
    off = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)
    ez = ext[seg, 2]
    ey = ext[seg, 1]
    cz = ilo[seg, 2] + off % ez
    cy = ilo[seg, 1] + (off // ez) % ey
    cx = ilo[seg, 0] + off // (ez * ey)
    keys = np.stack([cx, cy, cz], axis=1)
    order = np.lexsort((seg, cz, cy, cx))
    keys = keys[order]
    seg = seg[order]
    newcell = np.ones(len(seg), bool)
    newcell[1:] = np.any(keys[1:] != keys[:-1], axis=1)
    starts = np.nonzero(newcell)[0]
    ends = np.append(starts[1:], len(seg))
    pi_list, pj_list = [], []
    for s0, s1 in zip(starts, ends):
        k = s1 - s0
        if k < 2:
            continue
        members = seg[s0:s1]
        a, b = np.triu_indices(k, 1)
        pi_list.append(members[a])
        pj_list.append(members[b])
    if not pi_list:
        return np.zeros((0, 2), np.int64)
    pi = np.concatenate(pi_list)
    pj = np.concatenate(pj_list)
    i = np.minimum(pi, pj)
    j = np.maximum(pi, pj)
    key = np.unique(i * n + j)
    i, j = key // n, key % n
    ok = np.all(lo[i] <= hi[j], axis=1) & np.all(lo[j] <= hi[i], axis=1)
    return np.stack([i[ok], j[ok]], axis=1)


def broadphase_bruteforce(x0, x1, radii, *, margin: float = 0.0):
    """All (i<j) with inclusive AABB overlap, by an O(n^2) sweep (checker for
    ``broadphase_aabb``)."""
    lo, hi = aabb_bounds(np.asarray(x0, float), np.asarray(x1, float), radii, margin)
    n = len(lo)
    out = []
    for i in range(n - 1):
        ok = np.all(lo[i] <= hi[i + 1:], axis=1) & np.all(lo[i + 1:] <= hi[i], axis=1)
        js = np.nonzero(ok)[0] + i + 1
        out.append(np.stack([np.full(len(js), i), js], axis=1))
    return np.concatenate(out).astype(np.int64) if out else np.zeros((0, 2), np.int64)


def midpoint_prescreen(cand, conn, coords, r_arr, prescreening_margin: float):
    """Keep mask: max(0, |m_a - m_b| - L_a/2 - L_b/2) <= r_a + r_b + extra with
    extra = margin or 0.5 * mean(r_a + r_b) over the candidates
    (``_midpoint_prescreening``, _manager_process.py:259-286)."""
    conn = np.asarray(conn, np.int64)
    n0, n1 = conn[cand[:, 0], 0], conn[cand[:, 0], 1]
    m0, m1 = conn[cand[:, 1], 0], conn[cand[:, 1], 1]
    mid_a = 0.5 * (coords[n0] + coords[n1])
    mid_b = 0.5 * (coords[m0] + coords[m1])
    ha = 0.5 * np.linalg.norm(coords[n1] - coords[n0], axis=1)
    hb = 0.5 * np.linalg.norm(coords[m1] - coords[m0], axis=1)
    md = np.linalg.norm(mid_a - mid_b, axis=1)
    ra, rb = r_arr[cand[:, 0]], r_arr[cand[:, 1]]
    mind = np.maximum(0.0, md - ha - hb)
    extra = prescreening_margin
    if extra <= 0.0:
        extra = float(np.mean(ra + rb)) * 0.5
    return mind <= ra + rb + extra


def filter_candidates(raw, conn, coords, r_arr, cfg: ContactConfig):
    """Shared-node, same-layer and midpoint filters of
    ``DetectCandidatesProcess.process`` (_manager_process.py:200-217)."""
    conn = np.asarray(conn, np.int64)
    if len(raw) == 0:
        return raw
    ca, cb = conn[raw[:, 0]], conn[raw[:, 1]]
    share = ((ca[:, 0] == cb[:, 0]) | (ca[:, 0] == cb[:, 1]) |
             (ca[:, 1] == cb[:, 0]) | (ca[:, 1] == cb[:, 1]))
    keep = ~share
    lm = cfg.elem_layer_map
    if cfg.exclude_same_layer and lm is not None:
        la = np.array([lm.get(int(i), -1) for i in raw[:, 0]])
        lb = np.array([lm.get(int(j), -1) for j in raw[:, 1]])
        keep &= ~((la == lb) & (la >= 0))
    out = raw[keep]
    if cfg.midpoint_prescreening and len(out):
        out = out[midpoint_prescreen(out, conn, coords, r_arr, cfg.prescreening_margin)]
    return out


def detect_candidates(pairs: Pairs, coords, conn, radii, cfg: ContactConfig, *,
                      margin: float = 0.0, cell_size: float | None = None, core_radii=None):
    """``DetectCandidatesProcess.process`` (_manager_process.py:172-257):
    broadphase -> filters -> existing pairs not among the candidates become
    INACTIVE (state otherwise kept) -> new candidates appended with fresh state.

    Returns ``(pairs, candidates)``; new pairs are appended in candidate order
    (sorted (i<j) here)."""
    coords = np.asarray(coords, float)
    conn = np.asarray(conn, np.int64)
    ne = len(conn)
    r_arr = np.full(ne, float(radii)) if np.isscalar(radii) else np.asarray(radii, float)
    if core_radii is None:
        cr_arr = r_arr.copy()
    elif np.isscalar(core_radii):
        cr_arr = np.full(ne, float(core_radii))
    else:
        cr_arr = np.asarray(core_radii, float)
    raw = broadphase_aabb(coords[conn[:, 0]], coords[conn[:, 1]], r_arr,
                          margin=margin, cell_size=cell_size)
    cand = filter_candidates(raw, conn, coords, r_arr, cfg)
    pairs = pairs.copy()
    ck = set(map(tuple, cand.tolist()))
    ex_keys = [(min(a, b), max(a, b)) for a, b in zip(pairs.elem_a.tolist(), pairs.elem_b.tolist())]
    drop = np.array([k not in ck for k in ex_keys], bool)
    if len(pairs):
        pairs.status[drop] = INACTIVE
    exist = set(ex_keys)
    new = [k for k in map(tuple, cand.tolist()) if k not in exist]
    if new:
        new = np.array(new, np.int64)
        pairs = pairs.append(Pairs.fresh(new[:, 0], new[:, 1], conn, r_arr, cr_arr))
    return pairs, cand


# --------------------------------------------------------------------------
# narrowphase + frame
# --------------------------------------------------------------------------


def closest_point_segments(xA0, xA1, xB0, xB1, tol_parallel: float = 1e-10):
    """Clamped segment-segment closest points (_compute.py:55-151).

    Returns s, t, distance, normal, is_parallel."""
    dA = xA1 - xA0
    dB = xB1 - xB0
    w0 = xA0 - xB0
    a = np.einsum("ij,ij->i", dA, dA)
    b = np.einsum("ij,ij->i", dA, dB)
    c = np.einsum("ij,ij->i", dB, dB)
    d = np.einsum("ij,ij->i", dA, w0)
    e = np.einsum("ij,ij->i", dB, w0)
    det = a * c - b * b
    par = det < tol_parallel * np.maximum(a * c, 1e-30)
    sdet = np.where(par, 1.0, det)
    s_unc = np.where(par, 0.0, (b * e - c * d) / sdet)
    t_unc = np.where(par, 0.0, (a * e - b * d) / sdet)
    c_ok = c > tol_parallel
    a_ok = a > tol_parallel
    sc = np.where(c_ok, c, 1.0)
    sa = np.where(a_ok, a, 1.0)
    t_par = np.where(c_ok, np.clip(e / sc, 0.0, 1.0), 0.0)
    s = np.where(par, 0.0, np.clip(s_unc, 0.0, 1.0))
    t = np.where(par, t_par, np.clip(t_unc, 0.0, 1.0))
    s_cl = (s_unc < 0.0) | (s_unc > 1.0)
    t_cl = (t_unc < 0.0) | (t_unc > 1.0)
    t_re = np.where(c_ok, np.clip((b * s + e) / sc, 0.0, 1.0), 0.0)
    t = np.where(s_cl & ~par, t_re, t)
    need_s = t_cl & ~par
    s_re = np.where(a_ok, np.clip((b * t - d) / sa, 0.0, 1.0), 0.0)
    s = np.where(need_s, s_re, s)
    t_re2 = np.where(c_ok, np.clip((b * s + e) / sc, 0.0, 1.0), 0.0)
    t = np.where(need_s, t_re2, t)
    pa = xA0 + s[:, None] * dA
    pb = xB0 + t[:, None] * dB
    diff = pa - pb
    dist = np.sqrt(np.einsum("ij,ij->i", diff, diff))
    nrm = diff / np.maximum(dist, 1e-30)[:, None]
    nrm[dist < 1e-30] = [0.0, 0.0, 1.0]
    return s, t, dist, nrm, par


def build_contact_frame(normals, prev_t1, prev_n, has_prev, has_prev_n):
    """(n, t1, t2) by parallel transport of the previous t1 along the minimal
    rotation n_old -> n_new (when |n_old x n_new| >= 1e-12), projected and
    normalised (valid if > 1e-10); else Gram-Schmidt of the previous t1; else
    the unit axis of min |n| projected (_compute.py:154-261)."""
    N = len(normals)
    n = normals / np.maximum(np.sqrt(np.einsum("ij,ij->i", normals, normals)), 1e-30)[:, None]
    t1 = np.zeros((N, 3))
    t2 = np.zeros((N, 3))
    done = np.zeros(N, bool)
    idx = np.nonzero(has_prev & has_prev_n)[0]
    if len(idx):
        no = prev_n[idx]
        no = no / np.maximum(np.sqrt(np.einsum("ij,ij->i", no, no)), 1e-30)[:, None]
        nn = n[idx]
        to = prev_t1[idx]
        v = np.cross(no, nn)
        sv = np.sqrt(np.einsum("ij,ij->i", v, v))
        cv = np.einsum("ij,ij->i", no, nn)
        tt = to.copy()
        big = sv >= 1e-12
        if big.any():
            vb = v[big]
            vxt = np.cross(vb, to[big])
            vxvxt = np.cross(vb, vxt)
            fac = (1.0 - cv[big]) / (sv[big] * sv[big])
            tt[big] = to[big] + vxt + fac[:, None] * vxvxt
        raw = tt - np.einsum("ij,ij->i", tt, nn)[:, None] * nn
        rn = np.sqrt(np.einsum("ij,ij->i", raw, raw))
        ok = rn > 1e-10
        vi = idx[ok]
        t1[vi] = raw[ok] / rn[ok, None]
        t2[vi] = np.cross(n[vi], t1[vi])
        done[vi] = True
    idx = np.nonzero(has_prev & ~done)[0]
    if len(idx):
        pt = prev_t1[idx]
        ni = n[idx]
        raw = pt - np.einsum("ij,ij->i", pt, ni)[:, None] * ni
        rn = np.sqrt(np.einsum("ij,ij->i", raw, raw))
        ok = rn > 1e-10
        vi = idx[ok]
        t1[vi] = raw[ok] / rn[ok, None]
        t2[vi] = np.cross(n[vi], t1[vi])
        done[vi] = True
    idx = np.nonzero(~done)[0]
    if len(idx):
        ni = n[idx]
        ref = np.zeros_like(ni)
        ref[np.arange(len(idx)), np.argmin(np.abs(ni), axis=1)] = 1.0
        raw = ref - np.einsum("ij,ij->i", ref, ni)[:, None] * ni
        rn = np.sqrt(np.einsum("ij,ij->i", raw, raw))
        t1[idx] = raw / np.maximum(rn, 1e-30)[:, None]
        t2[idx] = np.cross(n[idx], t1[idx])
    return n, t1, t2


def update_geometry(pairs: Pairs, coords, cfg: ContactConfig, *,
                    allow_deactivation: bool = True, freeze_active_set: bool = False) -> Pairs:
    """``UpdateGeometryProcess.process`` (_manager_process.py:327-403) with
    ``_update_active_set_state`` (:475-494) when not frozen.

    With the coating on (coating_stiffness > 0) the gap is the core gap
    dist - (core_a + core_b), the coating compression is
    max(0, (r_a - core_a) + (r_b - core_b) - gap_core) (:351-360), and a pair
    with compression > 0 is activated and never deactivated."""
    p = pairs.copy()
    if len(p) == 0:
        return p
    coords = np.asarray(coords, float)
    s, t, dist, nrm, _ = closest_point_segments(
        coords[p.nodes_a[:, 0]], coords[p.nodes_a[:, 1]],
        coords[p.nodes_b[:, 0]], coords[p.nodes_b[:, 1]])
    coat = cfg.coating_stiffness > 0.0
    if coat:
        gap = dist - (p.core_radius_a + p.core_radius_b)
        coat_total = (p.radius_a - p.core_radius_a) + (p.radius_b - p.core_radius_b)
        p.coating_compression = np.maximum(0.0, coat_total - gap)
    else:
        gap = dist - (p.radius_a + p.radius_b)
    has_prev = np.sqrt(np.einsum("ij,ij->i", p.tangent1, p.tangent1)) > 1e-10
    has_prev_n = np.sqrt(np.einsum("ij,ij->i", p.normal, p.normal)) > 1e-10
    n, t1, t2 = build_contact_frame(nrm, p.tangent1, p.normal, has_prev, has_prev_n)
    p.s, p.t, p.gap, p.normal, p.tangent1, p.tangent2 = s, t, gap, n, t1, t2
    if not freeze_active_set:
        inact = p.status == INACTIVE
        coat_active = (p.coating_compression > 0.0) if coat else np.zeros(len(p), bool)
        on = inact & ((gap <= cfg.g_on) | coat_active)
        off = (~inact) & allow_deactivation & (gap >= cfg.g_off) & ~coat_active
        p.status = np.where(on, ACTIVE, np.where(off, INACTIVE, p.status))
    return p


# --------------------------------------------------------------------------
# ContactGeometryStrategy twin
# --------------------------------------------------------------------------


def strategy_update_geometry(pairs: Pairs, coords, config: ContactConfig | None = None) -> Pairs:
    """``_batch_update_geometry`` (geometry/strategy.py:205-305), the update of
    ``PointToPointProcess.update_geometry``: narrowphase and frame as
    ``UpdateGeometryProcess``, then ``_update_active_set_hysteresis`` (:47-80)
    with g_on / g_off / coating_stiffness from ``config`` (0, 0, 0 without
    one), deactivation always allowed, never frozen."""
    cfg = ContactConfig(g_on=0.0, g_off=0.0) if config is None else config
    return update_geometry(pairs, coords, cfg, allow_deactivation=True, freeze_active_set=False)


def strategy_detect(coords, conn, radii, *, exclude_same_layer: bool = True, margin: float = 0.0,
                    cell_size: float | None = None) -> Pairs:
    """``_detect_candidates`` (geometry/strategy.py:135-199): broadphase, the
    shared-node filter when ``exclude_same_layer`` (no layer map, no midpoint
    prescreen), every pair new with a fresh state and core radii 0.0 (the
    ``_ContactPairOutput`` defaults, _contact_pair.py:79-80), then the initial
    narrowphase without a config."""
    coords = np.asarray(coords, float)
    conn = np.asarray(conn, np.int64)
    ne = len(conn)
    if ne < 2:
        return Pairs()
    r_arr = np.full(ne, float(radii)) if np.isscalar(radii) else np.asarray(radii, float)
    raw = broadphase_aabb(coords[conn[:, 0]], coords[conn[:, 1]], r_arr, margin=margin, cell_size=cell_size)
    if exclude_same_layer and len(raw):
        ca, cb = conn[raw[:, 0]], conn[raw[:, 1]]
        share = ((ca[:, 0] == cb[:, 0]) | (ca[:, 0] == cb[:, 1]) |
                 (ca[:, 1] == cb[:, 0]) | (ca[:, 1] == cb[:, 1]))
        raw = raw[~share]
    if len(raw) == 0:
        return Pairs()
    p = Pairs.fresh(raw[:, 0], raw[:, 1], conn, r_arr, np.zeros(ne))
    return strategy_update_geometry(p, coords, None)


def constraint_jacobian_ptp(pairs: Pairs, ndof_total: int, ndof_per_node: int = 6):
    """``_build_constraint_jacobian_ptp`` (geometry/strategy.py:83-132):
    G = d g_n / d u, one row per pair with status != INACTIVE (pair order),
    entries c_k n_d at DOF node_k * ndof_per_node + d of [a0, a1, b0, b1],
    |c_k n_d| <= 1e-30 dropped, COO -> CSR (duplicates summed).
    Returns (G, active_indices)."""
    act = np.nonzero(pairs.status != INACTIVE)[0]
    if len(act) == 0:
        return sp.csr_matrix((0, ndof_total)), []
    c = shape_coeffs(pairs)[act]                                   # (m, 4)
    nodes = np.stack([pairs.nodes_a[act, 0], pairs.nodes_a[act, 1],
                      pairs.nodes_b[act, 0], pairs.nodes_b[act, 1]], axis=1)
    vals = c[:, :, None] * pairs.normal[act][:, None, :]           # (m, 4, 3)
    cols = nodes[:, :, None] * ndof_per_node + np.arange(3)[None, None, :]
    rows = np.broadcast_to(np.arange(len(act))[:, None, None], vals.shape)
    keep = np.abs(vals) > 1e-30
    G = sp.coo_matrix((vals[keep], (rows[keep], cols[keep])), shape=(len(act), ndof_total)).tocsr()
    return G, act.tolist()


# --------------------------------------------------------------------------
# Huber normal contact
# --------------------------------------------------------------------------


def huber(x, delta):
    """C1 max(0,x) (contact_force/strategy.py:104-113)."""
    x = np.asarray(x, float)
    if delta <= 0.0:
        return np.maximum(0.0, x)
    return np.where(x < -delta, 0.0, np.where(x > delta, x, (x + delta) ** 2 / (4.0 * delta)))


def huber_deriv(x, delta):
    """C0 derivative of ``huber`` (contact_force/strategy.py:115-124)."""
    x = np.asarray(x, float)
    if delta <= 0.0:
        return np.where(x > 0.0, 1.0, 0.0)
    return np.where(x < -delta, 0.0, np.where(x > delta, 1.0, (x + delta) / (2.0 * delta)))


def contact_dofs(p: Pairs) -> np.ndarray:
    """(N,4,3) translational global DOFs of nodes A0, A1, B0, B1 (_assembly_utils.py:8-25)."""
    nodes = np.stack([p.nodes_a[:, 0], p.nodes_a[:, 1], p.nodes_b[:, 0], p.nodes_b[:, 1]], axis=1)
    return nodes[:, :, None] * 6 + np.arange(3)[None, None, :]


def shape_coeffs(p: Pairs) -> np.ndarray:
    """c = [(1-s), s, -(1-t), -t] (contact_force/strategy.py:44-63)."""
    return np.stack([1.0 - p.s, p.s, -(1.0 - p.t), -p.t], axis=1)


def huber_evaluate(p: Pairs, k_pen: float, smoothing_delta: float, ndof: int):
    """``HuberContactForceProcess.evaluate`` (contact_force/strategy.py:126-170).

    Returns (f_c, residuals, pairs).  p_n is written for every non-INACTIVE
    pair; f_c += p_n * g_shape where p_n > 1e-30."""
    p = p.copy()
    dh = k_pen / smoothing_delta if smoothing_delta > 0.0 else 0.0
    act = p.status != INACTIVE
    pn = huber(k_pen * (-p.gap), dh)
    p.p_n = np.where(act, pn, p.p_n)
    res = np.where(p.p_n[act] > 0.0, k_pen * p.gap[act], 0.0)
    f = np.zeros(ndof)
    m = act & (p.p_n > 1e-30)
    if m.any():
        c = shape_coeffs(p)[m]
        vals = p.p_n[m, None, None] * c[:, :, None] * p.normal[m][:, None, :]
        np.add.at(f, contact_dofs(p)[m].ravel(), vals.ravel())
    return f, res, p


def huber_tangent_blocks(p: Pairs, k_pen: float, smoothing_delta: float):
    """Per-pair 4x4 node blocks of 3x3 of K_c (contact_force/strategy.py:172-254):
    cc [w_mat n n^T - w_geo (I - n n^T)], w_mat = h' k_pen, w_geo = p_n/dist
    (0 when dist <= 1e-15), over non-INACTIVE pairs not skipped by
    h' < 1e-30 and p_n < 1e-30; coefficient pairs with |c| < 1e-30 and entries
    with |val| <= 1e-30 are dropped.  Returns (mask, blocks (M,4,4,3,3))."""
    dh = k_pen / smoothing_delta if smoothing_delta > 0.0 else 0.0
    hd = huber_deriv(k_pen * (-p.gap), dh)
    m = (p.status != INACTIVE) & ~((hd < 1e-30) & (p.p_n < 1e-30))
    n = p.normal[m]
    wm = hd[m] * k_pen
    dist = p.gap[m] + p.radius_a[m] + p.radius_b[m]
    with np.errstate(divide="ignore", invalid="ignore"):
        wg = np.where(dist > 1e-15, p.p_n[m] / dist, 0.0)
    nn = n[:, :, None] * n[:, None, :]
    c = shape_coeffs(p)[m]
    cz = np.abs(c) < 1e-30
    cc = c[:, :, None] * c[:, None, :]
    # the reference forms val = w_mat*cc*n_i*n_j - w_geo*cc*(d_ij - n_i n_j) entrywise
    val = (wm[:, None, None, None, None] * cc[:, :, :, None, None] * nn[:, None, None]
           - wg[:, None, None, None, None] * cc[:, :, :, None, None]
           * (np.eye(3)[None, None, None] - nn[:, None, None]))
    skip = cz[:, :, None] | cz[:, None, :]
    val[skip] = 0.0
    val[np.abs(val) <= 1e-30] = 0.0
    return m, val


def blocks_to_csr(p: Pairs, mask, blocks, ndof: int) -> sp.csr_matrix:
    """Scatter (M,4,4,3,3) translational node blocks into an ndof x ndof CSR
    (COO with duplicate summation; zero entries dropped as the reference does)."""
    dofs = contact_dofs(p)[mask]  # (M,4,3)
    rows = np.broadcast_to(dofs[:, :, None, :, None], blocks.shape)
    cols = np.broadcast_to(dofs[:, None, :, None, :], blocks.shape)
    nz = blocks != 0.0
    return sp.csr_matrix((blocks[nz], (rows[nz], cols[nz])), shape=(ndof, ndof))


def huber_tangent(p: Pairs, k_pen: float, smoothing_delta: float, ndof: int) -> sp.csr_matrix:
    """``HuberContactForceProcess.tangent`` as CSR."""
    m, val = huber_tangent_blocks(p, k_pen, smoothing_delta)
    return blocks_to_csr(p, m, val, ndof)


# --------------------------------------------------------------------------
# Coulomb friction
# --------------------------------------------------------------------------


@dataclass
class FrictionResult:
    f_fric: np.ndarray
    residual: np.ndarray
    mask: np.ndarray      # pairs in friction_tangents / friction_forces_local
    q: np.ndarray         # (M,2) local friction forces of masked pairs
    D_t: np.ndarray       # (M,2,2)
    pairs: Pairs


def friction_evaluate(p: Pairs, u, u_ref, k_pen: float, k_t_ratio: float, mu_eff: float,
                      ndof: int) -> FrictionResult:
    """``_friction_return_mapping_loop`` (friction/_assembly.py:260-342).

    Every pair with p_n > 0 (regardless of status) is processed when mu_eff > 0:
    k_pen/k_t initialised if k_pen <= 0; du_t = [(du_B - du_A).t1, .t2] with
    du = u - u_ref at the closest points; q_trial = z_t + k_t du_t; stick if
    |q_trial| <= mu p_n else q = mu p_n q_trial/|q_trial|; the state gets
    z_t <- q, stick, q_trial_norm, dissipation = q.du_t and status
    ACTIVE (stick) / SLIDING (slip).  Pairs with |q| >= 1e-30 enter the force
    and tangent dictionaries with D_t from ``_tangent_2x2_core``."""
    p = p.copy()
    N = len(p)
    sel = (p.p_n > 0.0) & (mu_eff > 0.0) if N else np.zeros(0, bool)
    init = sel & (p.k_pen <= 0.0)
    p.k_pen = np.where(init, k_pen, p.k_pen)
    p.k_t = np.where(init, k_pen * k_t_ratio, p.k_t)
    du = np.asarray(u, float) - np.asarray(u_ref, float)
    dn = lambda nodes: du[(nodes[:, None] * 6 + np.arange(3)[None, :])]
    s, t = p.s[:, None], p.t[:, None]
    duA = (1.0 - s) * dn(p.nodes_a[:, 0]) + s * dn(p.nodes_a[:, 1])
    duB = (1.0 - t) * dn(p.nodes_b[:, 0]) + t * dn(p.nodes_b[:, 1])
    dur = duB - duA
    dut = np.stack([np.einsum("ij,ij->i", dur, p.tangent1), np.einsum("ij,ij->i", dur, p.tangent2)],
                   axis=1)
    qt = p.z_t + p.k_t[:, None] * dut
    qtn = np.linalg.norm(qt, axis=1)
    fy = mu_eff * p.p_n
    stick = qtn <= fy
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(stick[:, None], qt, fy[:, None] * qt / qtn[:, None])
    diss = np.einsum("ij,ij->i", q, dut)
    p.z_t = np.where(sel[:, None], q, p.z_t)
    p.stick = np.where(sel, stick, p.stick)
    p.q_trial_norm = np.where(sel, qtn, p.q_trial_norm)
    p.dissipation = np.where(sel, diss, p.dissipation)
    p.status = np.where(sel, np.where(stick, ACTIVE, SLIDING), p.status)
    qn = np.linalg.norm(q, axis=1)
    m = sel & (qn >= 1e-30)
    qm = q[m]
    res = np.maximum(0.0, qn[m] - fy[m])
    # D_t (_tangent_2x2_core, law_friction.py:126-153) with z_t = q (new state)
    kt = p.k_t[m]
    D = np.zeros((int(m.sum()), 2, 2))
    stk = stick[m]
    zn = np.linalg.norm(qm, axis=1)
    qtm = qtn[m]
    eye = np.eye(2)[None]
    fallback = stk | (zn < 1e-30) | (qtm < 1e-30)
    D[fallback] = kt[fallback, None, None] * eye
    sl = ~fallback
    if sl.any():
        qh = qm[sl] / zn[sl, None]
        ratio = (mu_eff * p.p_n[m][sl]) / qtm[sl]
        D[sl] = (ratio * kt[sl])[:, None, None] * (eye - qh[:, :, None] * qh[:, None, :])
    # friction force (_assemble_friction_force, friction/_assembly.py:85-105)
    f = np.zeros(ndof)
    if m.any():
        c = shape_coeffs(p)[m]
        tv = np.stack([p.tangent1[m], p.tangent2[m]], axis=1)  # (M,2,3)
        keep = np.abs(qm) >= 1e-30
        contrib = np.einsum("ma,mk,mad->mkd", np.where(keep, qm, 0.0), c, tv)
        np.add.at(f, contact_dofs(p)[m].ravel(), contrib.ravel())
    return FrictionResult(f_fric=f, residual=res, mask=m, q=qm, D_t=D, pairs=p)


def friction_material_blocks(p: Pairs, fr: FrictionResult):
    """K_fm = sum_ab D_ab g_ta g_tb^T over non-INACTIVE dictionary pairs,
    entries with |D_ab| < 1e-30 or |val| <= 1e-30 dropped
    (``_assemble_friction_tangent_stiffness``, friction/_assembly.py:108-157).
    Returns (mask, blocks (M,4,4,3,3))."""
    m = fr.mask & (p.status != INACTIVE)
    sub = m[fr.mask]
    D = fr.D_t[sub]
    c = shape_coeffs(p)[m]
    tv = np.stack([p.tangent1[m], p.tangent2[m]], axis=1)  # (M,2,3)
    g = c[:, None, :, None] * tv[:, :, None, :]  # (M,2,4,3) = g_ta
    out = np.zeros((int(m.sum()), 4, 4, 3, 3))
    for a1 in range(2):
        for a2 in range(2):
            d = D[:, a1, a2]
            term = d[:, None, None, None, None] * (
                g[:, a1][:, :, None, :, None] * g[:, a2][:, None, :, None, :])
            term[np.abs(d) < 1e-30] = 0.0
            term[np.abs(term) <= 1e-30] = 0.0
            out += term
    return m, out


def friction_geometric_blocks(p: Pairs, fr: FrictionResult):
    """K_fg blocks (c_k c_l / dist) M with
    M_ij = -q1 n_i t1_j - q2 t2_i n_j + q2 eps_ijk t1_k, over non-INACTIVE
    dictionary pairs with (|q1| >= 1e-30 or |q2| >= 1e-30) and dist >= 1e-15;
    coefficient pairs with |c| < 1e-30 and entries with |val| <= 1e-30 dropped
    (``_assemble_friction_geometric_stiffness``, friction/_assembly.py:160-257)."""
    m0 = fr.mask & (p.status != INACTIVE)
    q = fr.q[m0[fr.mask]]
    dist_all = p.gap + p.radius_a + p.radius_b
    qok = ~((np.abs(q[:, 0]) < 1e-30) & (np.abs(q[:, 1]) < 1e-30))
    dok = dist_all[m0] >= 1e-15
    ok = qok & dok
    m = m0.copy()
    m[np.nonzero(m0)[0][~ok]] = False
    q = q[ok]
    n, t1, t2 = p.normal[m], p.tangent1[m], p.tangent2[m]
    q1, q2 = q[:, 0], q[:, 1]
    M = -q1[:, None, None] * n[:, :, None] * t1[:, None, :] - q2[:, None, None] * t2[:, :, None] * n[:, None, :]
    M[:, 0, 1] += q2 * t1[:, 2]
    M[:, 0, 2] += q2 * (-t1[:, 1])
    M[:, 1, 0] += q2 * (-t1[:, 2])
    M[:, 1, 2] += q2 * t1[:, 0]
    M[:, 2, 0] += q2 * t1[:, 1]
    M[:, 2, 1] += q2 * (-t1[:, 0])
    c = shape_coeffs(p)[m]
    inv = 1.0 / dist_all[m]
    w = c[:, :, None] * c[:, None, :] * inv[:, None, None]
    val = w[:, :, :, None, None] * M[:, None, None]
    cz = np.abs(c) < 1e-30
    val[cz[:, :, None] | cz[:, None, :]] = 0.0
    val[np.abs(val) <= 1e-30] = 0.0
    return m, val


def friction_tangent(p: Pairs, fr: FrictionResult, ndof: int) -> sp.csr_matrix:
    """``CoulombReturnMappingProcess.tangent`` = K_fm + K_fg (friction/strategy.py:150-164)."""
    m1, b1 = friction_material_blocks(p, fr)
    m2, b2 = friction_geometric_blocks(p, fr)
    return blocks_to_csr(p, m1, b1, ndof) + blocks_to_csr(p, m2, b2, ndof)
