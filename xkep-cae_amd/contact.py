"""Drop-in contact path on the GPU: pair manager, DetectCandidates /
UpdateGeometry Processes, Huber normal force and Coulomb friction Strategies.

Mirrors the reference's interfaces:
  - ``_ContactConfigInput`` fields used by the path (contact/_contact_pair.py:84-170)
  - ``_ContactStateOutput`` / ``_ContactPairOutput`` views (:39-81)
  - ``DetectCandidatesProcess`` / ``UpdateGeometryProcess`` with their Input /
    Output dataclasses (contact/_manager_process.py:135-403)
  - ``HuberContactForceProcess`` (contact/contact_force/strategy.py:69-254)
    behind ``ContactForceStrategy`` (core/strategies/protocols.py:16-45)
  - ``CoulombReturnMappingProcess`` (contact/friction/strategy.py:50-168)
    behind ``FrictionStrategy`` (protocols.py:48-78)
  - ``NoCoatingProcess`` / ``KelvinVoigtCoatingProcess`` and
    ``_create_coating_strategy`` (contact/coating/strategy.py:19-408)

The pair state lives on the GPU (structure of arrays in the device context).
``manager.pairs`` materialises read-only views on demand; the strategies never
walk it -- they launch the pair kernel and gather kernels of csrc/.
"""

from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field
from enum import Enum
from typing import Any

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .context import DeviceContext


class ContactStatus(Enum):
    """contact/_types.py:8-11."""

    INACTIVE = 0
    ACTIVE = 1
    SLIDING = 2


@dataclass(frozen=True)
class ContactConfig:
    """The ``_ContactConfigInput`` fields read on this path, same names and defaults."""

    k_t_ratio: float = 0.5
    mu: float = 0.3
    g_on: float = 0.0
    g_off: float = 1e-6
    exclude_same_layer: bool = False
    elem_layer_map: dict | None = None
    midpoint_prescreening: bool = True
    prescreening_margin: float = 0.0
    coating_stiffness: float = 0.0
    coating_damping: float = 0.0
    coating_mu: float = 0.0
    coating_k_t_ratio: float = 0.5
    smoothing_delta: float = 0.0


def layer_array(cfg: ContactConfig, n_elems: int):
    """Per-element layer ids (int32, -1 = none) for the same-layer filter of
    DetectCandidatesProcess (contact/_manager_process.py:201-212), or None when
    the filter is off.  ``elem_layer_map`` is the reference's dict; an integer
    array of length n_elems is accepted as well."""
    if not (cfg.exclude_same_layer and cfg.elem_layer_map is not None):
        return None
    m = cfg.elem_layer_map
    if isinstance(m, dict):
        out = np.full(n_elems, -1, np.int32)
        if m:
            k = np.fromiter(m.keys(), np.int64, len(m))
            v = np.fromiter(m.values(), np.int64, len(m))
            ok = (k >= 0) & (k < n_elems)
            out[k[ok]] = v[ok]
        return out
    a = np.asarray(m)
    if a.shape != (n_elems,):
        raise ValueError("elem_layer_map array must have one entry per element")
    return a.astype(np.int32)


@dataclass(frozen=True)
class ContactState:
    """Read-only snapshot of one pair's ``_ContactStateOutput``."""

    s: float = 0.0
    t: float = 0.0
    gap: float = 0.0
    normal: np.ndarray = field(default_factory=lambda: np.zeros(3))
    tangent1: np.ndarray = field(default_factory=lambda: np.zeros(3))
    tangent2: np.ndarray = field(default_factory=lambda: np.zeros(3))
    lambda_n: float = 0.0
    k_pen: float = 0.0
    k_t: float = 0.0
    p_n: float = 0.0
    z_t: np.ndarray = field(default_factory=lambda: np.zeros(2))
    q_trial_norm: float = 0.0
    status: ContactStatus = ContactStatus.INACTIVE
    stick: bool = True
    dissipation: float = 0.0
    coating_compression: float = 0.0


@dataclass(frozen=True)
class ContactPair:
    """Read-only snapshot of one ``_ContactPairOutput``."""

    elem_a: int
    elem_b: int
    nodes_a: np.ndarray
    nodes_b: np.ndarray
    state: ContactState
    radius_a: float = 0.0
    radius_b: float = 0.0


class DevicePairList:
    """Sequence view of the device pair list (``manager.pairs``).

    Indexing and iteration download the state once and build frozen snapshots.
    Slice assignment from a list of the same manager is accepted as a no-op so
    that ``manager.pairs[:] = out.manager.pairs`` (contact/solver/_newton_steps.py:94)
    keeps working: the device list is already the updated one.
    """

    def __init__(self, manager: "ContactManager"):
        self.manager = manager
        self._snap = None

    def __len__(self):
        return self.manager.ctx.n_pairs

    def _materialise(self):
        if self._snap is None:
            ctx = self.manager.ctx
            d = ctx.get_pairs()
            cc = ctx.get_coating()["coating_compression"] if ctx.n_pairs else np.zeros(0)
            conn = ctx.conn_np
            rad = ctx.radii_np
            out = []
            for i in range(len(d["elem_a"])):
                ea, eb = int(d["elem_a"][i]), int(d["elem_b"][i])
                st = ContactState(s=d["s"][i], t=d["t"][i], gap=d["gap"][i], normal=d["normal"][i],
                                  tangent1=d["tangent1"][i], tangent2=d["tangent2"][i], k_pen=d["k_pen"][i],
                                  k_t=d["k_t"][i], p_n=d["p_n"][i], z_t=d["z_t"][i],
                                  q_trial_norm=d["q_trial_norm"][i], status=ContactStatus(int(d["status"][i])),
                                  stick=bool(d["stick"][i]), dissipation=d["dissipation"][i],
                                  coating_compression=float(cc[i]))
                out.append(ContactPair(ea, eb, conn[ea].astype(np.int64), conn[eb].astype(np.int64), st,
                                       float(rad[ea]), float(rad[eb])))
            self._snap = out
        return self._snap

    def __getitem__(self, k):
        return self._materialise()[k]

    def __iter__(self):
        return iter(self._materialise())

    def __setitem__(self, k, v):
        if isinstance(v, DevicePairList) and v.manager is self.manager:
            self._snap = None
            return
        raise TypeError("the device pair list is updated by the Processes; direct assignment is not supported")

    def __bool__(self):
        return len(self) > 0


class ContactManager:
    """Device-resident ``_ContactManagerInput`` (pairs + config)."""

    def __init__(self, ctx: DeviceContext, config: ContactConfig | None = None):
        self.ctx = ctx
        self.config = config or ContactConfig()

    @property
    def pairs(self) -> DevicePairList:
        return DevicePairList(self)

    @property
    def n_pairs(self) -> int:
        return self.ctx.n_pairs

    def n_active(self) -> int:
        """Pairs with status != INACTIVE (``_n_active``, contact/_contact_pair.py)."""
        return int(np.count_nonzero(self.ctx.get_pairs()["status"] != 0))


# ------------------------------------------------------------------ Processes


@dataclass(frozen=True)
class DetectCandidatesInput:
    manager: ContactManager
    node_coords: Any
    connectivity: Any = None
    radii: Any = 0.0
    margin: float = 0.0
    cell_size: float | None = None
    core_radii: Any = None


@dataclass(frozen=True)
class DetectCandidatesOutput:
    manager: ContactManager
    candidates: list
    n_pairs: int


class DetectCandidatesProcess:
    """Broadphase + filters + pair-list update on the GPU
    (contact/_manager_process.py:157-286).  ``connectivity`` and ``radii`` must
    be the ones the device context was built with."""

    def process(self, inp: DetectCandidatesInput) -> DetectCandidatesOutput:
        m = inp.manager
        ctx = m.ctx
        cfg = m.config
        if inp.connectivity is not None and not np.array_equal(np.asarray(inp.connectivity), ctx.conn_np):
            raise ValueError("connectivity differs from the device context's")
        if not np.isscalar(inp.radii) or float(inp.radii) != 0.0:
            r = np.full(ctx.n_elems, float(inp.radii)) if np.isscalar(inp.radii) else np.asarray(inp.radii, float)
            if not np.array_equal(r, ctx.radii_np):
                raise ValueError("radii differ from the device context's")
        if cfg.coating_stiffness > 0.0 or ctx.coat_on or inp.core_radii is not None:
            # core radii of the pairs created here (None: the contact radii, :184-189)
            ctx.ensure_coating(cfg.coating_stiffness, cfg.coating_damping, cfg.coating_mu, cfg.coating_k_t_ratio,
                               inp.core_radii)
        coords = ctx.to_dev(inp.node_coords)
        layer = layer_array(cfg, ctx.n_elems)
        ctx.detect_candidates(coords, margin=inp.margin, cell_size=inp.cell_size, elem_layer=layer,
                              midpoint_prescreening=cfg.midpoint_prescreening,
                              prescreening_margin=cfg.prescreening_margin)
        cand = [tuple(map(int, p)) for p in ctx.last_candidates()]
        return DetectCandidatesOutput(manager=m, candidates=cand, n_pairs=ctx.n_pairs)


@dataclass(frozen=True)
class UpdateGeometryInput:
    manager: ContactManager
    node_coords: Any
    allow_deactivation: bool = True
    freeze_active_set: bool = False


@dataclass(frozen=True)
class UpdateGeometryOutput:
    manager: ContactManager
    n_active: int


class UpdateGeometryProcess:
    """Narrowphase + frame + (unless frozen) hysteresis on the GPU
    (contact/_manager_process.py:312-403)."""

    def process(self, inp: UpdateGeometryInput) -> UpdateGeometryOutput:
        m = inp.manager
        cfg = m.config
        ctx = m.ctx
        if cfg.coating_stiffness > 0.0 or ctx.coat_on:
            # with the coating on: core gap, compression, activation by compression (:351-360, :486-494)
            ctx.ensure_coating(cfg.coating_stiffness, cfg.coating_damping, cfg.coating_mu, cfg.coating_k_t_ratio)
        coords = ctx.to_dev(inp.node_coords)
        modes = L.PAIR_GEOM | (0 if inp.freeze_active_set else L.PAIR_HYST)
        ctx.pairs_run(modes, coords=coords, g_on=cfg.g_on, g_off=cfg.g_off,
                      allow_deactivation=inp.allow_deactivation)
        return UpdateGeometryOutput(manager=m, n_active=m.n_active())


# ------------------------------------------------------------------ Strategies


class HuberContactForceProcess:
    """Huber smooth-penalty normal contact (contact/contact_force/strategy.py:69-262)."""

    def __init__(self, ndof: int, ndof_per_node: int = 6, *, smoothing_delta: float = 0.0) -> None:
        if ndof_per_node != 6:
            raise ValueError("the beam DOF layout has 6 DOFs per node")
        self._ndof = ndof
        self._ndof_per_node = ndof_per_node
        self._smoothing_delta = smoothing_delta

    @staticmethod
    def _huber(x: float, delta: float) -> float:
        if delta <= 0.0:
            return max(0.0, x)
        if x < -delta:
            return 0.0
        if x > delta:
            return x
        return (x + delta) ** 2 / (4.0 * delta)

    @staticmethod
    def _huber_deriv(x: float, delta: float) -> float:
        if delta <= 0.0:
            return 1.0 if x > 0.0 else 0.0
        if x < -delta:
            return 0.0
        if x > delta:
            return 1.0
        return (x + delta) / (2.0 * delta)

    def evaluate_device(self, u, manager: ContactManager, k_pen: float):
        ctx = manager.ctx
        ud = ctx.to_dev(u)
        ctx.pairs_run(L.PAIR_HUBER, u=ud, k_pen=k_pen, smoothing_delta=self._smoothing_delta)
        f = ctx.torch.empty(ctx.ndof, dtype=ctx.torch.float64, device=ctx.dev)
        ctx.pairs_gather(f=f)
        return f

    def evaluate(self, u, manager: ContactManager, k_pen: float):
        """(f_c, residuals): residuals = k_pen g for non-INACTIVE pairs with p_n > 0, else 0."""
        f = self.evaluate_device(u, manager, k_pen).cpu().numpy()
        d = manager.ctx.get_pairs()
        act = d["status"] != 0
        res = np.where(d["p_n"][act] > 0.0, k_pen * d["gap"][act], 0.0)
        return f, res

    def tangent(self, u, manager: ContactManager, k_pen: float) -> sp.csr_matrix:
        ctx = manager.ctx
        ctx.pairs_run(L.PAIR_TAN_C, k_pen=k_pen, smoothing_delta=self._smoothing_delta)
        K = ctx.new_K()
        ctx.pairs_gather(K_vals=K)
        return ctx.to_csr(K)


class _FrictionTangents(dict):
    """``friction_tangents``: pair index -> 2x2 D_t of the last evaluate."""


class CoulombReturnMappingProcess:
    """Coulomb friction return mapping (contact/friction/strategy.py:50-168)."""

    def __init__(self, ndof: int, ndof_per_node: int = 6, *, k_pen: float = 0.0, k_t_ratio: float = 1.0,
                 mu_ramp_counter: int = 0, mu_ramp_steps: int = 0) -> None:
        if ndof_per_node != 6:
            raise ValueError("the beam DOF layout has 6 DOFs per node")
        self._ndof = ndof
        self._ndof_per_node = ndof_per_node
        self._k_pen = k_pen
        self._k_t_ratio = k_t_ratio
        self._mu_ramp_counter = mu_ramp_counter
        self._mu_ramp_steps = mu_ramp_steps
        self._last_mu_eff = 0.0
        self._last_manager: ContactManager | None = None

    @property
    def friction_tangents(self) -> dict:
        """Pairs of the last evaluate with |q| >= 1e-30 mapped to their D_t
        (``_tangent_2x2_core``, friction/law_friction.py:126-153)."""
        m = self._last_manager
        out = _FrictionTangents()
        if m is None or self._last_mu_eff <= 0.0:
            return out
        d = m.ctx.get_pairs()
        zn = np.linalg.norm(d["z_t"], axis=1)
        sel = np.nonzero((d["p_n"] > 0.0) & (zn >= 1e-30))[0]
        for i in sel:
            kt, qtn = d["k_t"][i], d["q_trial_norm"][i]
            if d["stick"][i] or qtn < 1e-30:
                out[int(i)] = kt * np.eye(2)
            else:
                qh = d["z_t"][i] / zn[i]
                out[int(i)] = (self._last_mu_eff * d["p_n"][i] / qtn) * kt * (np.eye(2) - np.outer(qh, qh))
        return out

    def compute_k_t(self) -> float:
        return self._k_pen * self._k_t_ratio

    def compute_mu_effective(self, mu: float) -> float:
        """mu * min(1, counter / steps) (friction/law_friction.py:186-195)."""
        if self._mu_ramp_steps <= 0:
            return mu
        return mu * min(1.0, self._mu_ramp_counter / self._mu_ramp_steps)

    def set_k_pen(self, k_pen: float) -> None:
        self._k_pen = k_pen

    def set_k_t_ratio(self, k_t_ratio: float) -> None:
        self._k_t_ratio = k_t_ratio

    def set_mu_ramp_counter(self, counter: int) -> None:
        self._mu_ramp_counter = counter

    @staticmethod
    def _manager_of(contact_pairs) -> ContactManager:
        if isinstance(contact_pairs, DevicePairList):
            return contact_pairs.manager
        if isinstance(contact_pairs, ContactManager):
            return contact_pairs
        raise TypeError("contact_pairs must be the device manager's pair list")

    def evaluate_device(self, u, contact_pairs, mu: float, **kwargs):
        m = self._manager_of(contact_pairs)
        ctx = m.ctx
        mu_eff = self.compute_mu_effective(mu)
        ud = ctx.to_dev(u)
        u_ref = kwargs.get("u_ref")
        urd = ctx.to_dev(u_ref) if u_ref is not None else None
        ctx.pairs_run(L.PAIR_FRIC, u=ud, u_ref=urd, k_pen=self._k_pen, k_t_ratio=self._k_t_ratio, mu_eff=mu_eff)
        f = ctx.torch.empty(ctx.ndof, dtype=ctx.torch.float64, device=ctx.dev)
        ctx.pairs_gather(f=f)
        self._last_mu_eff = mu_eff
        self._last_manager = m
        return f

    def evaluate(self, u, contact_pairs, mu: float, **kwargs):
        """(f_fric, residuals): residuals = max(0, |q| - mu_eff p_n) over the
        pairs entering the force dictionary."""
        m = self._manager_of(contact_pairs)
        if m.ctx.n_pairs == 0:
            return np.zeros(self._ndof), np.zeros(0)
        f = self.evaluate_device(u, contact_pairs, mu, **kwargs).cpu().numpy()
        d = m.ctx.get_pairs()
        qn = np.linalg.norm(d["z_t"], axis=1)
        sel = (d["p_n"] > 0.0) & (qn >= 1e-30) if self._last_mu_eff > 0.0 else np.zeros(len(qn), bool)
        return f, np.maximum(0.0, qn[sel] - self._last_mu_eff * d["p_n"][sel])

    def tangent(self, u, contact_pairs, mu: float, **kwargs) -> sp.csr_matrix:
        """K_fm + K_fg of the last evaluate (friction/strategy.py:150-164)."""
        m = self._manager_of(contact_pairs)
        ctx = m.ctx
        ctx.pairs_run(L.PAIR_TAN_F, mu_eff=self._last_mu_eff)
        K = ctx.new_K()
        ctx.pairs_gather(K_vals=K)
        return ctx.to_csr(K)


def _coating_params(config):
    return (config.coating_stiffness, config.coating_damping, config.coating_mu, config.coating_k_t_ratio)


class NoCoatingProcess:
    """``NoCoatingProcess`` (contact/coating/strategy.py:19-77): every term is zero."""

    def forces(self, pairs, node_coords, config, dt: float) -> np.ndarray:
        return np.zeros(len(node_coords) * getattr(config, "ndof_per_node", 6))

    def stiffness(self, pairs, node_coords, config, ndof_total: int, dt: float) -> sp.csr_matrix:
        return sp.csr_matrix((ndof_total, ndof_total))

    def friction_forces(self, pairs, node_coords, config, u_cur, u_ref) -> np.ndarray:
        return np.zeros_like(u_cur)

    def friction_stiffness(self, pairs, node_coords, config, ndof_total: int) -> sp.csr_matrix:
        return sp.csr_matrix((ndof_total, ndof_total))

    def process(self, input_data) -> None:
        return None


class KelvinVoigtCoatingProcess:
    """``KelvinVoigtCoatingProcess`` (contact/coating/strategy.py:80-383): the
    coating terms of the pairs' current state (s, t, normal, frame and
    compression as the last UpdateGeometry left them), each method one pair
    kernel launch with its coating mode and one gather.  ``pairs`` is the
    device manager's pair list (or the manager); ``config`` carries the
    coating parameters; ``node_coords`` only sizes the result, as in the
    reference."""

    def _ctx(self, pairs, config, ndof=None):
        m = CoulombReturnMappingProcess._manager_of(pairs)
        ctx = m.ctx
        if config.coating_stiffness <= 0.0:
            raise ValueError("KelvinVoigtCoatingProcess needs coating_stiffness > 0 (_create_coating_strategy)")
        if ndof is not None and int(ndof) != ctx.ndof:
            raise ValueError("ndof_total differs from the device context's")
        ctx.ensure_coating(*_coating_params(config))
        return ctx

    def _force(self, ctx, modes, u=None, u_ref=None) -> np.ndarray:
        if ctx.n_pairs == 0:
            return np.zeros(ctx.ndof)
        f = ctx.torch.empty(ctx.ndof, dtype=ctx.torch.float64, device=ctx.dev)
        ctx.pairs_run(modes, u=None if u is None else ctx.to_dev(u), u_ref=None if u_ref is None else ctx.to_dev(u_ref))
        ctx.pairs_gather(f=f)
        return f.cpu().numpy()

    def _tangent(self, ctx, modes) -> sp.csr_matrix:
        if ctx.n_pairs == 0:
            return sp.csr_matrix((ctx.ndof, ctx.ndof))
        ctx.pairs_run(modes)
        K = ctx.new_K()
        ctx.pairs_gather(K_vals=K)
        return ctx.to_csr(K)

    def forces(self, pairs, node_coords, config, dt: float) -> np.ndarray:
        """Kelvin-Voigt normal force k cc + c (cc - cc_prev) / dt of every pair
        with cc > 0 or cc_prev > 0, -f_n n on A and +f_n n on B (:97-144)."""
        ctx = self._ctx(pairs, config)
        ctx.coating_set_dt(dt)
        return self._force(ctx, L.PAIR_COAT_F)

    def stiffness(self, pairs, node_coords, config, ndof_total: int, dt: float) -> sp.csr_matrix:
        """k_eff g_n g_n^T with k_eff = k + c / dt over the pairs with cc > 1e-15 (:146-204)."""
        ctx = self._ctx(pairs, config, ndof_total)
        ctx.coating_set_dt(dt)
        return self._tangent(ctx, L.PAIR_COAT_K)

    def friction_forces(self, pairs, node_coords, config, u_cur, u_ref) -> np.ndarray:
        """Coulomb return mapping of the coating history (k_t = k k_t_ratio,
        p_n = k cc) and its force; the history of separated pairs is reset (:206-297)."""
        if config.coating_mu <= 0.0:
            return np.zeros_like(u_cur)
        ctx = self._ctx(pairs, config)
        return self._force(ctx, L.PAIR_COAT_FRIC, u=u_cur, u_ref=u_ref)

    def friction_stiffness(self, pairs, node_coords, config, ndof_total: int) -> sp.csr_matrix:
        """(w_k w_l) T D T^T over the pairs with cc > 0, from the history the
        last friction_forces left (:299-379)."""
        if config.coating_mu <= 0.0:
            return sp.csr_matrix((ndof_total, ndof_total))
        ctx = self._ctx(pairs, config, ndof_total)
        return self._tangent(ctx, L.PAIR_COAT_KF)

    def process(self, input_data) -> None:
        return None


def _create_coating_strategy(*, coating_stiffness: float = 0.0):
    """Kelvin-Voigt when the stiffness is > 0, none otherwise (strategy.py:386-408)."""
    if coating_stiffness <= 0.0:
        return NoCoatingProcess()
    return KelvinVoigtCoatingProcess()
