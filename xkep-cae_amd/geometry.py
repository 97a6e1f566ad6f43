"""Drop-in ``ContactGeometryStrategy`` on the GPU.

Mirrors xkep_cae/contact/geometry/strategy.py behind the Protocol of
xkep_cae/core/strategies/protocols.py:107-147:

  ``PointToPointProcess``      :311-371   detect / compute_gap / update_geometry /
                                          build_constraint_jacobian / process
  ``LineToLineGaussProcess``   :374-473   the same geometry (its Gauss-point count
                                          is only recorded on pair objects that carry
                                          one, which the pair type does not)
  ``MortarSegmentProcess``     :476-536   the same geometry
  ``_create_contact_geometry_strategy``   :539-579
  ``ContactGeometryInput`` / ``ContactGeometryOutput``  :27-40

``detect`` is the Strategy twin of candidate detection (``_detect_candidates``,
:135-199): the GPU broadphase with the shared-node filter only (the Strategy's
``exclude_same_layer``), no layer filter, no midpoint prescreen, fresh pair
states with core radii 0.0, then the initial narrowphase without a config
(g_on = g_off = 0).  ``update_geometry`` runs the pair kernel's narrowphase,
frame transport and hysteresis with g_on / g_off / coating_stiffness from the
config and deactivation always allowed (``_batch_update_geometry``,
:205-305).  ``build_constraint_jacobian`` is a device kernel
(``xk_pairs_constraint_jacobian``).  The pairs stay in HBM: ``detect`` returns
a sequence view of the device pair list, and the methods run on it in place.
A plain list of pair objects (attributes elem_a, elem_b, state.*) of the same
mesh is accepted too: it is uploaded, updated on the device and written back
element by element, as the reference mutates its list.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Protocol, runtime_checkable

import numpy as np

from . import _lib as L
from .contact import ContactManager, ContactPair, ContactState, ContactStatus, DevicePairList
from .context import DeviceContext, Section

# the geometry path reads no section property; the context still carries one
_NO_SECTION = Section(1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0)


@dataclass(frozen=True)
class ContactGeometryInput:
    node_coords: np.ndarray
    connectivity: np.ndarray
    radii: Any


@dataclass(frozen=True)
class ContactGeometryOutput:
    contact_pairs: list


@runtime_checkable
class ContactGeometryStrategy(Protocol):
    """core/strategies/protocols.py:107-147."""

    def detect(self, node_coords, connectivity, radii) -> list: ...

    def compute_gap(self, pair, node_coords) -> float: ...

    def update_geometry(self, pairs, node_coords, *, config=None) -> None: ...

    def build_constraint_jacobian(self, pairs, ndof_total: int, ndof_per_node: int = 6): ...


class StrategyPairList(DevicePairList):
    """The pair list a geometry Strategy's ``detect`` returns: a read-only
    sequence view of the device pair list of the Strategy's context."""


def _cfg_value(config, name: str, default: float) -> float:
    return float(getattr(config, name, default)) if config is not None else default


class PointToPointProcess:
    """Closest-point pair contact geometry on the GPU (geometry/strategy.py:311-371)."""

    def __init__(self, *, exclude_same_layer: bool = True, precision: str = "fp64") -> None:
        self._exclude_same_layer = exclude_same_layer
        self._precision = precision
        self._ctx: DeviceContext | None = None

    # ------------------------------------------------------------ helpers
    def _context_for(self, node_coords, connectivity, radii) -> DeviceContext:
        X = np.ascontiguousarray(node_coords, dtype=np.float64)
        conn = np.ascontiguousarray(connectivity, dtype=np.int32).reshape(-1, 2)
        ne = len(conn)
        r = np.full(ne, float(radii)) if np.isscalar(radii) else np.asarray(radii, float)
        # a new context per detect: every detect returns an independent pair list
        return DeviceContext(X, conn, _NO_SECTION, radii=r, precision=self._precision)

    def _device_list(self, pairs) -> tuple[DeviceContext, bool]:
        """The context holding ``pairs`` on the device, and whether they had to
        be uploaded (a plain list of pair objects of the last detected mesh)."""
        if isinstance(pairs, DevicePairList):
            return pairs.manager.ctx, False
        ctx = self._ctx
        if ctx is None:
            raise ValueError("update a pair list of this Strategy's mesh: call detect first")
        pl = list(pairs)
        ea = np.array([p.elem_a for p in pl], np.int32)
        eb = np.array([p.elem_b for p in pl], np.int32)
        st = [p.state for p in pl]
        if len(pl) and (ea.max() >= ctx.n_elems or eb.max() >= ctx.n_elems):
            raise ValueError("pair elements outside the detected mesh")
        state = {
            "s": [x.s for x in st], "t": [x.t for x in st], "gap": [x.gap for x in st],
            "normal": np.array([x.normal for x in st], float).reshape(-1, 3),
            "tangent1": np.array([x.tangent1 for x in st], float).reshape(-1, 3),
            "p_n": [x.p_n for x in st], "k_pen": [x.k_pen for x in st], "k_t": [x.k_t for x in st],
            "q_trial_norm": [x.q_trial_norm for x in st], "dissipation": [x.dissipation for x in st],
            "z_t": np.array([x.z_t for x in st], float).reshape(-1, 2),
            "status": [int(getattr(x.status, "value", x.status)) for x in st], "stick": [bool(x.stick) for x in st],
        }
        ctx.set_pairs(ea, eb, state)
        return ctx, True

    @staticmethod
    def _write_back(pairs, ctx: DeviceContext) -> None:
        snap = StrategyPairList(ContactManager(ctx))._materialise()
        for i, p in enumerate(snap):
            pairs[i] = p

    # ------------------------------------------------------------ Protocol
    def detect(self, node_coords, connectivity, radii) -> list:
        """Broadphase AABB + shared-node filter + initial narrowphase (:135-199)."""
        ctx = self._context_for(node_coords, connectivity, radii)
        self._ctx = ctx
        coords = ctx.X
        if ctx.n_elems >= 2:
            ctx.strategy_detect(coords, shared_node_filter=self._exclude_same_layer)
            if ctx.n_pairs:
                ctx.pairs_run(L.PAIR_GEOM | L.PAIR_HYST, coords=coords, g_on=0.0, g_off=0.0, allow_deactivation=True)
        return StrategyPairList(ContactManager(ctx))

    def compute_gap(self, pair, node_coords) -> float:
        """The stored gap of the pair (geometry/strategy.py:339-343)."""
        if hasattr(pair, "state") and hasattr(pair.state, "gap"):
            return float(pair.state.gap)
        return 0.0

    def update_geometry(self, pairs, node_coords, *, config=None) -> None:
        """Narrowphase, frame and hysteresis of every pair, in place (:205-305)."""
        if len(pairs) == 0:
            return
        ctx, uploaded = self._device_list(pairs)
        cs = _cfg_value(config, "coating_stiffness", 0.0)
        if cs > 0.0:
            # the Strategy's pairs carry core radii 0.0: gap = dist, compression = r_a + r_b - dist
            ctx.ensure_coating(cs, 0.0, 0.0, 0.5, np.zeros(ctx.n_elems))
        elif ctx.coat_on:
            ctx.set_coating(0.0, core_radii=np.zeros(ctx.n_elems))
        coords = ctx.to_dev(np.ascontiguousarray(node_coords, dtype=np.float64))
        ctx.pairs_run(L.PAIR_GEOM | L.PAIR_HYST, coords=coords, g_on=_cfg_value(config, "g_on", 0.0),
                      g_off=_cfg_value(config, "g_off", 0.0), allow_deactivation=True)
        if uploaded:
            self._write_back(pairs, ctx)
        else:
            pairs._snap = None

    def build_constraint_jacobian(self, pairs, ndof_total: int, ndof_per_node: int = 6):
        """(G, active_indices) with G = d g_n / d u (:83-132)."""
        ctx, _ = self._device_list(pairs)
        return ctx.constraint_jacobian(ndof_total, ndof_per_node)

    def process(self, input_data: ContactGeometryInput) -> ContactGeometryOutput:
        return ContactGeometryOutput(
            contact_pairs=self.detect(input_data.node_coords, input_data.connectivity, input_data.radii))


class LineToLineGaussProcess(PointToPointProcess):
    """Line-to-line Gauss variant (geometry/strategy.py:374-473); its geometry
    update is the point-to-point one."""

    def __init__(self, *, n_gauss: int = 2, exclude_same_layer: bool = True, auto_gauss: bool = False,
                 precision: str = "fp64") -> None:
        super().__init__(exclude_same_layer=exclude_same_layer, precision=precision)
        self._n_gauss = n_gauss
        self._auto_gauss = auto_gauss


class MortarSegmentProcess(PointToPointProcess):
    """Mortar segment variant (geometry/strategy.py:476-536); its geometry
    update and Jacobian are the point-to-point ones."""

    def __init__(self, *, n_gauss: int = 2, exclude_same_layer: bool = True, precision: str = "fp64") -> None:
        super().__init__(exclude_same_layer=exclude_same_layer, precision=precision)
        self._n_gauss = n_gauss


def _create_contact_geometry_strategy(*, mode: str = "point_to_point", exclude_same_layer: bool = True,
                                      n_gauss: int = 2, auto_gauss: bool = False, line_contact: bool = False,
                                      use_mortar: bool = False):
    """Factory with the reference's precedence: use_mortar, then line_contact,
    then mode (geometry/strategy.py:539-579)."""
    if use_mortar:
        mode = "mortar"
    elif line_contact:
        mode = "line_to_line"
    if mode == "mortar":
        return MortarSegmentProcess(n_gauss=n_gauss, exclude_same_layer=exclude_same_layer)
    if mode == "line_to_line":
        return LineToLineGaussProcess(n_gauss=n_gauss, exclude_same_layer=exclude_same_layer, auto_gauss=auto_gauss)
    return PointToPointProcess(exclude_same_layer=exclude_same_layer)


__all__ = ["ContactGeometryInput", "ContactGeometryOutput", "ContactGeometryStrategy", "StrategyPairList",
           "PointToPointProcess", "LineToLineGaussProcess", "MortarSegmentProcess",
           "_create_contact_geometry_strategy", "ContactPair", "ContactState", "ContactStatus"]
