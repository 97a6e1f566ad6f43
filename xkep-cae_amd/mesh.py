"""Helical strand mesh generator (inputs of the assembly path).

Same semantics as the reference's ``_make_twisted_wire_mesh``
(xkep_cae/mesh/_twisted_wire.py:271-383) with the layout of
``_make_strand_layout`` (:150-263, the ``strand_diameter=None`` branch) and the
chord-sagitta safe gap of ``_compute_min_safe_gap`` (:127-147):

* a straight core wire (z from ``np.linspace``), then layers k = 1, 2, ... of
  ``min(6k, remaining)`` wires at lay radius ``k (d + gap)``, angle offsets
  ``2 pi j / (6k)``, lay direction alternating with layer parity;
* helix node i at ``z = i L/(n-1)``, ``theta = dir 2 pi z / pitch + offset``;
* every wire is a chain of ``n_elems`` 2-node elements, wires numbered
  core-first, nodes and elements contiguous per wire;
* ``gap`` is raised to ``2 * safety * d (1 - cos(pi / n_per_pitch))`` when
  smaller, and fewer than 16 elements per pitch is refused.
"""

from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class StrandMesh:
    node_coords: np.ndarray      # (n_nodes, 3) float64
    connectivity: np.ndarray     # (n_elems, 2) int64
    wire_layer: np.ndarray       # (n_wires,) layer index (0 = core)
    n_wires: int
    n_elems_per_wire: int
    wire_radius: float
    pitch: float
    length: float
    gap: float

    @property
    def n_nodes(self) -> int:
        return self.node_coords.shape[0]

    @property
    def n_elems(self) -> int:
        return self.connectivity.shape[0]

    @property
    def radii(self) -> np.ndarray:
        return np.full(self.n_elems, self.wire_radius)

    def elem_layer(self) -> np.ndarray:
        return np.repeat(self.wire_layer, self.n_elems_per_wire)


def layer_counts(n_wires: int) -> list[int]:
    """Wires per layer, core first (``_compute_layer_structure``, _twisted_wire.py:85-97)."""
    if n_wires == 3:
        return [0, 3]
    out = [1]
    rem = n_wires - 1
    k = 1
    while rem > 0:
        take = min(6 * k, rem)
        out.append(take)
        rem -= take
        k += 1
    return out


def pitch_for_lay_angle(n_wires: int, wire_radius: float, lay_angle_deg: float) -> float:
    """Pitch giving the outermost layer the stated lay angle at zero gap:
    2 pi r_lay,out / tan(alpha), r_lay,out = n_layers * d."""
    n_layers = len(layer_counts(n_wires)) - 1
    r_out = n_layers * 2.0 * wire_radius
    return 2.0 * math.pi * r_out / math.tan(math.radians(lay_angle_deg))


def min_safe_gap(n_wires, wire_radius, pitch, n_elems, n_pitches, safety=2.0, coating_thickness=0.0) -> float:
    """``_compute_min_safe_gap`` (_twisted_wire.py:127-147): twice the chord
    sagitta d_eff (1 - cos(pi / n_per_pitch)) times the safety factor."""
    if n_wires <= 1:
        return 0.0
    length = n_pitches * pitch
    n_per_pitch = n_elems * pitch / length if length > 0 else n_elems
    if n_per_pitch < 1:
        return 0.0
    d_eff = 2.0 * (wire_radius + coating_thickness)
    sag = d_eff * (1.0 - math.cos(math.pi / n_per_pitch))
    return 2.0 * sag * safety


@dataclass(frozen=True)
class StrandInfo:
    """One wire of the layout (``StrandInfoOutput``, _twisted_wire.py:17-26)."""

    strand_id: int
    layer: int
    angle_offset: float
    lay_radius: float
    wire_radius: float
    lay_direction: int = 1


@dataclass(frozen=True)
class TwistedWireMesh:
    """``TwistedWireMeshOutput`` (_twisted_wire.py:29-41)."""

    node_coords: np.ndarray
    connectivity: np.ndarray
    strand_node_ranges: tuple
    strand_elem_ranges: tuple
    strand_infos: tuple
    n_strands: int
    wire_radius: float
    pitch: float
    length: float
    n_elems_per_strand: int


def minimum_strand_diameter(n_strands: int, wire_diameter: float) -> float:
    """Smallest outer diameter that holds the layers (``_minimum_strand_diameter``,
    _twisted_wire.py:100-115): layer k at max(r_{k-1} + d, r / sin(pi / n_k))."""
    r = wire_diameter / 2.0
    d = wire_diameter
    if n_strands == 1:
        return wire_diameter
    if n_strands == 3:
        return 2.0 * (d / math.sqrt(3.0) + r)
    counts = layer_counts(n_strands)
    r_lay = [0.0]
    for k in range(1, len(counts)):
        n_k = counts[k]
        r_lay.append(max(r_lay[k - 1] + d, r / math.sin(math.pi / n_k) if n_k >= 2 else 0.0))
    return 2.0 * (r_lay[-1] + r)


def strand_layout(n_strands: int, wire_radius: float, *, gap: float = 0.0, lay_direction: int = 1,
                  strand_diameter: float | None = None, coating_thickness: float = 0.0) -> list[StrandInfo]:
    """Wire layout (``_make_strand_layout``, _twisted_wire.py:150-268): three
    wires around the axis; else a straight core and layers of min(6k, rest)
    wires, lay direction alternating with the layer's parity.  Without
    strand_diameter layer k sits at k (d_eff + gap), d_eff = 2 (r + coating);
    with it, each layer at its minimum radius (radial d_eff, circumferential
    r_eff / sin(pi / n_k)) plus k times an equal share of the surplus."""
    d = 2.0 * wire_radius
    d_eff = 2.0 * (wire_radius + coating_thickness)
    infos: list[StrandInfo] = []
    if n_strands == 3:
        if strand_diameter is not None:
            _check_strand_diameter(n_strands, d, strand_diameter)
            r_lay = strand_diameter / 2.0 - wire_radius
        else:
            r_lay = (d_eff + gap) / math.sqrt(3.0)
        return [StrandInfo(k, 1, 2.0 * math.pi * k / 3.0, r_lay, wire_radius, lay_direction) for k in range(3)]
    if n_strands >= 1:
        infos.append(StrandInfo(0, 0, 0.0, 0.0, wire_radius, 0))
    counts = layer_counts(n_strands) if n_strands >= 1 else [0]
    if strand_diameter is not None:
        _check_strand_diameter(n_strands, d, strand_diameter)
        n_layers = len(counts) - 1
        r_eff = wire_radius + coating_thickness
        r_min = [0.0]
        for k in range(1, n_layers + 1):
            n_k = counts[k]
            r_min.append(max(r_min[k - 1] + d_eff, r_eff / math.sin(math.pi / n_k) if n_k >= 2 else 0.0))
        surplus = (strand_diameter / 2.0 - wire_radius) - r_min[-1]
        if surplus > 1e-15 and n_layers > 0:
            delta = surplus / n_layers
            lay_radii = [0.0] + [r_min[k] + k * delta for k in range(1, n_layers + 1)]
        else:
            lay_radii = r_min
        for k in range(1, n_layers + 1):
            n_k = counts[k]
            direc = lay_direction if (k % 2 == 1) else -lay_direction
            for j in range(n_k):
                angle = 2.0 * math.pi * j / (6 * k) if n_k == 6 * k else 2.0 * math.pi * j / n_k
                infos.append(StrandInfo(len(infos), k, angle, lay_radii[k], wire_radius, direc))
        return infos
    for k in range(1, len(counts)):
        direc = lay_direction if (k % 2 == 1) else -lay_direction
        for j in range(counts[k]):
            infos.append(StrandInfo(len(infos), k, 2.0 * math.pi * j / (6 * k), k * (d_eff + gap), wire_radius,
                                    direc))
    return infos


def _check_strand_diameter(n_strands: int, wire_diameter: float, strand_diameter: float) -> None:
    """``_validate_strand_geometry`` (_twisted_wire.py:118-124)."""
    min_d = minimum_strand_diameter(n_strands, wire_diameter)
    if strand_diameter < min_d - 1e-15 * min_d:
        raise ValueError(f"strand diameter {strand_diameter:.6f} is too small for {n_strands} wires of diameter "
                         f"{wire_diameter:.6f}; minimum {min_d:.6f}")


def _wire_points(info: StrandInfo, n_points: int, pitch: float, length: float) -> np.ndarray:
    """``_straight_points`` / ``_helix_points`` (_twisted_wire.py:61-82): node i at
    z = i L / (n - 1), theta = dir 2 pi z / pitch + offset."""
    pts = np.zeros((n_points, 3))
    if info.lay_radius < 1e-15:
        pts[:, 2] = np.linspace(0.0, length, n_points)
        return pts
    for i in range(n_points):
        z = i * length / (n_points - 1)
        th = info.lay_direction * 2.0 * math.pi * z / pitch + info.angle_offset
        pts[i] = (info.lay_radius * math.cos(th), info.lay_radius * math.sin(th), z)
    return pts


def make_twisted_wire_mesh(n_strands: int, wire_diameter: float, pitch: float, length: float,
                           n_elems_per_strand: int, *, gap: float = 0.0, lay_direction: int = 1,
                           n_pitches: float | None = None, strand_diameter: float | None = None,
                           min_elems_per_pitch: int = 16, coating_thickness: float = 0.0) -> TwistedWireMesh:
    """``_make_twisted_wire_mesh`` (_twisted_wire.py:271-383): same arguments,
    checks (fewer than 16 elements per pitch refused; the gap raised to the
    chord-sagitta safe gap with a warning), layout and numbering (wires in
    layout order, nodes and elements contiguous per wire)."""
    wire_radius = wire_diameter / 2.0
    if n_pitches is not None:
        length = n_pitches * pitch
    if min_elems_per_pitch < 16:
        raise ValueError(f"min_elems_per_pitch={min_elems_per_pitch}: fewer than 16 elements per pitch "
                         "is not allowed")
    if n_strands > 1 and length > 0 and pitch > 0:
        n_per_pitch = n_elems_per_strand * pitch / length
        if n_per_pitch < min_elems_per_pitch:
            raise ValueError(f"{n_per_pitch:.1f} elements per pitch (at least {min_elems_per_pitch}); use "
                             f"n_elems_per_strand >= {math.ceil(min_elems_per_pitch * length / pitch)}")
    if n_strands > 1 and strand_diameter is None and gap >= 0:
        n_p = n_pitches if n_pitches is not None else (length / pitch if pitch > 0 else 1.0)
        g_min = min_safe_gap(n_strands, wire_radius, pitch, n_elems_per_strand, n_p,
                             coating_thickness=coating_thickness)
        if gap < g_min:
            warnings.warn(f"gap {gap:.6f} < minimum safe gap {g_min:.6f} (chord error): raised to {g_min:.6f}",
                          stacklevel=2)
            gap = g_min
    layout = strand_layout(n_strands, wire_radius, gap=gap, lay_direction=lay_direction,
                           strand_diameter=strand_diameter, coating_thickness=coating_thickness)
    nn = n_elems_per_strand + 1
    X = np.zeros((n_strands * nn, 3))
    conn = np.zeros((n_strands * n_elems_per_strand, 2), dtype=int)
    node_ranges, elem_ranges = [], []
    for w, info in enumerate(layout):
        n0, e0 = w * nn, w * n_elems_per_strand
        X[n0:n0 + nn] = _wire_points(info, nn, pitch, length)
        conn[e0:e0 + n_elems_per_strand, 0] = n0 + np.arange(n_elems_per_strand)
        conn[e0:e0 + n_elems_per_strand, 1] = n0 + np.arange(1, nn)
        node_ranges.append((n0, n0 + nn))
        elem_ranges.append((e0, e0 + n_elems_per_strand))
    return TwistedWireMesh(node_coords=X, connectivity=conn, strand_node_ranges=tuple(node_ranges),
                           strand_elem_ranges=tuple(elem_ranges), strand_infos=tuple(layout), n_strands=n_strands,
                           wire_radius=wire_radius, pitch=pitch, length=length,
                           n_elems_per_strand=n_elems_per_strand)


def make_strand(n_wires: int, wire_radius: float, length: float, n_elems_per_wire: int, *,
                lay_angle_deg: float | None = None, pitch: float | None = None,
                gap: float = 0.0, lay_direction: int = 1) -> StrandMesh:
    """The workloads' strand: ``make_twisted_wire_mesh`` with the pitch given
    directly or by the outer layer's lay angle, as a ``StrandMesh``."""
    if pitch is None:
        if lay_angle_deg is None:
            raise ValueError("give pitch or lay_angle_deg")
        pitch = pitch_for_lay_angle(n_wires, wire_radius, lay_angle_deg)
    if n_wires > 1 and length > 0 and pitch > 0:
        gap = max(gap, min_safe_gap(n_wires, wire_radius, pitch, n_elems_per_wire, length / pitch))
    m = make_twisted_wire_mesh(n_wires, 2.0 * wire_radius, pitch, length, n_elems_per_wire, gap=gap,
                               lay_direction=lay_direction)
    return StrandMesh(node_coords=m.node_coords, connectivity=m.connectivity.astype(np.int64),
                      wire_layer=np.array([i.layer for i in m.strand_infos], np.int64),
                      n_wires=n_wires, n_elems_per_wire=n_elems_per_wire,
                      wire_radius=wire_radius, pitch=pitch, length=length, gap=gap)


# ---------------------------------------------------------------- StrandMeshProcess
@dataclass(frozen=True)
class MeshData:
    """The mesh fields ``StrandMeshProcess`` hands on (core/data.py MeshData)."""

    node_coords: np.ndarray
    connectivity: np.ndarray
    radii: np.ndarray
    n_strands: int
    layer_ids: np.ndarray | None = None


@dataclass(frozen=True)
class StrandMeshConfig:
    """``StrandMeshConfig`` (mesh/process.py:15-28), same fields and defaults."""

    n_strands: int = 7
    wire_radius: float = 0.5
    pitch_length: float = 50.0
    gap: float = 0.0
    n_elements_per_pitch: int = 16
    n_pitches: float = 1.0
    core_radius: float | None = None
    coating_thickness: float = 0.0
    sheath_thickness: float = 0.0


@dataclass(frozen=True)
class StrandMeshResult:
    mesh: MeshData
    core_radii: np.ndarray | float | None = None


class StrandMeshProcess:
    """``StrandMeshProcess.process`` (mesh/process.py:37-80): length = pitch x
    n_pitches, elements per wire = int(per pitch x n_pitches)."""

    name = "StrandMesh"
    module = "pre"

    def process(self, input_data: StrandMeshConfig) -> StrandMeshResult:
        c = input_data
        length = c.pitch_length * c.n_pitches
        n_elems = int(c.n_elements_per_pitch * c.n_pitches)
        m = make_twisted_wire_mesh(c.n_strands, 2.0 * c.wire_radius, c.pitch_length, length, n_elems, gap=c.gap,
                                   n_pitches=c.n_pitches, coating_thickness=c.coating_thickness)
        layer_of_elem = np.repeat([i.layer for i in m.strand_infos], n_elems)
        data = MeshData(node_coords=m.node_coords, connectivity=m.connectivity,
                        radii=np.full(m.connectivity.shape[0], m.wire_radius), n_strands=c.n_strands,
                        layer_ids=layer_of_elem)
        return StrandMeshResult(mesh=data, core_radii=getattr(m, "core_radii", None))
