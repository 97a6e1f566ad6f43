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


def min_safe_gap(n_wires, wire_radius, pitch, n_elems, n_pitches, safety=2.0) -> float:
    if n_wires <= 1:
        return 0.0
    length = n_pitches * pitch
    n_per_pitch = n_elems * pitch / length if length > 0 else n_elems
    if n_per_pitch < 1:
        return 0.0
    d_eff = 2.0 * wire_radius
    sag = d_eff * (1.0 - math.cos(math.pi / n_per_pitch))
    return 2.0 * sag * safety


def make_strand(n_wires: int, wire_radius: float, length: float, n_elems_per_wire: int, *,
                lay_angle_deg: float | None = None, pitch: float | None = None,
                gap: float = 0.0, lay_direction: int = 1) -> StrandMesh:
    if pitch is None:
        if lay_angle_deg is None:
            raise ValueError("give pitch or lay_angle_deg")
        pitch = pitch_for_lay_angle(n_wires, wire_radius, lay_angle_deg)
    if n_wires > 1 and length > 0 and pitch > 0:
        if n_elems_per_wire * pitch / length < 16:
            raise ValueError("fewer than 16 elements per pitch")
        g_min = min_safe_gap(n_wires, wire_radius, pitch, n_elems_per_wire, length / pitch)
        if gap < g_min:
            gap = g_min
    d = 2.0 * wire_radius
    counts = layer_counts(n_wires)
    wires = []  # (layer, offset, r_lay, dir)
    if n_wires == 3:
        r_lay = (d + gap) / math.sqrt(3.0)
        wires = [(1, 2.0 * math.pi * k / 3.0, r_lay, lay_direction) for k in range(3)]
    else:
        wires.append((0, 0.0, 0.0, 0))
        for k in range(1, len(counts)):
            r_lay = k * (d + gap)
            direc = lay_direction if (k % 2 == 1) else -lay_direction
            for j in range(counts[k]):
                wires.append((k, 2.0 * math.pi * j / (6 * k), r_lay, direc))
    nn = n_elems_per_wire + 1
    X = np.zeros((n_wires * nn, 3))
    conn = np.zeros((n_wires * n_elems_per_wire, 2), np.int64)
    for w, (layer, off, r_lay, direc) in enumerate(wires):
        pts = np.zeros((nn, 3))
        if r_lay < 1e-15:
            pts[:, 2] = np.linspace(0.0, length, nn)
        else:
            for i in range(nn):
                z = i * length / (nn - 1)
                th = direc * 2.0 * math.pi * z / pitch + off
                pts[i] = (r_lay * math.cos(th), r_lay * math.sin(th), z)
        X[w * nn:(w + 1) * nn] = pts
        e0 = w * n_elems_per_wire
        conn[e0:e0 + n_elems_per_wire, 0] = w * nn + np.arange(n_elems_per_wire)
        conn[e0:e0 + n_elems_per_wire, 1] = w * nn + np.arange(1, nn)
    return StrandMesh(node_coords=X, connectivity=conn,
                      wire_layer=np.array([w[0] for w in wires], np.int64),
                      n_wires=n_wires, n_elems_per_wire=n_elems_per_wire,
                      wire_radius=wire_radius, pitch=pitch, length=length, gap=gap)
