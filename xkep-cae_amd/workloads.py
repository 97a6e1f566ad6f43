"""Synthetic strand workloads of the per-Newton-iteration assembly path.

Units are mm-ton-MPa.  Material E = 200e3 MPa, nu = 0.3, rho = 7.85e-9.
Contact: k_pen = 1e4 E r, mu = 0.3, Huber smoothing_delta = 5000 / r
(the reference's rule, numerical_tests/three_point_bend_jig.py:1121),
k_t_ratio = 0.5, g_on = 0, g_off = 1e-6.  Candidate filter: exclude_same_layer
with elem_layer_map = layer index of each element's wire (the map the
reference's MeshData.layer_ids carries, core/data.py:26), so only pairs between
layers remain; with it the rest configuration has no initial penetration,
which the reference's solver requires (contact/solver/process.py:312-324),
while wires of one layer (helices at one lay radius) and non-adjacent elements
of one wire would overlap.

State: u ~ N(0, (1e-3 r)^2) i.i.d. on all six DOFs of every node, from
``numpy.random.default_rng(seed)``; u_ref = 0; Generalized-alpha predicted
from rest (u_pred = v_pred = acc_old = 0).  Boundary: all DOFs of the first
node of every wire fixed; f_ext = end bending moment on the last node of every
wire.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import StrandMesh, make_strand

E_STEEL = 200e3
NU_STEEL = 0.3
RHO_STEEL = 7.85e-9


@dataclass(frozen=True)
class StrandCase:
    name: str
    n_wires: int
    wire_radius: float
    length: float
    n_elems_per_wire: int
    lay_angle_deg: float
    seed: int
    mu: float = 0.3
    rho_inf: float = 0.8
    end_moment: float = 0.0
    E: float = E_STEEL
    nu: float = NU_STEEL
    rho: float = RHO_STEEL

    @property
    def k_pen(self) -> float:
        return 1e4 * self.E * self.wire_radius

    @property
    def smoothing_delta(self) -> float:
        return 5000.0 / self.wire_radius

    def mesh(self) -> StrandMesh:
        return make_strand(self.n_wires, self.wire_radius, self.length, self.n_elems_per_wire,
                           lay_angle_deg=self.lay_angle_deg)

    def dt(self) -> float:
        """T1/40 of one wire's first bending mode, f1 = pi/(2 L^2) sqrt(EI/(rho A))
        (numerical_tests/three_point_bend_jig.py:841-858)."""
        r = self.wire_radius
        A = math.pi * r * r
        Iv = math.pi * (2 * r) ** 4 / 64.0
        f1 = (math.pi / (2.0 * self.length**2)) * math.sqrt(self.E * Iv / (self.rho * A))
        return 1.0 / f1 / 40.0


C1 = StrandCase("C1_7wire", 7, 1.0, 40.0, 40, 10.0, seed=7)
C2 = StrandCase("C2_37wire", 37, 0.5, 60.0, 100, 12.0, seed=37)
C3 = StrandCase("C3_91wire", 91, 0.4, 80.0, 150, 12.0, seed=91)
C5 = StrandCase("C5_1000wire", 1000, 0.3, 100.0, 100, 12.0, seed=1000)
CASES = {c.name: c for c in (C1, C2, C3, C5)}


def layer_map(elem_layer) -> dict:
    """elem_layer_map of the contact config: element -> layer index."""
    return {int(e): int(k) for e, k in enumerate(np.asarray(elem_layer))}


@dataclass
class StrandState:
    case: StrandCase
    mesh: StrandMesh
    u: np.ndarray
    u_ref: np.ndarray
    f_ext: np.ndarray
    fixed_dofs: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def elem_layer(self) -> np.ndarray:
        return self.mesh.elem_layer()


def make_state(case: StrandCase, *, seed: int | None = None) -> StrandState:
    mesh = case.mesh()
    rng = np.random.default_rng(case.seed if seed is None else seed)
    ndof = 6 * mesh.n_nodes
    u = rng.normal(0.0, 1e-3 * case.wire_radius, size=ndof)
    nn = mesh.n_elems_per_wire + 1
    first = np.arange(mesh.n_wires) * nn
    last = first + nn - 1
    fixed = (first[:, None] * 6 + np.arange(6)[None, :]).ravel()
    u[fixed] = 0.0
    f_ext = np.zeros(ndof)
    f_ext[last * 6 + 4] = case.end_moment
    return StrandState(case=case, mesh=mesh, u=u, u_ref=np.zeros(ndof), f_ext=f_ext,
                       fixed_dofs=fixed)


@dataclass
class Batch:
    """Several strand instances concatenated into one mesh.  Instance b is
    translated by ``offsets[b]`` along x so that no box of one instance meets a
    box of another (the broadphase then never pairs across instances); node,
    element and DOF ranges of instance b are ``node_ranges[b]``,
    ``elem_ranges[b]`` and 6x the node range."""

    states: list
    node_coords: np.ndarray
    connectivity: np.ndarray
    radii: np.ndarray
    u: np.ndarray
    u_ref: np.ndarray
    f_ext: np.ndarray
    fixed_dofs: np.ndarray
    elem_layer: np.ndarray
    node_ranges: list
    elem_ranges: list
    offsets: np.ndarray

    @property
    def n_instances(self) -> int:
        return len(self.states)


def concat_states(states: list, gap_factor: float = 4.0) -> Batch:
    xs, cs, rs, us, urs, fs, fx, ly = [], [], [], [], [], [], [], []
    nr, er, offs = [], [], []
    n0 = e0 = 0
    shift = 0.0
    for st in states:
        m = st.mesh
        X = m.node_coords.copy()
        lo, hi = X.min(axis=0), X.max(axis=0)
        pad = gap_factor * (2.0 * m.wire_radius + (hi[2] - lo[2]) / m.n_elems_per_wire)
        dx = shift - lo[0] + pad
        X[:, 0] += dx
        shift = X[:, 0].max() + pad
        offs.append(dx)
        xs.append(X)
        cs.append(m.connectivity + n0)
        rs.append(m.radii)
        us.append(st.u)
        urs.append(st.u_ref)
        fs.append(st.f_ext)
        fx.append(st.fixed_dofs + 6 * n0)
        ly.append(m.elem_layer())
        nr.append((n0, n0 + m.n_nodes))
        er.append((e0, e0 + m.n_elems))
        n0 += m.n_nodes
        e0 += m.n_elems
    return Batch(states=states, node_coords=np.concatenate(xs), connectivity=np.concatenate(cs),
                 radii=np.concatenate(rs), u=np.concatenate(us), u_ref=np.concatenate(urs),
                 f_ext=np.concatenate(fs), fixed_dofs=np.concatenate(fx),
                 elem_layer=np.concatenate(ly), node_ranges=nr, elem_ranges=er,
                 offsets=np.array(offs))


def squeeze_case(force: float, *, n_wires: int = 7, n_elems: int = 20, length: float = 20.0, radius: float = 1.0,
                 lay_angle_deg: float = 10.0, mu: float = 0.0) -> dict:
    """Dynamic radial squeeze of a strand: first node of every wire clamped, a
    radial inward force at the mid node of every outer wire, lumped mass,
    Generalized-alpha rho_inf = 0.8 over t_total = 4 T1/40 of one wire, steps
    capped at a quarter of it, penalty 0.5 c0 m_ii at that step (the dynamic
    estimate of contact/penalty/strategy.py:216-262), Huber delta 5000/r,
    same-layer exclusion by layer index, frictionless by default."""
    m = make_strand(n_wires, radius, length, n_elems, lay_angle_deg=lay_angle_deg)
    X, conn = m.node_coords, m.connectivity
    nn = n_elems + 1
    first = np.arange(n_wires) * nn
    fixed = (first[:, None] * 6 + np.arange(6)[None, :]).ravel()
    mid = first[1:] + nn // 2
    nrm = X[mid, :2] / np.linalg.norm(X[mid, :2], axis=1)[:, None]
    f = np.zeros(6 * len(X))
    f[mid * 6] = -force * nrm[:, 0]
    f[mid * 6 + 1] = -force * nrm[:, 1]
    A = math.pi * radius * radius
    Iv = math.pi * (2 * radius) ** 4 / 64.0
    f1 = (math.pi / (2.0 * length**2)) * math.sqrt(E_STEEL * Iv / (RHO_STEEL * A))
    t_total = 4.0 * (1.0 / f1 / 40.0)
    rho_inf, dt_max_fraction = 0.8, 0.25
    am = (2 * rho_inf - 1) / (rho_inf + 1)
    af = rho_inf / (rho_inf + 1)
    beta = 0.25 * (1 - am + af) ** 2
    dt_nom = t_total * dt_max_fraction
    k_pen = 0.5 * (1.0 / (beta * dt_nom**2)) * RHO_STEEL * A * (length / n_elems) / 2.0
    return dict(mesh=m, X=X, conn=conn, radii=m.radii, layer=m.elem_layer(), fixed=fixed, f=f, t_total=t_total,
                rho_inf=rho_inf, k_pen=k_pen, mu=mu, sd=5000.0 / radius, ktr=0.5,
                cfg=dict(max_nr_attempts=30, tol_force=1e-6, tol_disp=1e-8, dt_max_fraction=dt_max_fraction))


def uneven_batch(n_instances: int = 64, seed: int = 64):
    """The uneven batched workload assembled as one: the cases, their states
    (end moment on the last node of every wire), the concatenated batch and
    the per-element friction coefficient of each element's instance."""
    cases = batched_cases(n_instances, seed)
    states = [make_state(c) for c in cases]
    batch = concat_states(states)
    mu_elem = np.concatenate([np.full(st.mesh.n_elems, c.mu) for c, st in zip(cases, states)])
    return cases, states, batch, mu_elem


def batched_cases(n_instances: int = 64, seed: int = 64) -> list[StrandCase]:
    """The uneven batched workload: wires in {7, 19, 37, 61}, 50-150 elements per
    wire, lay angle U(8, 15) deg, mu U(0.1, 0.5), random end moment, r = 0.5,
    L = 60."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_instances):
        nw = int(rng.choice([7, 19, 37, 61]))
        m = int(rng.integers(50, 151))
        ang = float(rng.uniform(8.0, 15.0))
        mu = float(rng.uniform(0.1, 0.5))
        mom = float(rng.uniform(-1.0, 1.0)) * 10.0
        out.append(StrandCase(f"B{k}_{nw}w{m}", nw, 0.5, 60.0, m, ang, seed=seed * 1000 + k,
                              mu=mu, end_moment=mom))
    return out
