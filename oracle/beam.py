"""Oracle: corotational (CR) Timoshenko 3-D beam, FP64 NumPy.

Test infrastructure only (see ``oracle/__init__.py``).

Restates, element-batched, the reference's
``xkep_cae/elements/_beam_cr.py`` (batch helpers :672-998, scalar :100-183,
:267-296, :326-664) and ``xkep_cae/elements/_beam_assembly.py:211-325``
(``_assemble_cr_beam3d_batch_impl``) plus the CSR composition of
``assemble_cr_beam3d`` (:121-152) and ``ULCRBeamAssembler.assemble_mass``
(``_beam_assembler.py:148-168``).

DOF layout per node: [ux, uy, uz, thx, thy, thz]; element DOFs are
[node1 (6), node2 (6)].
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp


# --------------------------------------------------------------------------
# section
# --------------------------------------------------------------------------


@dataclass(frozen=True)
class Section:
    """Material + section constants for one beam family."""

    E: float
    G: float
    A: float
    Iy: float
    Iz: float
    J: float
    kappa_y: float
    kappa_z: float


def circle_section(radius: float, E: float, nu: float) -> Section:
    """Solid circle: ``BeamSectionInput.circle`` (_beam_section.py:129-135) and the
    Cowper factor ``6(1+nu)/(7+6nu)`` (_beam_section.py:59-66)."""
    d = 2.0 * radius
    A = math.pi * radius**2
    Iv = math.pi * d**4 / 64.0
    J = math.pi * d**4 / 32.0
    kap = 6.0 * (1.0 + nu) / (7.0 + 6.0 * nu)
    G = E / (2.0 * (1.0 + nu))
    return Section(E=E, G=G, A=A, Iy=Iv, Iz=Iv, J=J, kappa_y=kap, kappa_z=kap)


# --------------------------------------------------------------------------
# small 3x3 helpers
# --------------------------------------------------------------------------


def skew(v: np.ndarray) -> np.ndarray:
    """[v]x for a batch (N,3) -> (N,3,3) (_beam_cr.py:672-682)."""
    n = v.shape[0]
    S = np.zeros((n, 3, 3))
    S[:, 0, 1] = -v[:, 2]
    S[:, 0, 2] = v[:, 1]
    S[:, 1, 0] = v[:, 2]
    S[:, 1, 2] = -v[:, 0]
    S[:, 2, 0] = -v[:, 1]
    S[:, 2, 1] = v[:, 0]
    return S


def quat_to_rotmat(w, x, y, z) -> np.ndarray:
    """Unit quaternion -> rotation matrix, the form of _beam_cr.py:340-357/:705-726."""
    n = w.shape[0]
    R = np.empty((n, 3, 3))
    xx, yy, zz = x * x, y * y, z * z
    xy, xz, yz = x * y, x * z, y * z
    wx, wy, wz = w * x, w * y, w * z
    R[:, 0, 0] = 1.0 - 2.0 * (yy + zz)
    R[:, 0, 1] = 2.0 * (xy - wz)
    R[:, 0, 2] = 2.0 * (xz + wy)
    R[:, 1, 0] = 2.0 * (xy + wz)
    R[:, 1, 1] = 1.0 - 2.0 * (xx + zz)
    R[:, 1, 2] = 2.0 * (yz - wx)
    R[:, 2, 0] = 2.0 * (xz - wy)
    R[:, 2, 1] = 2.0 * (yz + wx)
    R[:, 2, 2] = 1.0 - 2.0 * (xx + yy)
    return R


def rotvec_to_rotmat(th: np.ndarray) -> np.ndarray:
    """Exponential map via quaternion, Taylor branch below |th| < 1e-12
    (``_batch_rotvec_to_rotmat``, _beam_cr.py:685-726)."""
    ang = np.linalg.norm(th, axis=1)
    half = ang / 2.0
    small = ang < 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        sinc = np.where(small, 0.5 - half * half / 48.0, np.sin(half) / ang)
    ch = np.where(small, 1.0 - half * half / 2.0, np.cos(half))
    v = sinc[:, None] * th
    return quat_to_rotmat(ch, v[:, 0], v[:, 1], v[:, 2])


def rotmat_to_rotvec(R: np.ndarray) -> np.ndarray:
    """Log map: Shepperd selection by argmax of [tr, R00, R11, R22] (first index
    on ties), w >= 0, normalise (floor 1e-15), ``2 atan2(|v|, w)/|v| v`` with the
    ``2 v`` branch below |v| < 1e-12 (``_batch_rotmat_to_rotvec``,
    _beam_cr.py:729-787)."""
    n = R.shape[0]
    tr = R[:, 0, 0] + R[:, 1, 1] + R[:, 2, 2]
    best = np.argmax(np.stack([tr, R[:, 0, 0], R[:, 1, 1], R[:, 2, 2]], axis=1), axis=1)
    q = np.empty((n, 4))
    m = best == 0
    if m.any():
        s = 2.0 * np.sqrt(tr[m] + 1.0)
        q[m, 0] = 0.25 * s
        q[m, 1] = (R[m, 2, 1] - R[m, 1, 2]) / s
        q[m, 2] = (R[m, 0, 2] - R[m, 2, 0]) / s
        q[m, 3] = (R[m, 1, 0] - R[m, 0, 1]) / s
    m = best == 1
    if m.any():
        s = 2.0 * np.sqrt(1.0 + R[m, 0, 0] - R[m, 1, 1] - R[m, 2, 2])
        q[m, 0] = (R[m, 2, 1] - R[m, 1, 2]) / s
        q[m, 1] = 0.25 * s
        q[m, 2] = (R[m, 0, 1] + R[m, 1, 0]) / s
        q[m, 3] = (R[m, 0, 2] + R[m, 2, 0]) / s
    m = best == 2
    if m.any():
        s = 2.0 * np.sqrt(1.0 + R[m, 1, 1] - R[m, 0, 0] - R[m, 2, 2])
        q[m, 0] = (R[m, 0, 2] - R[m, 2, 0]) / s
        q[m, 1] = (R[m, 0, 1] + R[m, 1, 0]) / s
        q[m, 2] = 0.25 * s
        q[m, 3] = (R[m, 1, 2] + R[m, 2, 1]) / s
    m = best == 3
    if m.any():
        s = 2.0 * np.sqrt(1.0 + R[m, 2, 2] - R[m, 0, 0] - R[m, 1, 1])
        q[m, 0] = (R[m, 1, 0] - R[m, 0, 1]) / s
        q[m, 1] = (R[m, 0, 2] + R[m, 2, 0]) / s
        q[m, 2] = (R[m, 1, 2] + R[m, 2, 1]) / s
        q[m, 3] = 0.25 * s
    neg = q[:, 0] < 0
    q[neg] = -q[neg]
    q = q / np.maximum(np.linalg.norm(q, axis=1, keepdims=True), 1e-15)
    v = q[:, 1:4]
    sh = np.linalg.norm(v, axis=1)
    small = sh < 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        coef = np.where(small, 2.0, 2.0 * np.arctan2(sh, q[:, 0]) / sh)
    return coef[:, None] * v


def rodrigues_min_rotation(e_from: np.ndarray, e_to: np.ndarray) -> np.ndarray:
    """Minimal rotation taking e_from to e_to (``_batch_rodrigues_rotation``,
    _beam_cr.py:790-828): I + [v]x + (1-c)/s^2 [v]x^2, identity when s < 1e-14
    and c > 0, a half-turn about a perpendicular axis when s < 1e-14 and c <= 0."""
    v = np.cross(e_from, e_to)
    s = np.linalg.norm(v, axis=1)
    c = np.sum(e_from * e_to, axis=1)
    small = s < 1e-14
    V = skew(v)
    V2 = V @ V
    den = np.where(small, 1.0, s * s)
    coef = np.where(small, 0.0, (1.0 - c) / den)
    out = np.eye(3)[None] + V + coef[:, None, None] * V2
    out[small & (c > 0)] = np.eye(3)
    for k in np.nonzero(small & (c <= 0))[0]:
        a = np.abs(e_from[k])
        if a[0] <= a[1] and a[0] <= a[2]:
            perp = np.array([1.0, 0.0, 0.0])
        elif a[1] <= a[2]:
            perp = np.array([0.0, 1.0, 0.0])
        else:
            perp = np.array([0.0, 0.0, 1.0])
        ax = np.cross(e_from[k], perp)
        ax = ax / np.linalg.norm(ax)
        out[k] = 2.0 * np.outer(ax, ax) - np.eye(3)
    return out


def local_axes(e_x: np.ndarray, v_ref: np.ndarray | None = None) -> np.ndarray:
    """Rows (e_x, e_y, e_z) with v_ref = unit axis of min |e_x| (argmin, first on
    ties) unless given; |e_x x v_ref| floored at 1e-10 (``_batch_build_local_axes``,
    _beam_cr.py:831-853)."""
    n = e_x.shape[0]
    if v_ref is not None:
        vr = np.tile(np.asarray(v_ref, dtype=float), (n, 1))
    else:
        vr = np.zeros((n, 3))
        vr[np.arange(n), np.argmin(np.abs(e_x), axis=1)] = 1.0
    e_z = np.cross(e_x, vr)
    e_z = e_z / np.maximum(np.linalg.norm(e_z, axis=1, keepdims=True), 1e-10)
    e_y = np.cross(e_z, e_x)
    return np.stack([e_x, e_y, e_z], axis=1)


def tangent_op(th: np.ndarray) -> np.ndarray:
    """T_s(th) = I + (1-cos a)/a^2 S + (1 - sin a / a)/a^2 S^2, Taylor
    I + S/2 + S^2/6 below a < 1e-10 (_beam_cr.py:856-879)."""
    a = np.linalg.norm(th, axis=1)
    S = skew(th)
    S2 = S @ S
    small = a < 1e-10
    with np.errstate(divide="ignore", invalid="ignore"):
        c1 = np.where(small, 0.5, (1.0 - np.cos(a)) / (a * a))
        c2 = np.where(small, 1.0 / 6.0, (1.0 - np.sin(a) / a) / (a * a))
    return np.eye(3)[None] + c1[:, None, None] * S + c2[:, None, None] * S2


def tangent_op_inv(th: np.ndarray) -> np.ndarray:
    """T_s^-1(th) = I - S/2 + (1/a^2 - (1+cos a)/(2 a sin a)) S^2, Taylor
    coefficient 1/12 below a < 1e-10 (_beam_cr.py:882-908)."""
    a = np.linalg.norm(th, axis=1)
    S = skew(th)
    S2 = S @ S
    small = a < 1e-10
    with np.errstate(divide="ignore", invalid="ignore"):
        c = np.where(small, 1.0 / 12.0, 1.0 / (a * a) - (1.0 + np.cos(a)) / (2.0 * a * np.sin(a)))
    return np.eye(3)[None] - 0.5 * S + c[:, None, None] * S2


def ke_local(sec: Section, L: np.ndarray, scf: float | None = None) -> np.ndarray:
    """12x12 local Timoshenko stiffness with Phi = 12 E I/(kappa G A L^2)
    (``_batch_timo_ke_local``, _beam_cr.py:911-998; scalar form :100-183).
    Local DOF order per node: [u, v, w, thx, thy, thz]."""
    E, G, A = sec.E, sec.G, sec.A
    n = L.shape[0]
    phy = 12.0 * E * sec.Iy / (sec.kappa_y * G * A * L * L)
    phz = 12.0 * E * sec.Iz / (sec.kappa_z * G * A * L * L)
    if scf is not None and scf > 0.0:
        phy = phy * (1.0 / (1.0 + scf * (L * L * A / (12.0 * sec.Iy))))
        phz = phz * (1.0 / (1.0 + scf * (L * L * A / (12.0 * sec.Iz))))
    dy, dz = 1.0 + phy, 1.0 + phz
    K = np.zeros((n, 12, 12))
    ea = E * A / L
    gj = G * sec.J / L
    for (i, j), val in {(0, 0): ea, (0, 6): -ea, (6, 0): -ea, (6, 6): ea,
                        (3, 3): gj, (3, 9): -gj, (9, 3): -gj, (9, 9): gj}.items():
        K[:, i, j] = val
    L2, L3 = L * L, L * L * L
    # bending in the local x-y plane (v, thz): DOFs 1, 5, 7, 11
    a3, a2, a1 = E * sec.Iz / L3, E * sec.Iz / L2, E * sec.Iz / L
    b = {
        (1, 1): 12 * a3, (1, 5): 6 * a2, (1, 7): -12 * a3, (1, 11): 6 * a2,
        (5, 5): (4 + phz) * a1, (5, 7): -6 * a2, (5, 11): (2 - phz) * a1,
        (7, 7): 12 * a3, (7, 11): -6 * a2, (11, 11): (4 + phz) * a1,
    }
    for (i, j), val in b.items():
        K[:, i, j] = val / dz
        K[:, j, i] = val / dz
    # bending in the local x-z plane (w, thy): DOFs 2, 4, 8, 10
    a3, a2, a1 = E * sec.Iy / L3, E * sec.Iy / L2, E * sec.Iy / L
    b = {
        (2, 2): 12 * a3, (2, 4): -6 * a2, (2, 8): -12 * a3, (2, 10): -6 * a2,
        (4, 4): (4 + phy) * a1, (4, 8): 6 * a2, (4, 10): (2 - phy) * a1,
        (8, 8): 12 * a3, (8, 10): 6 * a2, (10, 10): (4 + phy) * a1,
    }
    for (i, j), val in b.items():
        K[:, i, j] = val / dy
        K[:, j, i] = val / dy
    return K


# --------------------------------------------------------------------------
# element batch
# --------------------------------------------------------------------------


def element_dofs(conn: np.ndarray) -> np.ndarray:
    """(N_e, 12) global DOF indices, node-major (_beam_assembly.py:72-74)."""
    c = np.asarray(conn, dtype=np.int64)
    return (c[:, :, None] * 6 + np.arange(6, dtype=np.int64)[None, None, :]).reshape(len(c), 12)


def cr_beam_element_batch(
    X: np.ndarray,
    conn: np.ndarray,
    u: np.ndarray,
    sec: Section,
    *,
    v_ref: np.ndarray | None = None,
    scf: float | None = None,
    stiffness: bool = True,
    internal_force: bool = True,
):
    """Per-element CR internal force (N_e,12) and symmetrised analytical
    tangent (N_e,12,12) (``_assemble_cr_beam3d_batch_impl``,
    _beam_assembly.py:211-325).

    Returns ``(f_e, K_e)``; either is None when not requested.
    """
    c = np.asarray(conn, dtype=np.int64)
    edofs = element_dofs(c)
    ue = np.asarray(u, dtype=float)[edofs]
    P = np.asarray(X, dtype=float)[c]  # (N,2,3)

    d0 = P[:, 1] - P[:, 0]
    L0 = np.linalg.norm(d0, axis=1)
    ex0 = d0 / L0[:, None]
    R0 = local_axes(ex0, v_ref)

    x1 = P[:, 0] + ue[:, 0:3]
    x2 = P[:, 1] + ue[:, 6:9]
    dd = x2 - x1
    Ld = np.linalg.norm(dd, axis=1)
    exd = dd / Ld[:, None]

    Rrod = rodrigues_min_rotation(ex0, exd)
    Rcr = R0 @ np.transpose(Rrod, (0, 2, 1))  # R_0 R_rod^T
    Rn1 = rotvec_to_rotmat(ue[:, 3:6])
    Rn2 = rotvec_to_rotmat(ue[:, 9:12])
    R0T = np.transpose(R0, (0, 2, 1))
    th1 = rotmat_to_rotvec(Rcr @ Rn1 @ R0T)
    th2 = rotmat_to_rotvec(Rcr @ Rn2 @ R0T)

    n = len(c)
    dcr = np.zeros((n, 12))
    dcr[:, 3:6] = th1
    dcr[:, 6] = Ld - L0
    dcr[:, 9:12] = th2
    Kl = ke_local(sec, L0, scf)
    fcr = np.einsum("nij,nj->ni", Kl, dcr)
    RcrT = np.transpose(Rcr, (0, 2, 1))

    f_e = None
    if internal_force:
        f_e = np.empty((n, 12))
        for b in range(4):
            f_e[:, 3 * b : 3 * b + 3] = np.einsum("nij,nj->ni", RcrT, fcr[:, 3 * b : 3 * b + 3])

    K_e = None
    if stiffness:
        dpsi1 = (1.0 / Ld)[:, None, None] * (Rcr @ skew(exd))
        dpsi2 = -dpsi1
        Ti1 = tangent_op_inv(th1)
        Ti2 = tangent_op_inv(th2)
        Ts1 = tangent_op(ue[:, 3:6])
        Ts2 = tangent_op(ue[:, 9:12])
        B = np.zeros((n, 12, 12))
        B[:, 6, 0:3] = -exd
        B[:, 6, 6:9] = exd
        B[:, 3:6, 0:3] = Ti1 @ dpsi1
        B[:, 3:6, 6:9] = Ti1 @ dpsi2
        B[:, 3:6, 3:6] = Ti1 @ Rcr @ Ts1
        B[:, 9:12, 0:3] = Ti2 @ dpsi1
        B[:, 9:12, 6:9] = Ti2 @ dpsi2
        B[:, 9:12, 9:12] = Ti2 @ Rcr @ Ts2
        KB = Kl @ B
        Km = np.empty((n, 12, 12))
        for b in range(4):
            Km[:, 3 * b : 3 * b + 3, :] = RcrT @ KB[:, 3 * b : 3 * b + 3, :]
        Kg = np.zeros((n, 12, 12))
        for b in range(4):
            RtSf = RcrT @ skew(fcr[:, 3 * b : 3 * b + 3])
            Kg[:, 3 * b : 3 * b + 3, 0:3] += RtSf @ dpsi1
            Kg[:, 3 * b : 3 * b + 3, 6:9] += RtSf @ dpsi2
        KT = Km + Kg
        K_e = 0.5 * (KT + np.transpose(KT, (0, 2, 1)))
    return f_e, K_e


def assemble_cr_beam3d(
    X: np.ndarray,
    conn: np.ndarray,
    u: np.ndarray,
    sec: Section,
    *,
    v_ref: np.ndarray | None = None,
    scf: float | None = None,
    stiffness: bool = True,
    internal_force: bool = True,
):
    """Global (K_T csr, f_int) as ``assemble_cr_beam3d`` with analytical tangent
    and sparse output (_beam_assembly.py:121-152): f_int by ``np.add.at`` over
    element DOFs, K by COO (row-repeat / col-tile of the element DOFs) ->
    CSR with duplicate summation."""
    X = np.asarray(X, dtype=float)
    ndof = 6 * X.shape[0]
    c = np.asarray(conn, dtype=np.int64)
    if len(c) == 0:
        return (sp.csr_matrix((ndof, ndof)) if stiffness else None,
                np.zeros(ndof) if internal_force else None)
    f_e, K_e = cr_beam_element_batch(X, c, u, sec, v_ref=v_ref, scf=scf,
                                     stiffness=stiffness, internal_force=internal_force)
    edofs = element_dofs(c)
    f_int = None
    if internal_force:
        f_int = np.zeros(ndof)
        np.add.at(f_int, edofs.ravel(), f_e.ravel())
    K = None
    if stiffness:
        rows = np.repeat(edofs, 12, axis=1).ravel()
        cols = np.tile(edofs, (1, 12)).ravel()
        K = sp.csr_matrix((K_e.reshape(-1), (rows, cols)), shape=(ndof, ndof))
        K.sum_duplicates()
    return K, f_int


def lumped_mass_diag(X: np.ndarray, conn: np.ndarray, rho: float, sec: Section) -> np.ndarray:
    """Diagonal of the lumped mass (``ULCRBeamAssembler.assemble_mass(lumped)``,
    _beam_assembler.py:148-168 with ``timo_beam3d_lumped_mass_local``,
    _beam_cr.py:267-296): per element m/2 translations, rho(Iy+Iz)L/2 torsion,
    m L^2/78 for the two bending rotations, at each node."""
    X = np.asarray(X, dtype=float)
    c = np.asarray(conn, dtype=np.int64)
    L = np.linalg.norm(X[c[:, 1]] - X[c[:, 0]], axis=1)
    m = rho * sec.A * L
    mt = rho * (sec.Iy + sec.Iz) * L
    ri = m * L * L / 78.0
    d6 = np.stack([m / 2, m / 2, m / 2, mt / 2, ri, ri], axis=1)
    diag = np.zeros(6 * X.shape[0])
    for k in range(2):
        for j in range(6):
            np.add.at(diag, c[:, k] * 6 + j, d6[:, j])
    return diag


def ke_global_linear(X2: np.ndarray, sec: Section) -> np.ndarray:
    """T^T Ke_local T for one element at the reference configuration
    (``timo_beam3d_ke_global``, _beam_cr.py:186-204).  Used as the small-u limit."""
    d = X2[1] - X2[0]
    L = float(np.linalg.norm(d))
    R = local_axes((d / L)[None])[0]
    T = np.zeros((12, 12))
    for i in range(4):
        T[3 * i : 3 * i + 3, 3 * i : 3 * i + 3] = R
    return T.T @ ke_local(sec, np.array([L]))[0] @ T


def consistent_mass(X, conn, rho: float, sec, v_ref=None) -> sp.csr_matrix:
    """``assemble_mass(rho, lumped=False)`` (elements/_beam_assembler.py:148-196)
    with ``timo_beam3d_mass_local`` / ``timo_beam3d_mass_global``
    (_beam_cr.py:212-321): per element T^T Me T with the local axes of
    ``_build_local_axes`` (:59-84), entries |val| <= 1e-30 dropped, COO -> CSR."""
    X = np.asarray(X, float)
    conn = np.asarray(conn, np.int64)
    rows, cols, vals = [], [], []
    for n1, n2 in conn:
        d = X[n2] - X[n1]
        L = float(np.linalg.norm(d))
        ex = d / L
        if v_ref is None:
            a = np.abs(ex)
            vr = np.eye(3)[0 if (a[0] <= a[1] and a[0] <= a[2]) else (1 if a[1] <= a[2] else 2)]
        else:
            vr = np.asarray(v_ref, float)
        ez = np.cross(ex, vr)
        ez = ez / np.linalg.norm(ez)
        R = np.stack([ex, np.cross(ez, ex), ez])
        m = rho * sec.A * L
        Me = np.zeros((12, 12))
        Me[0, 0] = Me[6, 6] = m / 3.0
        Me[0, 6] = Me[6, 0] = m / 6.0
        mt = rho * (sec.Iy + sec.Iz) * L
        Me[3, 3] = Me[9, 9] = mt / 3.0
        Me[3, 9] = Me[9, 3] = mt / 6.0
        B = (m / 420.0) * np.array([[156.0, 22.0 * L, 54.0, -13.0 * L], [22.0 * L, 4.0 * L**2, 13.0 * L, -3.0 * L**2],
                                    [54.0, 13.0 * L, 156.0, -22.0 * L], [-13.0 * L, -3.0 * L**2, -22.0 * L, 4.0 * L**2]])
        for idx, sg in (([1, 5, 7, 11], np.ones(4)), ([2, 4, 8, 10], np.array([1.0, -1.0, 1.0, -1.0]))):
            Me[np.ix_(idx, idx)] = B * np.outer(sg, sg)
        T = np.kron(np.eye(4), R)
        Mg = T.T @ Me @ T
        ed = np.r_[6 * n1 + np.arange(6), 6 * n2 + np.arange(6)]
        ii, jj = np.nonzero(np.abs(Mg) > 1e-30)
        rows.append(ed[ii])
        cols.append(ed[jj])
        vals.append(Mg[ii, jj])
    n = 6 * len(X)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
