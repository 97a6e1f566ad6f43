"""Drop-in ``ULCRBeamAssembler`` on the GPU.

Mirrors xkep_cae/elements/_beam_assembler.py:19-196: same constructor,
``assemble_tangent(u) -> scipy CSR``, ``assemble_internal_force(u) -> ndarray``,
``update_reference``, ``checkpoint`` / ``rollback``, ``u_total_accum``,
``get_total_displacement`` and ``assemble_mass``.  The element force and
tangent come from the gfx950 element kernel (one element per lane, see
csrc/beam.hip); the ``*_device`` variants take and return torch tensors
without host copies.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .context import DeviceContext, Section


def _rotvec_to_rotmat(th: np.ndarray) -> np.ndarray:
    """Single rotation vector -> matrix via the unit quaternion (as _beam_cr.py:326-357),
    used only for the host-side reference-triad bookkeeping of ``update_reference``."""
    a = float(np.linalg.norm(th))
    if a < 1e-12:
        h2 = a * a / 4.0
        q = np.array([1.0 - h2 / 2.0, *((0.5 - h2 / 48.0) * th)])
        q = q / np.linalg.norm(q)
    else:
        q = np.array([np.cos(a / 2.0), *(np.sin(a / 2.0) / a * th)])
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


class ULCRBeamAssembler:
    """Corotational Timoshenko beam assembler backed by the gfx950 kernels."""

    def __init__(self, node_coords, connectivity, E, G, A, Iy, Iz, J, kappa_y, kappa_z=0.0, *,
                 v_ref=None, scf=None, precision: str = "fp64", radii=None, context: DeviceContext | None = None):
        self.coords_ref = np.array(node_coords, dtype=float, copy=True)
        self.connectivity = np.asarray(connectivity)
        self.n_nodes = len(self.coords_ref)
        self.ndof = 6 * self.n_nodes
        self.E, self.G, self.A, self.Iy, self.Iz, self.J = E, G, A, Iy, Iz, J
        self.kappa_y = kappa_y
        self.kappa_z = kappa_z if kappa_z > 0 else kappa_y
        self.v_ref = v_ref
        self.scf = scf
        self.precision = precision
        self.radii = radii
        self.R_ref = np.tile(np.eye(3), (self.n_nodes, 1, 1))
        self._u_total_accum = np.zeros(self.ndof)
        self._ckpt = None
        self._ctx = context

    @property
    def section(self) -> Section:
        return Section(self.E, self.G, self.A, self.Iy, self.Iz, self.J, self.kappa_y, self.kappa_z)

    @property
    def ctx(self) -> DeviceContext:
        if self._ctx is None:
            self._ctx = DeviceContext(self.coords_ref, self.connectivity, self.section, radii=self.radii,
                                      scf=self.scf, v_ref=self.v_ref, precision=self.precision)
        return self._ctx

    # ---------------------------------------------------------------- device API
    def assemble_internal_force_device(self, u):
        ctx = self.ctx
        ud = ctx.to_dev(u)
        f = ctx.torch.empty(self.ndof, dtype=ctx.torch.float64, device=ctx.dev)
        ctx.beam(ud, f_int=f)
        return f

    def assemble_tangent_device(self, u):
        """Tangent values in the context's CSR pattern (``ctx.pattern()``)."""
        ctx = self.ctx
        ud = ctx.to_dev(u)
        K = ctx.new_K()
        ctx.beam(ud, K_vals=K)
        return K

    # ---------------------------------------------------------------- drop-in API
    def _check_elements(self) -> None:
        """Raise as the reference's CR element routine does on a zero-length
        element (ValueError, elements/_beam_cr.py:54-55): the element kernel
        flags XK_FLAG_DEGENERATE in the device status word (read and cleared
        here, so this synchronises)."""
        if self.ctx.status_flags() & 2:
            raise ValueError("zero-length beam element (initial or deformed length below 1e-15)")

    def assemble_tangent(self, u_incr) -> sp.csr_matrix:
        """``ULCRBeamAssembler.assemble_tangent`` (_beam_assembler.py:73-93)."""
        K = self.ctx.to_csr(self.assemble_tangent_device(u_incr))
        self._check_elements()
        return K

    def assemble_internal_force(self, u_incr) -> np.ndarray:
        """``ULCRBeamAssembler.assemble_internal_force`` (_beam_assembler.py:95-115)."""
        f = self.assemble_internal_force_device(u_incr).cpu().numpy()
        self._check_elements()
        return f

    def update_reference(self, u_incr) -> None:
        """Updated-Lagrangian reference update (_beam_assembler.py:117-124); the
        device context is rebuilt on the new reference coordinates."""
        u_incr = np.asarray(u_incr, float)
        for i in range(self.n_nodes):
            self.coords_ref[i] += u_incr[6 * i:6 * i + 3]
            self.R_ref[i] = self.R_ref[i] @ _rotvec_to_rotmat(u_incr[6 * i + 3:6 * i + 6])
        self._u_total_accum += u_incr
        if self._ctx is not None:
            self._ctx.close()
        self._ctx = None

    def checkpoint(self) -> None:
        self._ckpt = (self.coords_ref.copy(), self.R_ref.copy(), self._u_total_accum.copy())

    def rollback(self) -> None:
        if self._ckpt is not None:
            changed = not np.array_equal(self._ckpt[0], self.coords_ref)
            self.coords_ref, self.R_ref, self._u_total_accum = (a.copy() for a in self._ckpt)
            if changed and self._ctx is not None:
                self._ctx.close()
                self._ctx = None

    @property
    def u_total_accum(self) -> np.ndarray:
        return self._u_total_accum

    def get_total_displacement(self, u_incr) -> np.ndarray:
        return self._u_total_accum + u_incr

    def lumped_mass_diag(self, rho: float) -> np.ndarray:
        """Diagonal of the lumped mass (``assemble_mass(lumped=True)``,
        _beam_assembler.py:148-168 with _beam_cr.py:267-296): per element m/2 on
        translations, rho (Iy+Iz) L/2 on torsion, m L^2/78 on both bending
        rotations, at each end node.  Set-up data, computed once."""
        c = np.asarray(self.connectivity, np.int64)
        L = np.linalg.norm(self.coords_ref[c[:, 1]] - self.coords_ref[c[:, 0]], axis=1)
        m = rho * self.A * L
        d6 = np.stack([m / 2, m / 2, m / 2, rho * (self.Iy + self.Iz) * L / 2, m * L * L / 78.0,
                       m * L * L / 78.0], axis=1)
        diag = np.zeros(self.ndof)
        for k in range(2):
            for j in range(6):
                np.add.at(diag, c[:, k] * 6 + j, d6[:, j])
        return diag

    def consistent_mass(self, rho: float) -> sp.csr_matrix:
        """Consistent mass (``assemble_mass(lumped=False)``, _beam_assembler.py:148-196,
        with timo_beam3d_mass_local / _global, _beam_cr.py:212-321): per element the
        local matrix -- axial and torsional m/3, m/6 couplings, the cubic bending
        blocks m/420 [156, 22L, 54, -13L; ...] in both bending planes (signs
        alternating in x-z) -- rotated by the element's local axes (reference axis
        v_ref, or the axis of least |e_x| component, first on ties), entries
        |val| <= 1e-30 dropped, summed into a CSR matrix.  Set-up data, built
        on the host once per mesh."""
        c = np.asarray(self.connectivity, np.int64)
        X = self.coords_ref
        d = X[c[:, 1]] - X[c[:, 0]]
        L = np.linalg.norm(d, axis=1)
        ex = d / L[:, None]
        if self.v_ref is None:
            a = np.abs(ex)
            k = np.where((a[:, 0] <= a[:, 1]) & (a[:, 0] <= a[:, 2]), 0, np.where(a[:, 1] <= a[:, 2], 1, 2))
            vr = np.eye(3)[k]
        else:
            vr = np.broadcast_to(np.asarray(self.v_ref, float), ex.shape)
        ez = np.cross(ex, vr)
        nz = np.linalg.norm(ez, axis=1)
        if np.any(nz < 1e-10):
            raise ValueError("reference vector parallel to a beam axis")
        ez = ez / nz[:, None]
        R = np.stack([ex, np.cross(ez, ex), ez], axis=1)            # rows e_x, e_y, e_z
        ne = len(c)
        m = rho * self.A * L
        mt = rho * (self.Iy + self.Iz) * L
        Me = np.zeros((ne, 12, 12))
        for i, j, w in ((0, 0, 1 / 3), (6, 6, 1 / 3), (0, 6, 1 / 6), (6, 0, 1 / 6)):
            Me[:, i, j] = w * m
        for i, j, w in ((3, 3, 1 / 3), (9, 9, 1 / 3), (3, 9, 1 / 6), (9, 3, 1 / 6)):
            Me[:, i, j] = w * mt
        q = m / 420.0
        Lc = [np.ones(ne), L, np.ones(ne), L]
        bend = np.array([[156.0, 22.0, 54.0, -13.0], [22.0, 4.0, 13.0, -3.0],
                         [54.0, 13.0, 156.0, -22.0], [-13.0, -3.0, -22.0, 4.0]])
        for idx, sg in (([1, 5, 7, 11], (1.0, 1.0, 1.0, 1.0)), ([2, 4, 8, 10], (1.0, -1.0, 1.0, -1.0))):
            for a_ in range(4):
                for b_ in range(4):
                    Me[:, idx[a_], idx[b_]] = (q * bend[a_, b_] * Lc[a_] * Lc[b_]) * (sg[a_] * sg[b_])
        M4 = Me.reshape(ne, 4, 3, 4, 3)
        Mg = np.einsum("eki,eakbl,elj->eaibj", R, M4, R).reshape(ne, 12, 12)
        ed = np.concatenate([6 * c[:, :1] + np.arange(6), 6 * c[:, 1:] + np.arange(6)], axis=1)
        keep = np.abs(Mg) > 1e-30
        rows = np.broadcast_to(ed[:, :, None], Mg.shape)[keep]
        cols = np.broadcast_to(ed[:, None, :], Mg.shape)[keep]
        return sp.csr_matrix((Mg[keep], (rows, cols)), shape=(self.ndof, self.ndof))

    def assemble_mass(self, rho: float, *, lumped: bool = True):
        """``assemble_mass`` (_beam_assembler.py:148-196): the lumped diagonal or the
        consistent matrix.  The fused assembly and the device solve take the lumped
        (diagonal) mass; the consistent one serves host-side consumers."""
        if not lumped:
            return self.consistent_mass(rho)
        return sp.diags(self.lumped_mass_diag(rho), format="csr")
