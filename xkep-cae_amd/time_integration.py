"""Generalized-alpha (Chung & Hulbert 1993) with a lumped mass, state on the GPU.

Mirrors ``GeneralizedAlphaProcess`` (xkep_cae/time_integration/strategy.py:96-232):
alpha_m = (2 rho_inf - 1)/(rho_inf + 1), alpha_f = rho_inf/(rho_inf + 1),
gamma = 1/2 - alpha_m + alpha_f, beta = (1 - alpha_m + alpha_f)^2 / 4;
predict / correct / effective_residual / effective_stiffness / checkpoint.
The fused assembly (``assembly.StrandAssembly``) performs ``correct`` and the
effective residual/stiffness inside its kernels on these device buffers.
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import _lib as L
from .context import require_gpu


class GeneralizedAlphaProcess:
    """``mass_matrix``: a lumped mass as its diagonal (1-D) or any mass matrix;
    ``damping_matrix``: optional C.  The fused assembly and the device solve
    take a diagonal mass without damping (``dyn_struct`` refuses anything else);
    the Strategy methods ``effective_residual`` / ``effective_stiffness`` take
    any M and C, as the reference's (time_integration/strategy.py:178-216)."""

    def __init__(self, mass_matrix, *, damping_matrix=None, rho_inf: float = 0.9):
        torch = require_gpu()
        self.torch = torch
        if sp.issparse(mass_matrix) or np.ndim(mass_matrix) == 2:
            M = sp.csr_matrix(mass_matrix)
            Md = M.diagonal()
            self.M = M
            self.lumped = M.count_nonzero() == np.count_nonzero(Md)
        else:
            Md = np.asarray(mass_matrix, float)
            self.M = sp.diags(Md, format="csr")
            self.lumped = True
        self.Md = np.ascontiguousarray(Md, dtype=float)
        self.C = None if damping_matrix is None else sp.csr_matrix(damping_matrix)
        self.rho_inf = rho_inf
        self.alpha_m = (2.0 * rho_inf - 1.0) / (rho_inf + 1.0)
        self.alpha_f = rho_inf / (rho_inf + 1.0)
        self.gamma = 0.5 - self.alpha_m + self.alpha_f
        self.beta = 0.25 * (1.0 - self.alpha_m + self.alpha_f) ** 2
        n = len(self.Md)
        dev = torch.device("cuda")
        z = lambda: torch.zeros(n, dtype=torch.float64, device=dev)
        self.d_M = torch.as_tensor(self.Md, device=dev)
        self.d_vel, self.d_acc, self.d_vel_old, self.d_acc_old, self.d_u_pred, self.d_v_pred = (z() for _ in range(6))
        self._ckpt = None

    @property
    def is_dynamic(self) -> bool:
        return True

    # host views (drop-in attribute access)
    @property
    def vel(self):
        return self.d_vel.cpu().numpy()

    @property
    def acc(self):
        return self.d_acc.cpu().numpy()

    def set_initial_state(self, velocity=None, acceleration=None):
        if velocity is not None:
            self.d_vel.copy_(self.torch.as_tensor(velocity))
        if acceleration is not None:
            self.d_acc.copy_(self.torch.as_tensor(acceleration))

    def predict_device(self, u, dt: float):
        self.d_acc_old.copy_(self.d_acc)
        self.d_vel_old.copy_(self.d_vel)
        self.d_u_pred.copy_(u + dt * self.d_vel + 0.5 * dt**2 * (1.0 - 2.0 * self.beta) * self.d_acc)
        self.d_v_pred.copy_(self.d_vel + dt * (1.0 - self.gamma) * self.d_acc)
        return self.d_u_pred.clone()

    def predict(self, u, dt: float):
        ud = self.torch.as_tensor(np.asarray(u, float), device=self.d_M.device)
        return self.predict_device(ud, dt).cpu().numpy()

    def correct(self, u, du, dt: float) -> None:
        if dt < 1e-30:
            return
        c0 = 1.0 / (self.beta * dt**2)
        ud = self.torch.as_tensor(np.asarray(u, float), device=self.d_M.device) if not isinstance(
            u, self.torch.Tensor) else u
        self.d_acc.copy_(c0 * (ud - self.d_u_pred))
        self.d_vel.copy_(self.d_v_pred + dt * self.gamma * self.d_acc)

    def effective_stiffness(self, K, dt: float):
        """K + (1 - alpha_m) c0 M + (1 - alpha_f) c1 C, c0 = 1 / (beta dt^2), c1 = gamma / (beta dt)."""
        if dt < 1e-30:
            return K
        K_eff = K + (1.0 - self.alpha_m) * (1.0 / (self.beta * dt**2)) * self.M
        if self.C is not None:
            K_eff = K_eff + (1.0 - self.alpha_f) * (self.gamma / (self.beta * dt)) * self.C
        return K_eff

    def effective_residual(self, R, dt: float):
        """R + M a_mid + C v_mid with a_mid = (1 - alpha_m) a + alpha_m a_old and
        v_mid = (1 - alpha_f) v + alpha_f v_old."""
        a_mid = ((1.0 - self.alpha_m) * self.d_acc + self.alpha_m * self.d_acc_old).cpu().numpy()
        R_eff = np.asarray(R, float) + (self.Md * a_mid if self.lumped else self.M @ a_mid)
        if self.C is not None:
            v_mid = ((1.0 - self.alpha_f) * self.d_vel + self.alpha_f * self.d_vel_old).cpu().numpy()
            R_eff = R_eff + self.C @ v_mid
        return R_eff

    def checkpoint(self) -> None:
        self._ckpt = (self.d_vel.clone(), self.d_acc.clone())

    def restore_checkpoint(self) -> None:
        if self._ckpt is not None:
            self.d_vel.copy_(self._ckpt[0])
            self.d_acc.copy_(self._ckpt[1])

    def dyn_struct(self, dt: float, *, state: bool = True) -> L.XkDyn:
        """Pointers of the device state for the fused assembly kernels (a lumped
        mass without damping: the kernels apply M as a diagonal).  Without
        ``state`` the assembly leaves acc and vel alone (``correct`` at the
        increment's end sets them)."""
        if not self.lumped or self.C is not None:
            raise ValueError("the fused assembly takes a lumped (diagonal) mass and no damping matrix")
        return L.XkDyn(M_diag=self.d_M.data_ptr(), u_pred=self.d_u_pred.data_ptr(), v_pred=self.d_v_pred.data_ptr(),
                       acc_old=self.d_acc_old.data_ptr(), acc=self.d_acc.data_ptr() if state else None,
                       vel=self.d_vel.data_ptr() if state else None,
                       dt=dt, alpha_m=self.alpha_m, beta=self.beta, gamma=self.gamma)
