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
    def __init__(self, mass_matrix, *, damping_matrix=None, rho_inf: float = 0.9):
        if damping_matrix is not None:
            raise NotImplementedError("the strand path runs without a damping matrix")
        torch = require_gpu()
        self.torch = torch
        Md = mass_matrix.diagonal() if sp.issparse(mass_matrix) else np.asarray(mass_matrix, float)
        if Md.ndim == 2:
            Md = np.diag(Md)
        self.Md = np.ascontiguousarray(Md, dtype=float)
        self.M = sp.diags(self.Md, format="csr")
        self.C = None
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
        if dt < 1e-30:
            return K
        return K + (1.0 - self.alpha_m) * (1.0 / (self.beta * dt**2)) * self.M

    def effective_residual(self, R, dt: float):
        a_mid = (1.0 - self.alpha_m) * self.d_acc + self.alpha_m * self.d_acc_old
        return np.asarray(R, float) + self.Md * a_mid.cpu().numpy()

    def checkpoint(self) -> None:
        self._ckpt = (self.d_vel.clone(), self.d_acc.clone())

    def restore_checkpoint(self) -> None:
        if self._ckpt is not None:
            self.d_vel.copy_(self._ckpt[0])
            self.d_acc.copy_(self._ckpt[1])

    def dyn_struct(self, dt: float) -> L.XkDyn:
        """Pointers of the device state for the fused assembly kernels."""
        return L.XkDyn(M_diag=self.d_M.data_ptr(), u_pred=self.d_u_pred.data_ptr(), v_pred=self.d_v_pred.data_ptr(),
                       acc_old=self.d_acc_old.data_ptr(), acc=self.d_acc.data_ptr(), vel=self.d_vel.data_ptr(),
                       dt=dt, alpha_m=self.alpha_m, beta=self.beta, gamma=self.gamma)
