"""Oracle: per-Newton-iteration composition of R and K_T, FP64.

Test infrastructure only (see ``oracle/__init__.py``).

Restates
  - ``ContactForceAssemblyProcess``  contact/solver/_newton_steps.py:79-135
  - ``TangentAssemblyProcess``       contact/solver/_newton_steps.py:289-313
  - the Kelvin-Voigt coating terms of both (``coating.py``)
  - ``GeneralizedAlphaProcess``      time_integration/strategy.py:96-232
  - the dynamic residual/stiffness steps of one Newton attempt
    (contact/solver/_newton_dynamic.py:164-192, :294-315) and the active count
    of ``ConvergenceCheckProcess`` (_newton_steps.py:194-213)
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import beam, coating, contact


class GeneralizedAlpha:
    """Chung-Hulbert Generalized-alpha (time_integration/strategy.py:96-232):
    a lumped mass given as its diagonal, or any mass matrix, and an optional
    damping matrix C (K_eff += (1 - alpha_f) c1 C, R_eff += C v_mid with
    c1 = gamma / (beta dt), v_mid = (1 - alpha_f) v + alpha_f v_old)."""

    def __init__(self, M_diag, rho_inf: float = 0.9, *, damping_matrix=None):
        if sp.issparse(M_diag) or np.ndim(M_diag) == 2:
            self.M = sp.csr_matrix(M_diag)
            self.Md = self.M.diagonal()
        else:
            self.Md = np.asarray(M_diag, float)
            self.M = sp.diags(self.Md, format="csr")
        self.C = None if damping_matrix is None else sp.csr_matrix(damping_matrix)
        self.rho_inf = rho_inf
        self.alpha_m = (2.0 * rho_inf - 1.0) / (rho_inf + 1.0)
        self.alpha_f = rho_inf / (rho_inf + 1.0)
        self.gamma = 0.5 - self.alpha_m + self.alpha_f
        self.beta = 0.25 * (1.0 - self.alpha_m + self.alpha_f) ** 2
        n = len(self.Md)
        self.vel = np.zeros(n)
        self.acc = np.zeros(n)
        self._vel_old = np.zeros(n)
        self._acc_old = np.zeros(n)
        self._u_pred = np.zeros(n)
        self._v_pred = np.zeros(n)

    def predict(self, u, dt):
        self._acc_old = self.acc.copy()
        self._vel_old = self.vel.copy()
        self._u_pred = u + dt * self.vel + 0.5 * dt**2 * (1.0 - 2.0 * self.beta) * self.acc
        self._v_pred = self.vel + dt * (1.0 - self.gamma) * self.acc
        return self._u_pred.copy()

    def correct(self, u, du, dt):
        if dt < 1e-30:
            return
        c0 = 1.0 / (self.beta * dt**2)
        self.acc = c0 * (u - self._u_pred)
        self.vel = self._v_pred + dt * self.gamma * self.acc

    def c0(self, dt):
        return 1.0 / (self.beta * dt**2)

    def effective_stiffness(self, K, dt):
        if dt < 1e-30:
            return K
        K_eff = K + (1.0 - self.alpha_m) * self.c0(dt) * self.M
        if self.C is not None:
            K_eff = K_eff + (1.0 - self.alpha_f) * (self.gamma / (self.beta * dt)) * self.C
        return K_eff

    def effective_residual(self, R, dt):
        a_mid = (1.0 - self.alpha_m) * self.acc + self.alpha_m * self._acc_old
        R_eff = R.copy() + self.M @ a_mid
        if self.C is not None:
            R_eff += self.C @ ((1.0 - self.alpha_f) * self.vel + self.alpha_f * self._vel_old)
        return R_eff

    def checkpoint(self):
        self._ckpt = (self.vel.copy(), self.acc.copy())

    def restore_checkpoint(self):
        if getattr(self, "_ckpt", None) is not None:
            self.vel, self.acc = self._ckpt[0].copy(), self._ckpt[1].copy()


@dataclass
class Problem:
    """Everything one assembly needs besides the state vectors."""

    X: np.ndarray
    conn: np.ndarray
    sec: beam.Section
    k_pen: float
    smoothing_delta: float
    k_t_ratio: float
    mu: float
    cfg: contact.ContactConfig

    @property
    def ndof(self) -> int:
        return 6 * self.X.shape[0]


def deformed_coords(X, u):
    """x = X + u[0:3] per node (``_deformed_coords``, contact/solver/_utils.py:16-28)."""
    return X + u.reshape(-1, 6)[:, :3]


@dataclass
class IterationResult:
    R: np.ndarray          # effective residual (fixed DOFs zeroed)
    K: sp.csr_matrix | None  # effective tangent
    res_norm: float
    n_active: int
    f_c: np.ndarray        # -(f_c + f_fric) as returned to the Newton loop
    f_int: np.ndarray
    pairs: contact.Pairs
    friction: contact.FrictionResult
    K_beam: sp.csr_matrix | None = None
    K_c: sp.csr_matrix | None = None
    K_fric: sp.csr_matrix | None = None


def assemble_iteration(prob: Problem, pairs: contact.Pairs, u, u_ref, f_ext, fixed_dofs,
                       ga: GeneralizedAlpha | None, dt: float, *, tangent: bool = True,
                       mu_eff: float | None = None, coat_dt: float = 1e-15) -> IterationResult:
    """One Newton attempt's assembly: narrowphase with frozen active set, Huber
    force (writes p_n), Coulomb friction (writes z_t/status), residual
    R = f_int - (f_c + f_fric) - f_ext, R[fixed] = 0, Generalized-alpha
    correct + effective residual, R[fixed] = 0, ||R||_2, n_active = #(p_n > 0),
    and (optionally) K_T = K_beam + K_c - K_fric + (1 - alpha_m) c0 M.

    With the coating on (cfg.coating_stiffness > 0) the Kelvin-Voigt terms
    enter after friction, f_c = ((f_huber + f_fric) + f_coat) + f_coat_fric
    (_newton_steps.py:110-121) and K_T = ((K_beam + K_c) + K_coat) + K_coat_fric
    - K_fric (:296-311); coat_dt = max(load_frac - load_frac_prev, 1e-15) is
    the dashpot's time step (:112, :297)."""
    ndof = prob.ndof
    mu_eff = prob.mu if mu_eff is None else mu_eff
    xd = deformed_coords(prob.X, u)
    p = contact.update_geometry(pairs, xd, prob.cfg, freeze_active_set=True)
    f_c, _, p = contact.huber_evaluate(p, prob.k_pen, prob.smoothing_delta, ndof)
    fr = contact.friction_evaluate(p, u, u_ref, prob.k_pen, prob.k_t_ratio, mu_eff, ndof)
    p = fr.pairs
    f_c = f_c + fr.f_fric
    coat = prob.cfg.coating_stiffness > 0.0
    if coat:
        f_c = f_c + coating.forces(p, prob.cfg, coat_dt, ndof)
        if prob.cfg.coating_mu > 0.0:
            f_cf, p = coating.friction_forces(p, prob.cfg, u, u_ref, ndof)
            f_c = f_c + f_cf
    fc_neg = -f_c
    _, f_int = beam.assemble_cr_beam3d(prob.X, prob.conn, u, prob.sec, stiffness=False)
    R = f_int + fc_neg - f_ext
    R[fixed_dofs] = 0.0
    if ga is not None and dt > 1e-30:
        ga.correct(u, np.zeros_like(u), dt)
        R = ga.effective_residual(R, dt)
        R[fixed_dofs] = 0.0
    res_norm = float(np.linalg.norm(R))
    n_active = int(np.sum(p.p_n > 0.0))
    out = IterationResult(R=R, K=None, res_norm=res_norm, n_active=n_active, f_c=fc_neg,
                          f_int=f_int, pairs=p, friction=fr)
    if tangent:
        K_beam, _ = beam.assemble_cr_beam3d(prob.X, prob.conn, u, prob.sec, internal_force=False)
        K_c = contact.huber_tangent(p, prob.k_pen, prob.smoothing_delta, ndof)
        K = K_beam + K_c
        if coat:
            m, b = coating.stiffness_blocks(p, prob.cfg, coat_dt)
            K = K + contact.blocks_to_csr(p, m, b, ndof)
            if prob.cfg.coating_mu > 0.0:
                m, b = coating.friction_stiffness_blocks(p, prob.cfg)
                K = K + contact.blocks_to_csr(p, m, b, ndof)
        K_fric = None
        if fr.mask.any():
            K_fric = contact.friction_tangent(p, fr, ndof)
            K = K - K_fric
        if ga is not None and dt > 1e-30:
            K = ga.effective_stiffness(K, dt)
        out.K, out.K_beam, out.K_c, out.K_fric = K.tocsr(), K_beam, K_c, K_fric
    return out
