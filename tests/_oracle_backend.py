"""Solver backend made of the oracle (FP64 NumPy restatement): the product's
Newton driver run on it reproduces the reference solver's trajectory on the
CPU.  Test infrastructure (imports the oracle)."""

from __future__ import annotations

import numpy as np

from oracle import assembly as oasm
from oracle import beam as obeam
from oracle import coating as ocoat
from oracle import contact as ocon


class OracleBackend:
    def __init__(self, X, conn, radii, sec, *, M_diag, rho_inf, k_pen, smoothing_delta, mu, k_t_ratio=0.5,
                 cfg: ocon.ContactConfig):
        self.X, self.conn, self.radii = X, conn, radii
        self.cfg = cfg
        self.prob = oasm.Problem(X, conn, sec, k_pen, smoothing_delta, k_t_ratio, mu, cfg)
        self.ga = oasm.GeneralizedAlpha(M_diag, rho_inf)
        self.pairs = ocon.Pairs()
        self.fixed = None
        self.coating = cfg.coating_stiffness > 0.0
        self.coat_dt = 1e-15

    def set_coat_dt(self, coat_dt):
        self.coat_dt = coat_dt

    def commit_coating(self):
        self.pairs = ocoat.commit(self.pairs)

    def _x(self, u):
        return self.X + u.reshape(-1, 6)[:, :3]

    def detect(self, u):
        self.pairs, _ = ocon.detect_candidates(self.pairs, self._x(u), self.conn, self.radii, self.cfg)

    def update_geometry(self, u):
        self.pairs = ocon.update_geometry(self.pairs, self._x(u), self.cfg)

    def initial_penetrations(self, u):
        if len(self.pairs) == 0:
            return 0
        p = ocon.update_geometry(self.pairs.copy(), self._x(u), self.cfg)
        return int(np.sum(p.gap < 0.0))

    def predict(self, u, dt):
        return self.ga.predict(u, dt)

    def assemble(self, u, u_ref, f_ext, dt):
        r = oasm.assemble_iteration(self.prob, self.pairs, u, u_ref, f_ext, self.fixed, self.ga, dt,
                                    coat_dt=self.coat_dt)
        self.pairs = r.pairs
        return r.R, r.K, r.n_active, r.f_c

    def internal_force(self, u):
        return obeam.assemble_cr_beam3d(self.X, self.conn, u, self.prob.sec, stiffness=False)[1]

    def correct(self, u, dt):
        self.ga.correct(u, np.zeros_like(u), dt)

    def checkpoint(self):
        self.ga.checkpoint()

    def restore_checkpoint(self):
        self.ga.restore_checkpoint()
