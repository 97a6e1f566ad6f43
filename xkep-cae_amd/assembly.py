"""Fused per-Newton-iteration assembly of one strand on the GPU.

One call of ``StrandAssembly.assemble`` performs what the reference does in
``ContactForceAssemblyProcess`` + Generalized-alpha correct/effective_residual
+ ``TangentAssemblyProcess`` + effective_stiffness
(contact/solver/_newton_steps.py:79-135, :289-313;
contact/solver/_newton_dynamic.py:164-192, :294-315): narrowphase with the
active set frozen, Huber force, Coulomb friction, element force and tangent,
R = f_int - (f_c + f_fric) - f_ext with fixed DOFs zeroed, inertia terms,
||R||^2 and #(p_n > 0) -- five kernel launches on one stream, no host
synchronisation, capturable in a HIP graph.

``begin_increment`` runs the once-per-increment part (contact/solver/process.py:406-429):
deformed coordinates, DetectCandidates on the GPU, UpdateGeometry with
hysteresis, and the tangent pattern of the increment.

With ``config.coating_stiffness > 0`` the Kelvin-Voigt coating terms
(contact/coating/strategy.py:80-384) join the pair kernel; each assembly then
first probes the coating compression at the current u, and a pair that
starts pressing its coating within the increment without a slot in the
tangent pattern makes the assembly rebuild the pattern before it runs.
"""

from __future__ import annotations

import numpy as np

from . import _lib as L
from .contact import ContactConfig, layer_array
from .context import DeviceContext
from .time_integration import GeneralizedAlphaProcess


class StrandAssembly:
    def __init__(self, ctx: DeviceContext, *, k_pen: float, smoothing_delta: float, k_t_ratio: float = 0.5,
                 mu: float = 0.3, f_ext=None, fixed_dofs=None, config: ContactConfig | None = None,
                 time: GeneralizedAlphaProcess | None = None, dt: float = 0.0, core_radii=None,
                 ga_state: bool = True):
        torch = ctx.torch
        self.ctx = ctx
        self.torch = torch
        self.k_pen = float(k_pen)
        self.smoothing_delta = float(smoothing_delta)
        self.k_t_ratio = float(k_t_ratio)
        self.mu = float(mu)
        self.cfg = config or ContactConfig(mu=mu, k_t_ratio=k_t_ratio)
        self._layer = layer_array(self.cfg, ctx.n_elems)
        self.coating = self.cfg.coating_stiffness > 0.0
        self.n_coat_rebuilds = 0
        if self.coating:
            ctx.set_coating(self.cfg.coating_stiffness, self.cfg.coating_damping, self.cfg.coating_mu,
                            self.cfg.coating_k_t_ratio, core_radii)
        n = ctx.ndof
        dev = ctx.dev
        self.f_ext = ctx.to_dev(np.zeros(n) if f_ext is None else f_ext)
        mask = np.zeros(n, np.uint8)
        if fixed_dofs is not None:
            mask[np.asarray(fixed_dofs, np.int64)] = 1
        self.fixed_mask = torch.as_tensor(mask, device=dev)
        self.u = torch.zeros(n, dtype=torch.float64, device=dev)
        self.u_ref = torch.zeros(n, dtype=torch.float64, device=dev)
        self.R = torch.zeros(n, dtype=torch.float64, device=dev)
        self.res2 = torch.zeros(1, dtype=torch.float64, device=dev)
        self.n_active = torch.zeros(1, dtype=torch.int32, device=dev)
        self.K = None
        self.time = time
        self.dt = float(dt)
        # each assembly stores the attempt's acceleration and velocity into the time
        # integrator (GeneralizedAlphaProcess.correct, as the reference's loop does per
        # attempt, _newton_dynamic.py:190); a driver that corrects at the increment's
        # end (process.py:549-551), before anything reads them, passes ga_state=False
        self.ga_state = bool(ga_state)
        self._dyn = time.dyn_struct(self.dt, state=self.ga_state) if time is not None else None
        self.graph = None
        # learned prescreen (set_prescreen): network, feature layer ids, logging
        self.prescreen = None
        self.lp_layer = None
        self.lp_scale = 0.125
        self.lp_log = None
        self.n_watch = None
        self._lp_feat = None

    def set_prescreen(self, net, *, elem_layer=None, layer_scale: float = 0.125):
        """Run the per-iteration pair kernel over the learned watch list (None: all pairs)."""
        self.prescreen = net
        self.lp_layer = self._layer_dev(elem_layer if elem_layer is not None else self._layer)
        self.lp_scale = layer_scale
        if net is not None:
            self.ctx.lp_set(net.host_params(), net.threshold)

    def log_features(self, on: bool = True, *, elem_layer=None, layer_scale: float = 0.125):
        """Record (features at increment start, pair gaps at the end of the
        Newton call) for every call, as training data."""
        self.lp_log = [] if on else None
        if elem_layer is not None or self.lp_layer is None:
            self.lp_layer = self._layer_dev(elem_layer if elem_layer is not None else self._layer)
        self.lp_scale = layer_scale

    def _layer_dev(self, layer):
        """Per-element layer ids as a device int32 tensor (converted once, read every increment)."""
        return None if layer is None else self.ctx.to_dev(layer, self.torch.int32)

    def set_u(self, u):
        self.u.copy_(self.ctx.to_dev(u))

    def set_u_ref(self, u_ref):
        self.u_ref.copy_(self.ctx.to_dev(u_ref))

    def set_f_ext(self, f_ext):
        self.f_ext.copy_(self.ctx.to_dev(f_ext))

    def set_coat_dt(self, coat_dt: float):
        """Time step of the coating dashpot: max(load_frac - load_frac_prev, 1e-15)
        (contact/solver/_newton_steps.py:112)."""
        self.ctx.coating_set_dt(coat_dt)

    def commit_coating(self):
        """End of a converged step: compression_prev <- compression."""
        if self.coating:
            self.ctx.coating_commit()

    def set_dt(self, dt: float):
        """Time step of the inertia terms (a captured graph is dropped)."""
        self.dt = float(dt)
        self._dyn = self.time.dyn_struct(self.dt, state=self.ga_state) if self.time is not None else None
        self.graph = None

    def deformed_coords(self, u=None):
        """x = X + u[0:3] per node (``_deformed_coords``, contact/solver/_utils.py:16-28)."""
        u = self.u if u is None else self.ctx.to_dev(u)
        return (self.ctx.X + u.view(-1, 6)[:, :3]).contiguous()

    def begin_increment(self, u=None, *, detect: bool = True, geometry: bool = True, margin: float = 0.0,
                        cell_size=None):
        """Once-per-increment: GPU DetectCandidates and/or UpdateGeometry
        (unfrozen) on x = X + u, then the tangent pattern of the increment."""
        if u is not None:
            self.set_u(u)
        x = self.deformed_coords()
        cfg = self.cfg
        if detect:
            self.ctx.detect_candidates(x, margin=margin, cell_size=cell_size, elem_layer=self._layer,
                                       midpoint_prescreening=cfg.midpoint_prescreening,
                                       prescreening_margin=cfg.prescreening_margin)
        if geometry:
            self.ctx.pairs_run(L.PAIR_GEOM | L.PAIR_HYST, coords=x, g_on=cfg.g_on, g_off=cfg.g_off)
            self.K = self.ctx.new_K()  # builds the pattern (synchronous)
            self.n_watch = None
            if self.prescreen is not None or self.lp_log is not None:
                f = self.ctx.lp_features(self.u, self.lp_layer, self.lp_scale, log=self.lp_log is not None)
                self._lp_feat = None if f is None else f.cpu().numpy()
                if self.prescreen is not None:
                    self.n_watch = self.ctx.lp_select()
        else:
            self.K = None
        self.graph = None
        return self

    def end_call(self) -> int:
        """After a Newton call: the verification pass (exact narrowphase over the pairs the
        prescreen left out, at the last assembled u) -> #misses; logs the
        training rows when logging."""
        miss = 0
        if self.n_watch is not None:
            miss = self.ctx.lp_verify(self.u, self.cfg.g_on)
            self.n_watch = None
        if self.lp_log is not None and self._lp_feat is not None:
            self.lp_log.append((self._lp_feat, self.ctx.get_pairs()["gap"].copy()))
            self._lp_feat = None
        return miss

    def count_penetrations(self, u=None) -> int:
        """#(gap < 0) over the pair list at x = X + u without changing any pair
        state (the initial-penetration check, contact/solver/_initial_penetration.py:117-142)."""
        if u is not None:
            self.set_u(u)
        snap = self.ctx.get_pairs()
        if len(snap["elem_a"]) == 0:
            return 0
        coat = self.ctx.get_coating() if self.coating else None
        self.ctx.pairs_run(L.PAIR_GEOM, coords=self.deformed_coords())
        gap = self.ctx.get_pairs()["gap"]   # the core gap with the coating on (_initial_penetration.py:134-137)
        self.ctx.set_pairs(snap["elem_a"], snap["elem_b"], state=snap)
        if coat is not None:
            self.ctx.set_coating_state(coat)
        return int(np.sum(gap < 0.0))

    def assemble(self, u=None, *, tangent: bool = True, controlled: bool = False):
        """Launch one fused assembly; returns device tensors (R, K, ||R||^2, n_active).
        With the coating on, the pattern probe reads its count on the host, or,
        with ``controlled`` (a Newton control record attached), leaves its outcome
        in the record (XK_NC_REBUILD skips the rest of the attempt; the host then
        calls ``coating_rebuild``), so the assembly needs no host read."""
        if u is not None:
            self.set_u(u)
        if self.coating and not controlled and self.ctx.coating_probe(self.u) > 0:
            self.K = None   # a pair pressing its coating has no slot yet: new pattern
            self.graph = None
            self.n_coat_rebuilds += 1
        if self.K is None:
            self.K = self.ctx.new_K()
        if self.coating and controlled:
            self.ctx.coating_probe_control(self.u)
        self.ctx.assemble_iteration(self.u, self.u_ref, self.f_ext, self.fixed_mask, self.R,
                                    self.K if tangent else None, self._dyn, k_pen=self.k_pen,
                                    smoothing_delta=self.smoothing_delta, k_t_ratio=self.k_t_ratio,
                                    mu_eff=self.mu, res_norm2=self.res2, n_active=self.n_active)
        return self.R, (self.K if tangent else None), self.res2, self.n_active

    def coating_rebuild(self):
        """After a controlled attempt stopped at XK_NC_REBUILD (its probe met a
        compressed pair without a slot, at the same u): probe on the host, which
        marks the pattern stale, so that the next assembly rebuilds it."""
        if self.ctx.coating_probe(self.u) == 0:
            raise RuntimeError("XK_NC_REBUILD without an unplaced pair at the same displacement")
        self.K = None
        self.graph = None
        self.n_coat_rebuilds += 1

    # -------------------------------------------------------------- HIP graph
    def capture(self, *, tangent: bool = True):
        """Capture one assembly (reading ``self.u``) into a HIP graph."""
        if self.coating:
            raise RuntimeError("an assembly with the coating on probes the pattern on the host: not capturable")
        torch = self.torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.assemble(tangent=tangent)  # warm (pattern already built)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.assemble(tangent=tangent)
        self.graph = g
        return g

    def replay(self):
        self.graph.replay()
        return self.R, self.K, self.res2, self.n_active

    def K_csr(self):
        return self.ctx.to_csr(self.K)
