"""Dynamic frictional-contact Newton driver around the device assembly.

Restates, for the strand path (smooth-penalty Huber contact, Coulomb friction,
Generalized-alpha dynamics, optionally the Kelvin-Voigt coating, no prescribed
displacements), the
reference's solver loop so that the assembly on the GPU can be driven through
whole solves and its Newton trajectory compared with the reference's:

  ContactFrictionProcess.process   contact/solver/process.py:139-674
      increment loop, predictor, per-increment DetectCandidates +
      UpdateGeometry, cutback with checkpoint restore, accept / checkpoint
  NewtonDynamicProcess.process     contact/solver/_newton_dynamic.py:115-389
      attempts, force / displacement / energy convergence, divergence
      detection, line search, update
  ConvergenceCheckProcess          contact/solver/_newton_steps.py:182-243
  LinearSolveProcess               contact/solver/_newton_steps.py:338-366
  LineSearchUpdateProcess          contact/solver/_newton_steps.py:396-434
  _ncp_line_search                 contact/solver/_utils.py:31-61
  UnifiedTimeStepProcess           contact/solver/_unified_time_controller.py:103-215
  AdaptiveSteppingProcess          contact/solver/_adaptive_stepping.py

The per-attempt work (narrowphase, forces, tangent, effective terms, norm) is
one call of the backend's ``assemble``; the linear solve is SciPy's SuperLU on
the host, as in the reference.  The backend protocol is small so that the same
driver runs on the device (``DeviceBackend``) or on any other assembly with
the reference's semantics.  ``DeviceBackend(device_solve=True)`` instead runs
the Newton attempts of an increment on the device: the assembly, the linear
solve (``xk_linear_solve``), the line search and the updates stay in HBM, and
the driver's convergence and divergence tests run on the device as well
(``xk_newton_control_*``): attempts are enqueued back to back and the host
reads each attempt's control record afterwards.
"""

from __future__ import annotations

import math
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import scipy.sparse as sp

from . import _lib as L

# device status word bits (XK_FLAG_* of include/xkamd.h)
FLAG_NONFINITE, FLAG_DEGENERATE, FLAG_SINGULAR = 1, 2, 4


@dataclass(frozen=True)
class SolverConfig:
    """Solver knobs with the reference's names and defaults
    (ContactFrictionInputData, core/data.py:156-193; _ContactConfigInput
    adaptive-stepping fields, contact/_contact_pair.py:146-153;
    NewtonDynamicInput, _newton_dynamic.py:36-47)."""

    max_nr_attempts: int = 50
    tol_force: float = 1e-8
    tol_disp: float = 1e-8
    divergence_window: int = 5
    du_norm_cap: float = 0.0
    max_increments: int = 10000
    use_line_search: bool = True
    line_search_max_steps: int = 5
    dt_grow_factor: float = 1.5
    dt_shrink_factor: float = 0.5
    dt_grow_attempt_threshold: int = 5
    dt_shrink_attempt_threshold: int = 15
    dt_contact_change_threshold: float = 0.3
    dt_min_fraction: float = 0.0
    dt_max_fraction: float = 0.0


# ---------------------------------------------------------------- stepping
class AdaptiveStepping:
    """Load-fraction queue with growth / shrink / cutback
    (AdaptiveSteppingProcess, _adaptive_stepping.py) behind the physical-time
    interface of UnifiedTimeStepProcess (_unified_time_controller.py:103-215),
    configured as ContactFrictionProcess does (process.py:332-354)."""

    def __init__(self, t_total: float, cfg: SolverConfig):
        if t_total <= 0.0:
            raise ValueError("t_total must be positive")
        self.t_total = t_total
        dt_initial = t_total
        dt_min = cfg.dt_min_fraction * t_total if cfg.dt_min_fraction > 0 else t_total / 32.0
        dt_max = cfg.dt_max_fraction * t_total if cfg.dt_max_fraction > 0 else t_total
        self.grow = cfg.dt_grow_factor
        self.shrink = cfg.dt_shrink_factor
        self.grow_thr = cfg.dt_grow_attempt_threshold if cfg.dt_grow_attempt_threshold > 0 else 5
        self.shrink_thr = cfg.dt_shrink_attempt_threshold
        self.contact_thr = cfg.dt_contact_change_threshold
        f_init = dt_initial / t_total
        self.f_min = dt_min / t_total
        self.f_max = dt_max / t_total
        n_eff = max(1, int(1.0 / f_init) if f_init > 0 else 1)
        if self.f_min <= 0.0:
            self.f_min = 1.0 / (n_eff * 32)
        if self.f_max <= 0.0:
            self.f_max = min(8.0 / n_eff, 1.0)
        base = f_init if f_init > 0.0 else 1.0
        if self.f_max > 0.0:
            base = min(base, self.f_max)
        self.queue: deque[float] = deque([min(base, 1.0)])
        self.consecutive_good = 0
        self.n_cutbacks = 0
        self.last_prev = 0.0

    def query(self, load_frac_prev: float):
        """(has_more, load_frac, dt_sub)."""
        while self.queue:
            nxt = self.queue[0]
            if nxt <= load_frac_prev + 1e-15:
                self.queue.popleft()
                continue
            dt_sub = max((nxt - self.last_prev) * self.t_total, 0.0)
            return True, nxt, dt_sub
        return False, 0.0, 0.0

    def success(self, load_frac, load_frac_prev, n_attempts, n_active, prev_n_active):
        self.last_prev = load_frac
        if self.queue and self.queue[0] <= load_frac + 1e-15:
            self.queue.popleft()
        if load_frac >= 1.0 - 1e-12:
            return
        cur = load_frac - load_frac_prev
        nxt = cur
        if n_attempts <= self.grow_thr:
            self.consecutive_good += 1
            if self.consecutive_good <= 2:
                nxt = cur * self.grow
            else:
                damp = max(0.1, 1.0 / self.consecutive_good)
                nxt = cur * (1.0 + (self.grow - 1.0) * damp)
        elif n_attempts >= self.shrink_thr:
            nxt = cur * self.shrink
            self.consecutive_good = 0
        else:
            self.consecutive_good = 0
        changed = abs(n_active - prev_n_active)
        if changed > 0:
            base = max(prev_n_active, n_active, 1)
            if changed / base > self.contact_thr:
                nxt = min(nxt, cur * self.shrink)
                self.consecutive_good = 0
        nxt = max(nxt, self.f_min)
        nxt = min(nxt, self.f_max)
        nf = min(load_frac + nxt, 1.0)
        if 1.0 - nf < self.f_min * 0.5:
            nf = 1.0
        self.queue.appendleft(nf)

    def failure(self, load_frac, load_frac_prev, diverged) -> bool:
        """True if the increment can be retried with a smaller step."""
        delta = load_frac - load_frac_prev
        if delta <= self.f_min + 1e-15:
            return False
        self.consecutive_good = 0
        self.n_cutbacks += 1
        shrink = self.shrink * self.shrink if diverged else self.shrink
        mid = load_frac_prev + delta * shrink
        self.queue.appendleft(load_frac)
        self.queue.appendleft(mid)
        return True


# ---------------------------------------------------------------- Newton pieces
def linear_solve(K: sp.spmatrix, R: np.ndarray, fixed_dofs: np.ndarray):
    """du solving K du = -R with fixed rows / columns replaced by identity
    (LinearSolveProcess, _newton_steps.py:351-366); None if SuperLU fails."""
    from scipy.sparse.linalg import spsolve

    n = K.shape[0]
    free = np.ones(n)
    free[fixed_dofs] = 0.0
    D = sp.diags(free)
    K_eff = (D @ K.astype(np.float64) @ D + sp.diags(1.0 - free)).tocsc()
    K_eff.eliminate_zeros()
    rhs = -np.asarray(R, float).copy()
    rhs[fixed_dofs] = 0.0
    try:
        du = spsolve(K_eff, rhs)
    except Exception:
        return None
    return du


def line_search(u, du, f_ext, fixed_dofs, internal_force, res_u_norm, max_steps, f_c, diverge_factor=3.0) -> float:
    """Step length: halve until the frozen-contact static residual
    f_int(u + a du) + f_c - f_ext is finite and below diverge_factor ||R||
    (_ncp_line_search, contact/solver/_utils.py:31-61)."""
    alpha = 1.0
    for _ in range(max_steps):
        try:
            f_try = internal_force(u + alpha * du)
        except Exception:
            alpha *= 0.5
            continue
        R_try = f_try + f_c - f_ext
        R_try[fixed_dofs] = 0.0
        if not np.all(np.isfinite(R_try)):
            alpha *= 0.5
            continue
        if float(np.linalg.norm(R_try)) < diverge_factor * max(res_u_norm, 1e-30):
            return alpha
        alpha *= 0.5
    return alpha


@dataclass
class NewtonOutcome:
    converged: bool
    n_attempts: int
    n_active: int
    diverged: bool
    f_c: np.ndarray
    res_history: list = field(default_factory=list)
    active_history: list = field(default_factory=list)


def newton_dynamic(backend, u: np.ndarray, u_ref: np.ndarray, f_ext: np.ndarray, fixed_dofs: np.ndarray,
                   dt: float, f_ext_ref_norm: float, dynamic_ref: bool, cfg: SolverConfig) -> NewtonOutcome:
    """One increment's Newton iterations (NewtonDynamicProcess.process,
    _newton_dynamic.py:115-389).  ``u`` is updated in place.  A backend that
    runs the whole loop itself (``device_newton``) is handed the call."""
    if getattr(backend, "device_newton", False):
        return backend.newton(u, u_ref, f_ext, fixed_dofs, dt, f_ext_ref_norm, dynamic_ref, cfg)
    energy_ref = None
    incr_f_ref = 0.0
    consecutive_increase = 0
    prev_ratio = float("inf")
    out = NewtonOutcome(False, 0, 0, False, np.zeros_like(u))
    for att in range(cfg.max_nr_attempts):
        out.n_attempts += 1
        R, K, n_active, f_c = backend.assemble(u, u_ref, f_ext, dt)
        out.f_c = f_c
        # ConvergenceCheckProcess, force part (_newton_steps.py:194-213)
        res = float(np.linalg.norm(R))
        f_ref = incr_f_ref if incr_f_ref > 1e-30 else f_ext_ref_norm
        if dynamic_ref and att == 0 and res > 1e-30:
            f_ref = res
        if att == 0 and dynamic_ref:
            incr_f_ref = f_ref
        out.n_active = n_active
        ratio = res / f_ref
        out.res_history.append(ratio)
        out.active_history.append(n_active)
        if ratio < cfg.tol_force:
            out.converged = True
            break
        # divergence (_newton_dynamic.py:258-285)
        if att > 0 and ratio > prev_ratio * 1.01:
            consecutive_increase += 1
        else:
            consecutive_increase = 0
        prev_ratio = ratio
        if (consecutive_increase >= cfg.divergence_window and not ratio < 0.01) or (att >= 10 and ratio > 100.0):
            out.diverged = True
            break
        if not math.isfinite(res):
            break   # a non-finite system has no solution: LinearSolveProcess success=False (_newton_steps.py:360-366)
        du = linear_solve(K, R, fixed_dofs)
        if du is None:
            break
        scale = 1.0
        if cfg.use_line_search:
            scale *= line_search(u, du, f_ext, fixed_dofs, backend.internal_force, res,
                                 cfg.line_search_max_steps, f_c)
        if cfg.du_norm_cap > 0.0:
            dn = float(np.linalg.norm(scale * du))
            un = max(float(np.linalg.norm(u)), 1.0)
            if dn > cfg.du_norm_cap * un:
                scale *= cfg.du_norm_cap * un / dn
        du = scale * du
        u += du
        # ConvergenceCheckProcess, displacement / energy part (_newton_steps.py:215-243)
        du_norm = float(np.linalg.norm(du))
        u_norm = float(np.linalg.norm(u))
        if u_norm > 1e-30 and du_norm / u_norm < cfg.tol_disp:
            out.converged = True
            break
        energy = abs(float(np.dot(du, R)))
        if energy_ref is None:
            energy_ref = energy if energy > 1e-30 else 1.0
        if energy_ref > 1e-30 and energy / energy_ref < 1e-10:
            out.converged = True
            break
    return out


# ---------------------------------------------------------------- driver
@dataclass
class NewtonCall:
    """One Newton call (an increment attempt, cutbacks included)."""

    load_frac: float
    load_frac_prev: float
    dt: float
    converged: bool
    n_attempts: int
    n_active: int
    diverged: bool
    res_history: list
    active_history: list


@dataclass
class SolverResult:
    u: np.ndarray
    converged: bool
    n_increments: int
    total_attempts: int
    n_cutbacks: int
    calls: list
    n_rechecks: int = 0          # Newton calls re-run on the full path after a prescreen miss
    watch: list = field(default_factory=list)   # (pairs watched, pairs) per call with the prescreen on


class ContactFrictionSolver:
    """The dynamic increment loop of ContactFrictionProcess
    (contact/solver/process.py:139-674) over a backend that assembles.

    backend protocol (the prescreen hooks watch_count / snapshot_pairs /
    restore_pairs / end_call are optional):
      detect(u)                 DetectCandidatesProcess on X + u
      update_geometry(u)        UpdateGeometryProcess (hysteresis on) on X + u
      initial_penetrations(u)   #(gap < 0) over the pairs, no state change
      predict(u, dt) -> u_pred  Generalized-alpha predictor
      assemble(u, u_ref, f_ext, dt) -> (R_eff, K_eff, n_active, f_c)
      internal_force(u) -> f_int
      correct(u, dt); checkpoint(); restore_checkpoint()
    and, when its ``coating`` attribute is true, set_coat_dt(coat_dt) and
    commit_coating().
    """

    def __init__(self, backend, *, fixed_dofs, f_ext_total, t_total: float, f_ext_base=None,
                 config: SolverConfig | None = None):
        self.b = backend
        self.cfg = config or SolverConfig()
        self.fixed = np.asarray(fixed_dofs, np.int64)
        self.f_total = np.asarray(f_ext_total, float)
        self.f_base = np.zeros_like(self.f_total) if f_ext_base is None else np.asarray(f_ext_base, float)
        self.t_total = float(t_total)

    def solve(self, u0=None) -> SolverResult:
        cfg = self.cfg
        ndof = len(self.f_total)
        u = np.zeros(ndof) if u0 is None else np.asarray(u0, float).copy()
        u_ref = u.copy()
        f_ref_norm = float(np.linalg.norm(self.f_total))
        dynamic_ref = f_ref_norm < 1e-30
        if dynamic_ref:
            f_ref_norm = 1.0
        b = self.b
        # set-up candidates and the initial-penetration check (process.py:258-324;
        # with the coating on a core penetration is allowed, :320)
        coat = bool(getattr(b, "coating", False))
        b.detect(u)
        n_pen = b.initial_penetrations(u)
        if n_pen > 0 and not coat:
            raise ValueError(f"initial penetration detected in {n_pen} pairs")
        ckpt = (u.copy(), u_ref.copy())
        b.checkpoint()
        stepping = AdaptiveStepping(self.t_total, cfg)
        calls, watches, n_rechecks = [], [], 0
        load_frac_prev, prev_n_active, n_incr, incr_count, total = 0.0, 0, 0, 0, 0
        while True:
            if cfg.max_increments > 0 and incr_count >= cfg.max_increments:
                break
            more, load_frac, dt = stepping.query(load_frac_prev)
            if not more:
                break
            incr_count += 1
            n_incr += 1
            f_ext = self.f_base + load_frac * self.f_total
            if coat:
                b.set_coat_dt(max(load_frac - load_frac_prev, 1e-15))   # _newton_steps.py:112, :297
            u = b.predict(u, dt)
            b.detect(u)
            b.update_geometry(u)
            watch = b.watch_count() if hasattr(b, "watch_count") else None
            if watch is not None:
                u_start = u.copy()
                b.snapshot_pairs()
                watches.append(watch)
            step = newton_dynamic(b, u, u_ref, f_ext, self.fixed, dt, f_ref_norm, dynamic_ref, cfg)
            if hasattr(b, "end_call") and b.end_call() > 0:
                # a pruned pair ended the call within contact range: re-run the
                # call on the full path from the same start
                n_rechecks += 1
                total += step.n_attempts
                u = u_start
                b.restore_pairs()
                step = newton_dynamic(b, u, u_ref, f_ext, self.fixed, dt, f_ref_norm, dynamic_ref, cfg)
                b.end_call()
            total += step.n_attempts
            calls.append(NewtonCall(load_frac, load_frac_prev, dt, step.converged, step.n_attempts, step.n_active,
                                    step.diverged, step.res_history, step.active_history))
            if not step.converged:
                if stepping.failure(load_frac, load_frac_prev, step.diverged):
                    u, u_ref = ckpt[0].copy(), ckpt[1].copy()
                    b.restore_checkpoint()
                    n_incr -= 1
                    continue
                return SolverResult(u, False, n_incr, total, stepping.n_cutbacks, calls, n_rechecks, watches)
            if coat:
                b.commit_coating()   # coating_compression_prev <- coating_compression (process.py:539-547)
            b.correct(u, dt)
            stepping.success(load_frac, load_frac_prev, step.n_attempts, step.n_active, prev_n_active)
            prev_n_active = step.n_active
            load_frac_prev = load_frac
            u_ref = u.copy()
            ckpt = (u.copy(), u_ref.copy())
            b.checkpoint()
        return SolverResult(u, True, n_incr, total, stepping.n_cutbacks, calls, n_rechecks, watches)


# ---------------------------------------------------------------- device backend
class DeviceBackend:
    """The solver backend on one GPU: candidates, geometry, the fused assembly
    (R_eff, K_eff, ||R||^2, #active in one graph-capturable launch sequence),
    the element-only internal force of the line search and the
    Generalized-alpha state all live on the device.  By default R, K and f_c
    are copied to the host for the SuperLU solve; with ``device_solve`` whole
    Newton calls run on the device (``newton``)."""

    def __init__(self, ctx, *, M_diag, rho_inf: float, k_pen: float, smoothing_delta: float, mu: float,
                 k_t_ratio: float = 0.5, fixed_dofs=None, config=None, device_solve: bool = False,
                 use_graph: bool = False, core_radii=None, device_control: bool = True):
        from .assembly import StrandAssembly
        from .time_integration import GeneralizedAlphaProcess

        self.ctx = ctx
        self.torch = ctx.torch
        self.device_newton = bool(device_solve)
        # device Newton attempts: the driver's tests on the device (attempts enqueued
        # back to back) or on the host (a host read after each attempt's stages)
        self.device_control = bool(device_control)
        self.solve_residuals = []   # device solve: relative residual of every linear solve
        self.solve_tol = 1e-11      # device solve: target relative residual (SuperLU reaches about 1e-12)
        self.max_refine = 16
        self.use_graph = bool(device_solve and use_graph)   # attempts after the first as HIP-graph replays
        self.graph_refine = 3       # fixed refinement steps inside the captured solve (host tests)
        self.ctl_refine = 2         # most refinement steps of a device-tested attempt (strands need 2);
                                    # a solve they leave above solve_tol is re-solved by the host
        self.graph_attempts_captured = 0
        self.graph_error = None
        self.ga = GeneralizedAlphaProcess(M_diag, rho_inf=rho_inf)
        # the attempts' acceleration and velocity are set by correct() at the increment's end
        self.asm = StrandAssembly(ctx, k_pen=k_pen, smoothing_delta=smoothing_delta, k_t_ratio=k_t_ratio, mu=mu,
                                  fixed_dofs=fixed_dofs, config=config, time=self.ga, dt=0.0,
                                  core_radii=core_radii, ga_state=False)
        self._f = self.torch.zeros(ctx.ndof, dtype=self.torch.float64, device=ctx.dev)
        self._dt = None
        self.flags = 0              # every XK_FLAG_* bit the device status word has reported
        self.solve_fail_tol = 1e-6  # a solve whose refined residual stays above this counts as failed
        self.eager_resolves = 0     # graph attempts whose fixed refinement missed solve_tol, re-solved eagerly
        self.control_reasons = []   # device-controlled calls: XK_NCR_* bits of each call's end

    def _read_flags(self) -> int:
        """Read (and clear) the device status word, keep its bits, and raise as
        the reference does on a degenerate element (the scalar CR routine's
        ValueError, elements/_beam_cr.py:54-55)."""
        f = self.ctx.status_flags()
        self.flags |= f
        if f & FLAG_DEGENERATE:
            raise ValueError("zero-length beam element met by the element kernel (XK_FLAG_DEGENERATE)")
        return f

    def detect(self, u):
        self.asm.begin_increment(u, geometry=False)

    def update_geometry(self, u):
        self.asm.begin_increment(u, detect=False)

    def initial_penetrations(self, u) -> int:
        return self.asm.count_penetrations(u)

    @property
    def coating(self) -> bool:
        return self.asm.coating

    def set_coat_dt(self, coat_dt: float):
        self.asm.set_coat_dt(coat_dt)

    def commit_coating(self):
        self.asm.commit_coating()

    def predict(self, u, dt):
        return self.ga.predict(u, dt)

    def assemble(self, u, u_ref, f_ext, dt):
        asm = self.asm
        if dt != self._dt:
            asm.set_dt(dt)
            self._dt = dt
        asm.set_u(u)
        asm.set_u_ref(u_ref)
        asm.set_f_ext(f_ext)
        R, K, _, n_act = asm.assemble()
        self.ctx.pairs_gather(f=self._f)
        self._read_flags()
        return (R.cpu().numpy(), asm.K_csr(), int(n_act.item()), -self._f.cpu().numpy())

    def internal_force(self, u):
        self.ctx.beam(self.ctx.to_dev(u), f_int=self._f)
        self._read_flags()
        return self._f.cpu().numpy()

    def correct(self, u, dt):
        self.ga.correct(u, None, dt)

    def newton(self, u, u_ref, f_ext, fixed_dofs, dt, f_ext_ref_norm, dynamic_ref, cfg) -> NewtonOutcome:
        """``newton_dynamic`` with the vectors in HBM: the linear solve
        (LinearSolveProcess -> xk_linear_solve) and the line search + update
        (LineSearchUpdateProcess -> xk_newton_update) on the device, with the
        same control flow, thresholds and order of updates.  With
        ``device_control`` (the default) the driver's tests run on the device
        too (``_newton_controlled``).  Otherwise the host reads (||R||^2,
        n_active) after the assembly, the solve's residual, and (alpha,
        ||du||, ||u||, |du.R|) after the update of every attempt; with
        ``use_graph`` the attempts after the first replay three captured HIP
        graphs: G1 = assembly + pair force, G2 = linear solve (a fixed number
        of refinement steps), G3 = line search + update.  ``u`` (host) is
        updated in place at the end."""
        torch, asm, ctx = self.torch, self.asm, self.ctx
        if self.device_control:
            return self._newton_controlled(u, u_ref, f_ext, dt, f_ext_ref_norm, dynamic_ref, cfg)
        if dt != self._dt:
            asm.set_dt(dt)
            self._dt = dt
        asm.set_u(u)
        asm.set_u_ref(u_ref)
        asm.set_f_ext(f_ext)
        ud = asm.u
        du = torch.empty_like(ud)
        status = torch.zeros(24, dtype=torch.float64, device=ctx.dev)
        ls_steps = cfg.line_search_max_steps if cfg.use_line_search else 0
        graphs = None
        # with the coating on every assembly probes the pattern first (host sync): eager attempts
        graph_ok = self.use_graph and asm.prescreen is None and not asm.coating
        energy_ref = None
        incr_f_ref = 0.0
        consecutive_increase = 0
        prev_ratio = float("inf")
        out = NewtonOutcome(False, 0, 0, False, None)

        def solve(rel, refine, tol):
            ctx.linear_solve(ud, asm.fixed_mask, asm.R, du, asm._dyn, max_refine=refine, tol=tol, res_rel=rel)

        def update():
            ctx.newton_update(ud, du, asm.R, self._f, asm.f_ext, asm.fixed_mask, asm.res2, status,
                              line_search_steps=ls_steps, diverge_factor=3.0, du_norm_cap=cfg.du_norm_cap)

        for att in range(cfg.max_nr_attempts):
            out.n_attempts += 1
            if graphs is None:
                asm.assemble()
                ctx.pairs_gather(f=self._f)
            else:
                graphs[0].replay()
            out.f_c = self._f
            # ConvergenceCheckProcess, force part (_newton_steps.py:194-213)
            res = math.sqrt(float(asm.res2.item()))
            n_active = int(asm.n_active.item())
            flags = self._read_flags()
            f_ref = incr_f_ref if incr_f_ref > 1e-30 else f_ext_ref_norm
            if dynamic_ref and att == 0 and res > 1e-30:
                f_ref = res
            if att == 0 and dynamic_ref:
                incr_f_ref = f_ref
            out.n_active = n_active
            ratio = res / f_ref
            out.res_history.append(ratio)
            out.active_history.append(n_active)
            if ratio < cfg.tol_force:
                out.converged = True
                break
            # divergence (_newton_dynamic.py:258-285)
            if att > 0 and ratio > prev_ratio * 1.01:
                consecutive_increase += 1
            else:
                consecutive_increase = 0
            prev_ratio = ratio
            if (consecutive_increase >= cfg.divergence_window and not ratio < 0.01) or (att >= 10 and ratio > 100.0):
                out.diverged = True
                break
            if flags & FLAG_NONFINITE or not math.isfinite(res):
                break   # a non-finite system has no solution: LinearSolveProcess success=False
            rel = torch.zeros(1, dtype=torch.float64, device=ctx.dev)
            if graphs is None:
                solve(rel, self.max_refine, self.solve_tol)
            else:
                rel = graphs[3]
                graphs[1].replay()
                if not float(rel.item()) <= self.solve_tol:   # NaN included
                    # the captured solve's fixed refinement fell short: solve this attempt eagerly
                    self.eager_resolves += 1
                    rel = torch.zeros(1, dtype=torch.float64, device=ctx.dev)
                    solve(rel, self.max_refine, self.solve_tol)
            rel_v = float(rel.item())
            self.solve_residuals.append(rel.clone())
            if self._read_flags() & FLAG_SINGULAR or not rel_v <= self.solve_fail_tol:
                break   # zero pivot or no accurate solution: the solve failed (LinearSolveProcess success=False)
            if graphs is None:
                update()
            else:
                graphs[2].replay()
            st = status[:4].cpu().numpy()
            du_norm, u_norm, energy = float(st[1]), float(st[2]), float(st[3])
            # ConvergenceCheckProcess, displacement / energy part (_newton_steps.py:215-243)
            if u_norm > 1e-30 and du_norm / u_norm < cfg.tol_disp:
                out.converged = True
                break
            if energy_ref is None:
                energy_ref = energy if energy > 1e-30 else 1.0
            if energy_ref > 1e-30 and energy / energy_ref < 1e-10:
                out.converged = True
                break
            if graph_ok and graphs is None:
                graphs = self._capture_attempt(solve, update)
                if graphs is None:
                    graph_ok = False
                else:
                    self.graph_attempts_captured += 1
        u[:] = ud.cpu().numpy()
        return out

    def _newton_controlled(self, u, u_ref, f_ext, dt, f_ext_ref_norm, dynamic_ref, cfg) -> NewtonOutcome:
        """``newton_dynamic`` with its tests on the device (xk_newton_control_*).
        Each attempt -- assembly, pair force, force and divergence tests,
        linear solve, solve check, line search, update, displacement / energy
        tests and the attempt limit -- is enqueued without a host read; the
        refinement of the solve stops on the device (at most ``ctl_refine``
        steps); the host keeps one
        attempt queued ahead and reads each attempt's control record
        afterwards from pinned memory, so the device never waits on it.
        Every attempt kernel skips once the call has ended, so the attempt
        enqueued past the end changes nothing.  An attempt whose solve missed
        ``solve_tol`` is handed to the host (XK_NC_ASSIST), which re-solves it
        eagerly as the host-tested path does.  With the coating on, an attempt
        whose pattern probe met a compressed pair without a slot stops at
        XK_NC_REBUILD: the host rebuilds the pattern and enqueues the attempt
        again.  With ``use_graph`` the attempts after each increment's first
        (and after a rebuild or an assist) replay one captured graph."""
        torch, asm, ctx = self.torch, self.asm, self.ctx
        if dt != self._dt:
            asm.set_dt(dt)
            self._dt = dt
        asm.set_u(u)
        asm.set_u_ref(u_ref)
        asm.set_f_ext(f_ext)
        ud = asm.u
        du = torch.empty_like(ud)
        status = torch.zeros(24, dtype=torch.float64, device=ctx.dev)
        ls_steps = cfg.line_search_max_steps if cfg.use_line_search else 0
        cap = int(cfg.max_nr_attempts)
        ctl = torch.zeros(L.NC_HIST + L.NC_HSTRIDE * cap, dtype=torch.float64, device=ctx.dev)
        rel = ctl[L.NC_REL:L.NC_REL + 1]   # the solve writes its residual into the record
        pin = torch.zeros((2, L.NC_HIST), dtype=torch.float64, pin_memory=True)
        evs = [torch.cuda.Event(), torch.cuda.Event()]
        graph_ok = self.use_graph and asm.prescreen is None
        graph = None

        def solve(refine, tol):
            ctx.linear_solve(ud, asm.fixed_mask, asm.R, du, asm._dyn, max_refine=refine, tol=tol, res_rel=rel)

        def update():
            ctx.newton_update(ud, du, asm.R, self._f, asm.f_ext, asm.fixed_mask, asm.res2, status,
                              line_search_steps=ls_steps, diverge_factor=3.0, du_norm_cap=cfg.du_norm_cap)

        def attempt():
            asm.assemble(controlled=True)
            ctx.pairs_gather(f=self._f)
            ctx.newton_control_force(asm.res2, asm.n_active)
            solve(self.ctl_refine, self.solve_tol)   # refinement stopped on the device
            update()

        def post(slot):
            pin[slot].copy_(ctl[:L.NC_HIST], non_blocking=True)
            evs[slot].record()

        self._read_flags()   # flags of the increment's set-up (raises on a degenerate element)
        ctx.newton_control_begin(ctl, hist_cap=cap, tol_force=cfg.tol_force, tol_disp=cfg.tol_disp,
                                 f_ext_ref_norm=f_ext_ref_norm, dynamic_ref=dynamic_ref,
                                 divergence_window=cfg.divergence_window, max_attempts=cap,
                                 solve_tol=self.solve_tol)
        ctx.newton_control_attach(ctl)
        try:
            k = 0               # next attempt to enqueue
            pending = deque()   # attempts enqueued and not yet read
            rebuilt_at = -1     # attempt whose pattern was last rebuilt
            while True:
                if k < cap:
                    if graph is not None:
                        graph.replay()
                    else:
                        attempt()
                    post(k % 2)
                    pending.append(k)
                    k += 1
                    if graph is None and graph_ok:
                        graph = self._capture_controlled(attempt)   # recorded, not run
                        graph_ok = graph is not None
                        self.graph_attempts_captured += graph is not None
                    if len(pending) < 2 and k < cap:
                        continue   # keep one attempt queued ahead of the reads
                if not pending:
                    break
                p = pending.popleft()
                evs[p % 2].synchronize()
                state = int(pin[p % 2, L.NC_STATE])
                if state == L.NC_RUNNING:
                    continue
                if state == L.NC_REBUILD:
                    # the attempt's coating probe met a pair without a pattern slot: it and the
                    # attempt enqueued after it were no-ops; rebuild, then enqueue it again
                    torch.cuda.synchronize()
                    pending.clear()
                    k = int(pin[p % 2, L.NC_ATT])
                    if k == rebuilt_at:
                        raise RuntimeError("the rebuilt tangent pattern still lacks a coating slot")
                    rebuilt_at = k
                    asm.coating_rebuild()
                    ctx.newton_control_set(L.NC_RUNNING)
                    graph = None   # the captured attempt holds the old pattern's K
                    graph_ok = self.use_graph and asm.prescreen is None
                    continue
                if state != L.NC_ASSIST:
                    break
                # the attempt's device refinement missed solve_tol: the attempt enqueued
                # after it was a no-op; re-solve this one eagerly, then update
                torch.cuda.synchronize()
                pending.clear()
                k = int(pin[p % 2, L.NC_ATT])
                self.eager_resolves += 1
                ctx.newton_control_attach(None)   # the host's solve: norms read per step, LU fallback
                solve(self.max_refine, self.solve_tol)
                ctx.newton_control_attach(ctl)
                rel_v = float(rel.item())
                fl = self._read_flags()
                if fl & FLAG_SINGULAR or not rel_v <= self.solve_fail_tol:
                    # zero pivot or no accurate solution: the solve failed (LinearSolveProcess success=False)
                    ctx.newton_control_set(L.NC_STOPPED, L.NCR_SINGULAR if fl & FLAG_SINGULAR else L.NCR_SOLVE)
                    break
                ctx.newton_control_set(L.NC_RUNNING)
                update()
                post(p % 2)
                pending.append(p)
                graph = None   # a fallback to partial pivoting changes the captured solve
                graph_ok = self.use_graph and asm.prescreen is None
        finally:
            torch.cuda.synchronize()   # no enqueued attempt may outlive ctl, also on an exception
            ctx.newton_control_attach(None)
        c = ctl.cpu().numpy()
        self._read_flags()   # raises on a degenerate element, as the reference does
        att = int(c[L.NC_ATT])
        h = c[L.NC_HIST:L.NC_HIST + L.NC_HSTRIDE * att].reshape(att, L.NC_HSTRIDE)
        state = int(c[L.NC_STATE])
        out = NewtonOutcome(state == L.NC_CONVERGED, att, int(c[L.NC_NACT]), state == L.NC_DIVERGED, self._f,
                            [float(x) for x in h[:, 0]], [int(x) for x in h[:, 1]])
        self.solve_residuals.extend(np.float64(x) for x in h[:, 2] if not np.isnan(x))
        self.control_reasons.append(int(c[L.NC_REASON]))
        u[:] = ud.cpu().numpy()
        return out

    def _capture_controlled(self, attempt):
        """One graph of a whole device-controlled attempt, recorded without being
        run (the eager attempt before it has set the solve up); None if capture
        is not possible, the reason kept in ``graph_error``."""
        torch = self.torch
        try:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                attempt()
        except Exception as e:  # noqa: BLE001 - any capture failure falls back to eager attempts
            self.graph_error = repr(e)
            torch.cuda.synchronize()
            return None
        return g

    def _capture_attempt(self, solve, update):
        """G1 (assembly + pair force), G2 (solve with graph_refine fixed
        refinement steps, writing its relative residual) and G3 (line search +
        update), recorded without being run (the eager first attempt has set
        the solve up).  The host reads the residual between G2 and G3 and
        solves eagerly when the fixed refinement missed ``solve_tol``.  None if
        capture is not possible; the reason is kept in ``graph_error``."""
        torch, asm, ctx = self.torch, self.asm, self.ctx
        rel = torch.zeros(1, dtype=torch.float64, device=ctx.dev)
        try:
            g1 = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g1):
                asm.assemble()
                ctx.pairs_gather(f=self._f)
            g2 = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g2):
                solve(rel, self.graph_refine, 0.0)
            g3 = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g3):
                update()
        except Exception as e:  # noqa: BLE001 - any capture failure falls back to eager attempts
            self.graph_error = repr(e)
            torch.cuda.synchronize()
            return None
        return g1, g2, g3, rel

    def checkpoint(self):
        self.ga.checkpoint()

    def restore_checkpoint(self):
        self.ga.restore_checkpoint()

    # ---- learned prescreen hooks (optional part of the backend protocol)
    def use_prescreen(self, net, *, elem_layer=None, layer_scale: float = 0.125):
        self.asm.set_prescreen(net, elem_layer=elem_layer, layer_scale=layer_scale)

    def log_features(self, on: bool = True, **kw):
        self.asm.log_features(on, **kw)

    @property
    def feature_log(self):
        return self.asm.lp_log

    def watch_count(self):
        """(watched pairs, pairs) of the increment when the prescreen is on, else None."""
        if self.asm.n_watch is None:
            return None
        return self.asm.n_watch, self.ctx.n_pairs

    def snapshot_pairs(self):
        self.ctx.pairs_snapshot()

    def restore_pairs(self):
        self.ctx.pairs_restore()
        self.ctx.lp_off()
        self.asm.n_watch = None

    def end_call(self) -> int:
        return self.asm.end_call()
