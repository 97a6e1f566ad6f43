"""Device context: one mesh + section + pair list on one GPU.

Thin owner of an ``xk_ctx`` (include/xkamd.h).  Device memory for inputs and
outputs comes from torch tensors on ``cuda``; the library keeps its own
persistent pair state, staging and pattern buffers.  There is no CPU path:
constructing a context without a ROCm GPU raises.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import scipy.sparse as sp

from . import _lib as L


def _torch():
    import torch

    return torch


def require_gpu():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("xkep-cae_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


@dataclass
class Section:
    """[E, G, A, Iy, Iz, J, kappa_y, kappa_z] of ULCRBeamAssembler
    (xkep_cae/elements/_beam_assembler.py:36-62)."""

    E: float
    G: float
    A: float
    Iy: float
    Iz: float
    J: float
    kappa_y: float
    kappa_z: float

    def as_array(self) -> np.ndarray:
        return np.array([self.E, self.G, self.A, self.Iy, self.Iz, self.J, self.kappa_y, self.kappa_z])


def circle_section(radius: float, E: float, nu: float) -> Section:
    """Solid circle with the Cowper shear factor 6(1+nu)/(7+6nu)
    (xkep_cae/elements/_beam_section.py:59-66, :129-135)."""
    d = 2.0 * radius
    Iv = np.pi * d**4 / 64.0
    kap = 6.0 * (1.0 + nu) / (7.0 + 6.0 * nu)
    return Section(E, E / (2.0 * (1.0 + nu)), np.pi * radius**2, Iv, Iv, np.pi * d**4 / 32.0, kap, kap)


PAIR_FIELDS_F64 = ("s", "t", "gap", "p_n", "k_pen", "k_t", "q_trial_norm", "dissipation")


class DeviceContext:
    """Mesh, section, contact radii and pair list resident on the GPU."""

    def __init__(self, node_coords, connectivity, section: Section, *, radii=None, scf=None, v_ref=None,
                 precision: str = "fp64", staged: bool = False, host_pattern: bool = False):
        """``staged`` forces the staged element path; otherwise meshes made of
        element chains use the direct path (see ``direct``), whose tangent
        pattern is built on the device unless ``host_pattern``."""
        torch = require_gpu()
        self.torch = torch
        self.precision = precision
        X = np.ascontiguousarray(node_coords, dtype=np.float64)
        conn = np.ascontiguousarray(connectivity, dtype=np.int32)
        self.n_nodes = X.shape[0]
        self.n_elems = conn.shape[0]
        self.ndof = 6 * self.n_nodes
        self.X_np = X
        self.conn_np = conn
        r = np.zeros(self.n_elems) if radii is None else (
            np.full(self.n_elems, float(radii)) if np.isscalar(radii) else np.asarray(radii, float))
        self.radii_np = np.ascontiguousarray(r)
        sec = np.ascontiguousarray(section.as_array())
        vr = None if v_ref is None else np.ascontiguousarray(v_ref, dtype=np.float64)
        h = C.c_void_p()
        L.check(L.lib().xk_ctx_create(C.byref(h), self.n_nodes, self.n_elems, X.ctypes.data, conn.ctypes.data,
                                      sec.ctypes.data, float(scf) if scf else 0.0,
                                      vr.ctypes.data if vr is not None else None, self.radii_np.ctypes.data,
                                      L.PREC[precision] | (L.OPT_STAGED if staged else 0)
                                      | (L.OPT_HOST_PATTERN if host_pattern else 0)), "xk_ctx_create")
        self.h = h
        self.direct = bool(L.lib().xk_ctx_direct(h))
        self.dev = torch.device("cuda")
        self.X = torch.from_numpy(X).to(self.dev)
        self.kdtype = torch.float64 if precision == "fp64" else torch.float32
        self._pattern_epoch = 0
        self.coat_on = False
        self.coat_key = None
        self._pattern_cache = None

    # ------------------------------------------------------------ plumbing
    def close(self):
        if getattr(self, "h", None):
            L.lib().xk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return self.torch.cuda.current_stream().cuda_stream

    def to_dev(self, a, dtype=None):
        torch = self.torch
        dtype = dtype or torch.float64
        if isinstance(a, torch.Tensor):
            return a.to(device=self.dev, dtype=dtype).contiguous()
        return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=self.dev)

    def invalidate_pattern(self):
        self._pattern_epoch += 1

    # ------------------------------------------------------------ pattern
    def pattern(self):
        """(indptr, indices, nnz) of the current tangent pattern (scalar CSR)."""
        if self._pattern_cache is not None and self._pattern_cache[0] == self._pattern_epoch:
            return self._pattern_cache[1]
        nnz = C.c_int64()
        nb = C.c_int64()
        ns = C.c_int64()
        L.check(L.lib().xk_pattern_info(self.h, C.byref(nnz), C.byref(nb), C.byref(ns)), "xk_pattern_info")
        indptr = np.empty(self.ndof + 1, np.int64)
        indices = np.empty(max(nnz.value, 1), np.int32)
        L.check(L.lib().xk_pattern_export(self.h, indptr.ctypes.data, indices.ctypes.data), "xk_pattern_export")
        out = (indptr, indices[: nnz.value], nnz.value)
        self._pattern_cache = (self._pattern_epoch, out)
        return out

    def pattern_counts(self):
        """(scalar entries, 3x3 blocks, contributing pairs) of the current
        tangent pattern (builds it if needed)."""
        nnz = C.c_int64()
        nb = C.c_int64()
        ns = C.c_int64()
        L.check(L.lib().xk_pattern_info(self.h, C.byref(nnz), C.byref(nb), C.byref(ns)), "xk_pattern_info")
        return nnz.value, nb.value, ns.value

    def pattern_nnz(self) -> int:
        """Scalar entries of the current tangent pattern (builds it if needed)."""
        return self.pattern_counts()[0]

    def new_K(self):
        """Device buffer for the tangent values of the current pattern (the host
        copy of the pattern is made only when ``pattern()`` asks for it)."""
        return self.torch.empty(max(self.pattern_nnz(), 1), dtype=self.kdtype, device=self.dev)

    def to_csr(self, K_vals) -> sp.csr_matrix:
        indptr, indices, nnz = self.pattern()
        v = K_vals[:nnz].detach().to("cpu").numpy().astype(np.float64)
        return sp.csr_matrix((v, indices, indptr), shape=(self.ndof, self.ndof))

    # ------------------------------------------------------------ pairs
    @property
    def n_pairs(self) -> int:
        n = C.c_int64()
        L.check(L.lib().xk_pairs_count(self.h, C.byref(n)))
        return n.value

    def set_pairs(self, elem_a, elem_b, state: dict | None = None):
        ea = np.ascontiguousarray(elem_a, dtype=np.int32)
        eb = np.ascontiguousarray(elem_b, dtype=np.int32)
        st = state or {}
        keep = []

        def p(name, dtype, shape=None):
            a = st.get(name)
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data

        L.check(L.lib().xk_pairs_set(
            self.h, len(ea), ea.ctypes.data, eb.ctypes.data, p("s", np.float64), p("t", np.float64),
            p("gap", np.float64), p("normal", np.float64), p("tangent1", np.float64), p("p_n", np.float64),
            p("k_pen", np.float64), p("k_t", np.float64), p("q_trial_norm", np.float64),
            p("dissipation", np.float64), p("z_t", np.float64), p("status", np.int8), p("stick", np.int8)),
            "xk_pairs_set")
        self.invalidate_pattern()

    def get_pairs(self) -> dict:
        n = self.n_pairs
        out = {
            "elem_a": np.empty(n, np.int32), "elem_b": np.empty(n, np.int32),
            **{k: np.empty(n) for k in PAIR_FIELDS_F64},
            "normal": np.empty((n, 3)), "tangent1": np.empty((n, 3)), "z_t": np.empty((n, 2)),
            "status": np.empty(n, np.int8), "stick": np.empty(n, np.int8),
        }
        if n:
            d = lambda k: out[k].ctypes.data
            L.check(L.lib().xk_pairs_get(
                self.h, d("elem_a"), d("elem_b"), d("s"), d("t"), d("gap"), d("normal"), d("tangent1"),
                d("p_n"), d("k_pen"), d("k_t"), d("q_trial_norm"), d("dissipation"), d("z_t"), d("status"),
                d("stick")), "xk_pairs_get")
        out["tangent2"] = np.cross(out["normal"], out["tangent1"])
        return out

    # ------------------------------------------------------------ coating
    def set_coating(self, stiffness: float, damping: float = 0.0, mu: float = 0.0, k_t_ratio: float = 0.5,
                    core_radii=None):
        """Kelvin-Voigt coating terms on (stiffness > 0) or off; core radii per
        element (None: the contact radii)."""
        cr = None
        if core_radii is not None:
            cr = np.ascontiguousarray(np.broadcast_to(np.asarray(core_radii, float), (self.n_elems,)))
        L.check(L.lib().xk_ctx_set_coating(self.h, float(stiffness), float(damping), float(mu), float(k_t_ratio),
                                           None if cr is None else cr.ctypes.data), "xk_ctx_set_coating")
        self.coat_on = stiffness > 0.0
        self.coat_key = (float(stiffness), float(damping), float(mu), float(k_t_ratio),
                         None if cr is None else cr.tobytes())
        self._core_radii = cr
        self.invalidate_pattern()

    KEEP = object()

    def ensure_coating(self, stiffness: float, damping: float = 0.0, mu: float = 0.0, k_t_ratio: float = 0.5,
                       core_radii=KEEP):
        """``set_coating`` unless the context holds these parameters already;
        ``core_radii=KEEP`` keeps the core radii set before (the contact radii
        if none were)."""
        if core_radii is self.KEEP:
            cr = getattr(self, "_core_radii", None)
        else:
            cr = None if core_radii is None else np.ascontiguousarray(
                np.broadcast_to(np.asarray(core_radii, float), (self.n_elems,)))
        key = (float(stiffness), float(damping), float(mu), float(k_t_ratio), None if cr is None else cr.tobytes())
        if key != self.coat_key:
            self.set_coating(stiffness, damping, mu, k_t_ratio, cr)

    def coating_set_dt(self, coat_dt: float):
        L.check(L.lib().xk_coating_set_dt(self.h, float(coat_dt)), "xk_coating_set_dt")

    def coating_probe(self, u) -> int:
        """#(pairs pressing their coating at X + u without a slot of the pattern); > 0
        means the next assembly rebuilds the pattern first."""
        n = C.c_int64()
        L.check(L.lib().xk_coating_probe(self.h, L.ptr(u), C.byref(n), self.stream), "xk_coating_probe")
        if n.value > 0:
            self.invalidate_pattern()   # the host copy of the pattern is stale from here
        return n.value

    def coating_probe_control(self, u):
        """The probe on the attached Newton control record, with no host read: a
        pair without a slot moves the record to XK_NC_REBUILD (the rest of the
        attempt is a no-op); the host then calls ``coating_probe`` and rebuilds."""
        L.check(L.lib().xk_coating_probe_control(self.h, L.ptr(u), self.stream), "xk_coating_probe_control")

    def coating_commit(self):
        L.check(L.lib().xk_coating_commit(self.h, self.stream), "xk_coating_commit")

    def get_coating(self) -> dict:
        n = self.n_pairs
        out = {k: np.empty(n) for k in ("core_radius_a", "core_radius_b", "coating_compression",
                                         "coating_compression_prev", "coating_q_trial_norm",
                                         "coating_dissipation")}
        out["coating_z_t"] = np.empty((n, 2))
        out["coating_stick"] = np.empty(n, np.int8)
        if n:
            d = lambda k: out[k].ctypes.data
            L.check(L.lib().xk_pairs_get_coating(
                self.h, d("core_radius_a"), d("core_radius_b"), d("coating_compression"),
                d("coating_compression_prev"), d("coating_z_t"), d("coating_stick"), d("coating_q_trial_norm"),
                d("coating_dissipation")), "xk_pairs_get_coating")
        return out

    def set_coating_state(self, state: dict):
        keep = []

        def p(name, dtype):
            a = state.get(name)
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data

        L.check(L.lib().xk_pairs_set_coating(
            self.h, p("coating_compression", np.float64), p("coating_compression_prev", np.float64),
            p("coating_z_t", np.float64), p("coating_stick", np.int8), p("coating_q_trial_norm", np.float64),
            p("coating_dissipation", np.float64)), "xk_pairs_set_coating")
        self.invalidate_pattern()

    # ------------------------------------------------------------ kernels
    def _check_vec(self, a, n, name):
        if a is not None and a.numel() != n:
            raise ValueError(f"{name} has {a.numel()} entries, expected {n}")

    @staticmethod
    def _cap(K):
        return 0 if K is None else K.numel()

    def beam(self, u, f_int=None, K_vals=None):
        self._check_vec(u, self.ndof, "u")
        self._check_vec(f_int, self.ndof, "f_int")
        L.check(L.lib().xk_beam_assemble(self.h, L.ptr(u), L.ptr(f_int), L.ptr(K_vals), self._cap(K_vals),
                                         self.stream), "xk_beam_assemble")

    def pairs_run(self, modes, *, coords=None, u=None, u_ref=None, k_pen=0.0, smoothing_delta=0.0,
                  k_t_ratio=0.5, mu_eff=0.0, g_on=0.0, g_off=1e-6, allow_deactivation=True):
        coords = self.X if coords is None else coords
        if coords.numel() != 3 * self.n_nodes:
            raise ValueError("coords must hold n_nodes x 3 entries")
        self._check_vec(u, self.ndof, "u")
        self._check_vec(u_ref, self.ndof, "u_ref")
        L.check(L.lib().xk_pairs_run(self.h, modes, L.ptr(coords), L.ptr(u), L.ptr(u_ref), float(k_pen),
                                     float(smoothing_delta), float(k_t_ratio), float(mu_eff), float(g_on),
                                     float(g_off), int(bool(allow_deactivation)), self.stream), "xk_pairs_run")
        if (modes & L.PAIR_HYST) or (self.coat_on and modes & L.PAIR_GEOM):
            self.invalidate_pattern()

    def pairs_gather(self, f=None, K_vals=None):
        self._check_vec(f, self.ndof, "f")
        L.check(L.lib().xk_pairs_gather(self.h, L.ptr(f), L.ptr(K_vals), self._cap(K_vals), self.stream),
                "xk_pairs_gather")

    def broadphase(self, coords, *, margin=0.0, cell_size=None):
        """Raw inclusive-AABB-overlap pairs (i<j), sorted, as an (n,2) int64 array."""
        n = C.c_int64()
        L.check(L.lib().xk_broadphase(self.h, L.ptr(coords), float(margin), float(cell_size or 0.0), None, 0,
                                      C.byref(n), self.stream), "xk_broadphase")
        out = self.torch.empty(2 * max(n.value, 1), dtype=self.torch.int32, device=self.dev)
        L.check(L.lib().xk_broadphase(self.h, L.ptr(coords), float(margin), float(cell_size or 0.0),
                                      L.ptr(out), out.numel(), C.byref(n), self.stream), "xk_broadphase")
        a = out[: 2 * n.value].cpu().numpy().astype(np.int64).reshape(2, -1).T
        return np.ascontiguousarray(a)

    def detect_candidates(self, coords, *, margin=0.0, cell_size=None, elem_layer=None,
                          midpoint_prescreening=True, prescreening_margin=0.0):
        """DetectCandidatesProcess on the device; returns the candidate count.

        The midpoint cutoff margin is ``prescreening_margin`` when positive,
        else 0.5 * mean(r_a + r_b) over the shared-node-filtered candidates
        with NumPy's mean (as _manager_process.py:280-283)."""
        layer = None if elem_layer is None else self.to_dev(elem_layer, self.torch.int32)
        n1 = C.c_int64()
        extra = -1.0
        if midpoint_prescreening:
            if prescreening_margin > 0.0:
                L.check(L.lib().xk_prescreen_radii(self.h, L.ptr(coords), float(margin), float(cell_size or 0.0),
                                                   L.ptr(layer), None, 0, C.byref(n1), self.stream))
                extra = float(prescreening_margin)
            else:
                uniform = self.radii_np.size and np.all(self.radii_np == self.radii_np[0])
                if uniform:
                    L.check(L.lib().xk_prescreen_radii(self.h, L.ptr(coords), float(margin),
                                                       float(cell_size or 0.0), L.ptr(layer), None, 0,
                                                       C.byref(n1), self.stream))
                    r0 = float(self.radii_np[0])
                    extra = float(np.mean(np.full(n1.value, r0 + r0))) * 0.5 if n1.value else 0.0
                else:
                    L.check(L.lib().xk_prescreen_radii(self.h, L.ptr(coords), float(margin),
                                                       float(cell_size or 0.0), L.ptr(layer), None, 0,
                                                       C.byref(n1), self.stream))
                    rs = np.empty(max(n1.value, 1))
                    L.check(L.lib().xk_prescreen_radii(self.h, L.ptr(coords), float(margin),
                                                       float(cell_size or 0.0), L.ptr(layer), rs.ctypes.data,
                                                       rs.size, C.byref(n1), self.stream))
                    extra = float(np.mean(rs[: n1.value])) * 0.5 if n1.value else 0.0
        nr = C.c_int64()
        nc = C.c_int64()
        L.check(L.lib().xk_detect_candidates(self.h, L.ptr(coords), float(margin), float(cell_size or 0.0),
                                             L.ptr(layer), extra, C.byref(nr), C.byref(nc), self.stream),
                "xk_detect_candidates")
        self.invalidate_pattern()
        self._n_cand = nc.value
        return nc.value

    def strategy_detect(self, coords, *, shared_node_filter: bool = True, margin: float = 0.0,
                        cell_size=None) -> int:
        """Replace the pair list by the broadphase pairs of ``coords`` with a
        fresh state (``_detect_candidates`` before its narrowphase); returns
        the number of pairs."""
        n = C.c_int64()
        L.check(L.lib().xk_strategy_detect(self.h, L.ptr(coords), float(margin), float(cell_size or 0.0),
                                           int(bool(shared_node_filter)), C.byref(n), self.stream),
                "xk_strategy_detect")
        self.invalidate_pattern()
        return n.value

    def constraint_jacobian(self, ndof_total: int, ndof_per_node: int = 6):
        """(G, active_indices): the constraint Jacobian of the pair list as a
        SciPy CSR matrix (rows = pairs with status != INACTIVE, in pair order)."""
        lib = L.lib()
        rows, nnz = C.c_int64(), C.c_int64()
        args = (self.h, int(ndof_per_node), int(ndof_total), C.byref(rows), C.byref(nnz))
        L.check(lib.xk_pairs_constraint_jacobian(*args, None, None, None, None, 0, 0, self.stream),
                "xk_pairs_constraint_jacobian")
        if rows.value == 0:
            return sp.csr_matrix((0, int(ndof_total))), []
        t = self.torch
        act = t.empty(rows.value, dtype=t.int32, device=self.dev)
        indptr = t.empty(rows.value + 1, dtype=t.int64, device=self.dev)
        idx = t.empty(nnz.value, dtype=t.int32, device=self.dev)
        data = t.empty(nnz.value, dtype=t.float64, device=self.dev)
        L.check(lib.xk_pairs_constraint_jacobian(*args, L.ptr(act), L.ptr(indptr), L.ptr(idx), L.ptr(data),
                                                 act.numel(), idx.numel(), self.stream),
                "xk_pairs_constraint_jacobian")
        G = sp.csr_matrix((data.cpu().numpy(), idx.cpu().numpy(), indptr.cpu().numpy()),
                          shape=(rows.value, int(ndof_total)))
        return G, act.cpu().numpy().astype(np.int64).tolist()

    def last_candidates(self) -> np.ndarray:
        """Filtered candidate pairs (i<j) of the last ``detect_candidates``, sorted."""
        n = getattr(self, "_n_cand", 0)
        buf = np.empty(2 * max(n, 1), np.int32)
        L.check(L.lib().xk_last_candidates(self.h, buf.ctypes.data, buf.size), "xk_last_candidates")
        return np.ascontiguousarray(buf[: 2 * n].reshape(2, -1).T.astype(np.int64))

    def assemble_iteration(self, u, u_ref, f_ext, fixed_mask, R, K_vals=None, dyn: "L.XkDyn | None" = None, *,
                           k_pen, smoothing_delta, k_t_ratio, mu_eff, res_norm2=None, n_active=None):
        for a, nm in ((u, "u"), (u_ref, "u_ref"), (f_ext, "f_ext"), (fixed_mask, "fixed_mask"), (R, "R")):
            self._check_vec(a, self.ndof, nm)
        L.check(L.lib().xk_assemble_iteration(
            self.h, L.ptr(u), L.ptr(u_ref), L.ptr(f_ext), L.ptr(fixed_mask),
            C.byref(dyn) if dyn is not None else None, float(k_pen), float(smoothing_delta), float(k_t_ratio),
            float(mu_eff), L.ptr(R), L.ptr(K_vals), self._cap(K_vals), L.ptr(res_norm2), L.ptr(n_active),
            self.stream), "xk_assemble_iteration")

    def solve_lu(self, kind: int | None = None) -> tuple[bool, int]:
        """Capacitance LU of the device solve: kind 0 = without row
        interchanges, partial pivoting after a failure (default); 1 = partial
        pivoting always; None = unchanged.  Returns (the current pattern's
        solves pivot, solves that fell back so far)."""
        piv, fb = C.c_int32(), C.c_int64()
        L.check(L.lib().xk_solve_set_lu(self.h, -1 if kind is None else int(kind), C.byref(piv), C.byref(fb)),
                "xk_solve_set_lu")
        return bool(piv.value), int(fb.value)

    def linear_solve(self, u, fixed_mask, R, du, dyn: "L.XkDyn | None" = None, *, max_refine: int = 16,
                     tol: float = 1e-11, res_rel=None) -> int:
        """du = -(D K D + I - D)^-1 D R on the device (LinearSolveProcess,
        _newton_steps.py:338-366), after ``assemble_iteration`` with a tangent
        at the same u, fixed_mask and dyn; refined until the relative residual
        is <= tol or stops decreasing (tol <= 0: exactly max_refine steps, no
        host read).
        res_rel: optional device float64 scalar receiving the final relative
        residual.  Returns the refinement steps taken (-1 when a device Newton
        control record is attached and tol > 0: the refinement then stops on
        the device, without a host read)."""
        for a, nm in ((u, "u"), (fixed_mask, "fixed_mask"), (R, "R"), (du, "du")):
            self._check_vec(a, self.ndof, nm)
        steps = C.c_int32()
        L.check(L.lib().xk_linear_solve(
            self.h, L.ptr(u), L.ptr(fixed_mask), C.byref(dyn) if dyn is not None else None, L.ptr(R),
            L.ptr(du), int(max_refine), float(tol), L.ptr(res_rel), C.byref(steps), self.stream),
            "xk_linear_solve")
        return steps.value

    def newton_update(self, u, du, R, f_contact, f_ext, fixed_mask, res_norm2, status, *,
                      line_search_steps: int = 5, diverge_factor: float = 3.0, du_norm_cap: float = 0.0):
        """Line search + update + norms of a Newton attempt on the device
        (LineSearchUpdateProcess with _ncp_line_search, _newton_steps.py:396-434,
        _utils.py:31-61); ``status`` (device float64, >= 24) receives alpha,
        ||du||, ||u||, |du.R| and the accepted candidate.  No host sync."""
        for a, nm in ((u, "u"), (du, "du"), (R, "R"), (f_contact, "f_contact"), (f_ext, "f_ext"),
                      (fixed_mask, "fixed_mask")):
            self._check_vec(a, self.ndof, nm)
        if status.numel() < 24:
            raise ValueError("status needs 24 entries")
        L.check(L.lib().xk_newton_update(
            self.h, L.ptr(u), L.ptr(du), L.ptr(R), L.ptr(f_contact), L.ptr(f_ext), L.ptr(fixed_mask),
            L.ptr(res_norm2), int(line_search_steps), float(diverge_factor), float(du_norm_cap), L.ptr(status),
            self.stream), "xk_newton_update")

    # ---- device Newton control (the driver's tests on the device)
    def newton_control_begin(self, ctl, *, hist_cap: int, tol_force: float, tol_disp: float,
                             f_ext_ref_norm: float, dynamic_ref: bool, divergence_window: int, max_attempts: int,
                             solve_tol: float):
        """Reset the control record ``ctl`` (device float64, NC_HIST +
        NC_HSTRIDE * hist_cap) for a new Newton call and clear the device
        status word; stream-ordered."""
        if ctl.dtype != self.torch.float64 or not ctl.is_cuda or ctl.numel() < L.NC_HIST + L.NC_HSTRIDE * hist_cap:
            raise ValueError("ctl must be a device float64 tensor of NC_HIST + NC_HSTRIDE * hist_cap entries")
        L.check(L.lib().xk_newton_control_begin(
            self.h, L.ptr(ctl), int(hist_cap), float(tol_force), float(tol_disp), float(f_ext_ref_norm),
            int(bool(dynamic_ref)), int(divergence_window), int(max_attempts), float(solve_tol), self.stream),
            "xk_newton_control_begin")

    def newton_control_attach(self, ctl):
        """Attach ``ctl`` (None detaches): the attempt kernels skip once its call has ended."""
        L.check(L.lib().xk_newton_control_attach(self.h, L.ptr(ctl) if ctl is not None else None),
                "xk_newton_control_attach")

    def newton_control_force(self, res_norm2, n_active):
        """Force and divergence tests of the attempt, on the attached record."""
        L.check(L.lib().xk_newton_control_force(self.h, L.ptr(res_norm2), L.ptr(n_active), self.stream),
                "xk_newton_control_force")

    def newton_control_set(self, state: int, reason: int = 0):
        """Host decision after an eager re-solve or a pattern rebuild (XK_NC_RUNNING
        after XK_NC_ASSIST accepts the attempt's solve)."""
        L.check(L.lib().xk_newton_control_set(self.h, int(state), int(reason), self.stream),
                "xk_newton_control_set")

    SOLVE_BUFFERS = ("Kb", "Ab", "Lf", "Sinv", "slist", "slot_nodes", "slot_G", "slot_c", "pos", "cslot", "cn",
                     "cnode", "ch_m", "ch_cn0", "ch_yoff", "ch_coff", "Y", "C", "S", "ipiv", "info", "W", "b", "x",
                     "r", "t", "red", "perm", "W2", "piv32")

    def solve_buffer(self, name: str, dtype) -> np.ndarray:
        """Host copy of one of the linear solve's device buffers (tests)."""
        k = self.SOLVE_BUFFERS.index(name)
        n = C.c_int64()
        L.check(L.lib().xk_solve_buffer(self.h, k, None, 0, C.byref(n)), "xk_solve_buffer")
        out = np.empty(n.value // np.dtype(dtype).itemsize, dtype)
        L.check(L.lib().xk_solve_buffer(self.h, k, out.ctypes.data, out.nbytes, None), "xk_solve_buffer")
        return out

    def profile_begin(self, max_calls: int):
        L.check(L.lib().xk_profile_begin(self.h, int(max_calls)), "xk_profile_begin")

    def profile_end(self):
        """(stage_ms[4] summed over calls, n_calls): pair kernel, element kernel,
        gather kernel, finish kernel."""
        ms = np.zeros(4)
        n = C.c_int()
        L.check(L.lib().xk_profile_end(self.h, ms.ctypes.data, C.byref(n)), "xk_profile_end")
        return ms, n.value

    def set_friction(self, mu_elem=None):
        """Per-element friction coefficients (pair: mu_eff x mu_elem[elem_a]); None: off."""
        if mu_elem is None:
            L.check(L.lib().xk_ctx_set_friction(self.h, None), "xk_ctx_set_friction")
            return
        m = np.ascontiguousarray(np.asarray(mu_elem, float))
        if m.shape != (self.n_elems,):
            raise ValueError("one coefficient per element")
        L.check(L.lib().xk_ctx_set_friction(self.h, m.ctypes.data), "xk_ctx_set_friction")

    # ------------------------------------------------------------ launch shapes
    def get_tuning(self) -> dict:
        t = L.XkTuning()
        L.check(L.lib().xk_ctx_get_tuning(self.h, C.byref(t)), "xk_ctx_get_tuning")
        return {"cell_factor": t.cell_factor, "pair_block": t.pair_block, "gather_block": t.gather_block}

    def set_tuning(self, *, cell_factor=None, pair_block=None, gather_block=None):
        cur = self.get_tuning()
        t = L.XkTuning(cell_factor=float(cell_factor if cell_factor is not None else cur["cell_factor"]),
                       pair_block=int(pair_block if pair_block is not None else cur["pair_block"]),
                       gather_block=int(gather_block if gather_block is not None else cur["gather_block"]))
        L.check(L.lib().xk_ctx_set_tuning(self.h, C.byref(t)), "xk_ctx_set_tuning")

    # ------------------------------------------------------------ learned prescreen
    def lp_set(self, params, threshold: float):
        p = np.ascontiguousarray(np.asarray(params, np.float32))
        L.check(L.lib().xk_lp_set_params(self.h, p.ctypes.data, int(p.size), float(threshold)), "xk_lp_set_params")

    def lp_features(self, u=None, elem_layer=None, layer_scale: float = 0.125, log: bool = False):
        """Features of every pair at X + u (after the increment's geometry update);
        returns them as an (n_pairs, 8) float32 device tensor when ``log``."""
        torch = self.torch
        u = None if u is None else self.to_dev(u)
        layer = None if elem_layer is None else self.to_dev(elem_layer, torch.int32)
        out = torch.empty((max(self.n_pairs, 1), 8), dtype=torch.float32, device=self.dev) if log else None
        L.check(L.lib().xk_lp_features(self.h, L.ptr(self.X), L.ptr(u), L.ptr(layer), float(layer_scale),
                                       L.ptr(out), self.stream), "xk_lp_features")
        return None if out is None else out[: self.n_pairs]

    def lp_select(self) -> int:
        n = C.c_int64()
        L.check(L.lib().xk_lp_select(self.h, C.byref(n), self.stream), "xk_lp_select")
        return n.value

    def lp_verify(self, u, g_on: float = 0.0) -> int:
        n = C.c_int64()
        L.check(L.lib().xk_lp_verify(self.h, L.ptr(self.X), L.ptr(self.to_dev(u)), float(g_on), C.byref(n),
                                    self.stream), "xk_lp_verify")
        return n.value

    def lp_off(self):
        L.check(L.lib().xk_lp_off(self.h), "xk_lp_off")

    def pairs_snapshot(self):
        L.check(L.lib().xk_pairs_snapshot(self.h), "xk_pairs_snapshot")

    def pairs_restore(self):
        L.check(L.lib().xk_pairs_restore(self.h), "xk_pairs_restore")

    def status_flags(self) -> int:
        f = C.c_int()
        L.check(L.lib().xk_status(self.h, C.byref(f)))
        return f.value
