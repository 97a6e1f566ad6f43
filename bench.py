"""Benchmark of the per-Newton-iteration assembly on one MI355X.

Workload: the 91-wire strand (layers 1+6+12+18+24+30, r = 0.4 mm, L = 80 mm,
150 elements per wire, lay angle 12 deg), Generalized-alpha (rho_inf = 0.8)
dynamic step, Huber contact k = 1e4 E r, Coulomb mu = 0.3, seeded
perturbation u ~ N(0, (1e-3 r)^2), candidates between layers (same-layer
exclusion by layer index, so the rest state has no initial penetration).  One step = one complete fused assembly
(narrowphase over every candidate pair, Huber, friction, element f_int and K,
R and K composition with inertia terms, ||R||^2, #(p_n > 0)) replayed from a
HIP graph, inputs resident in HBM.

value: the FP32-mixed form over a batch of B independent 91-wire instances
(different seeds) assembled by the same launches, (sum N_elem + N_pair) per
second.  Beside it: the single-instance FP64 plain and FP32-mixed forms.

Prints ONE JSON line (one GPU).
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(case, states, precision, torch, batch=None, mu_elem=None, coating=None, layer_filter=True):
    """Context + assembly for one instance, a concatenated batch of instances
    of one case, or a prepared uneven batch (per-element friction); ``coating``
    (config fields, core radius fraction, coat_dt) switches the Kelvin-Voigt
    coating terms on; ``layer_filter`` False keeps every candidate pair of the
    broadphase (no same-layer exclusion, no layer map)."""
    wl = importlib.import_module("xkep-cae_amd.workloads")
    C = importlib.import_module("xkep-cae_amd.context")
    A = importlib.import_module("xkep-cae_amd.assembly")
    T = importlib.import_module("xkep-cae_amd.time_integration")
    E = importlib.import_module("xkep-cae_amd.elements")
    K = importlib.import_module("xkep-cae_amd.contact")
    if batch is not None:
        b = batch
        X, conn, radii, u, f_ext, fixed = b.node_coords, b.connectivity, b.radii, b.u, b.f_ext, b.fixed_dofs
        layer = b.elem_layer
    elif len(states) == 1:
        st = states[0]
        X, conn, radii, u, f_ext, fixed = (st.mesh.node_coords, st.mesh.connectivity, st.mesh.radii, st.u,
                                           st.f_ext, st.fixed_dofs)
        layer = st.elem_layer
    else:
        b = wl.concat_states(states)
        X, conn, radii, u, f_ext, fixed = b.node_coords, b.connectivity, b.radii, b.u, b.f_ext, b.fixed_dofs
        layer = b.elem_layer
    mu = case.mu if mu_elem is None else 1.0
    coat_kw, core, coat_dt = ({}, None, 0.0) if coating is None else coating
    cfg = (K.ContactConfig(mu=mu, exclude_same_layer=True, elem_layer_map=wl.layer_map(layer), **coat_kw)
           if layer_filter else K.ContactConfig(mu=mu, exclude_same_layer=False, **coat_kw))
    sec = C.circle_section(case.wire_radius, case.E, case.nu)
    ctx = C.DeviceContext(X, conn, sec, radii=radii, precision=precision)
    if mu_elem is not None:
        ctx.set_friction(mu_elem)
    Md = E.ULCRBeamAssembler(X, conn, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J, sec.kappa_y,
                             sec.kappa_z).lumped_mass_diag(case.rho)
    dt = case.dt()
    ga = T.GeneralizedAlphaProcess(Md, rho_inf=case.rho_inf)
    ga.predict_device(torch.zeros(ctx.ndof, dtype=torch.float64, device=ctx.dev), dt)
    # the assembly as the Newton driver runs it each attempt: the attempt's acceleration and
    # velocity are not stored (the driver's correct() at the increment's end sets them)
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=mu,
                           f_ext=f_ext, fixed_dofs=fixed, config=cfg, time=ga, dt=dt,
                           core_radii=None if core is None else core * radii, ga_state=False)
    if coating is not None:
        asm.set_coat_dt(coat_dt)
    asm.bench_layer = layer
    asm.begin_increment(np.zeros(ctx.ndof))   # set-up broadphase on the reference configuration
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    asm.begin_increment(u)                     # increment: broadphase + hysteresis + pattern
    torch.cuda.synchronize()
    t_inc = time.perf_counter() - t0
    return ctx, asm, t_inc


def _timed(torch, f, reps=3):
    """Median wall time (ms) of f() over reps calls, synchronised on both sides."""
    ts, r = [], None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = f()
        torch.cuda.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(ts)), r


def increment_breakdown(asm, ctx, torch, reps=3):
    """A steady-state increment (buffers already sized by the first one) at the
    same displacement: DetectCandidates (broadphase, filters, pair-list merge),
    UpdateGeometry with hysteresis, and the tangent pattern of the increment,
    each timed alone, then begin_increment as a whole."""
    L = importlib.import_module("xkep-cae_amd._lib")
    x = asm.deformed_coords()
    cfg = asm.cfg
    t_det, n_cand = _timed(torch, lambda: ctx.detect_candidates(
        x, elem_layer=asm._layer, midpoint_prescreening=cfg.midpoint_prescreening,
        prescreening_margin=cfg.prescreening_margin), reps)
    t_geo, _ = _timed(torch, lambda: ctx.pairs_run(L.PAIR_GEOM | L.PAIR_HYST, coords=x, g_on=cfg.g_on,
                                                   g_off=cfg.g_off), reps)
    t_pat, _ = _timed(torch, lambda: ctx.new_K(), reps)
    t_all, _ = _timed(torch, lambda: asm.begin_increment(), reps)
    return {"steady_ms": round(t_all, 3), "detect_ms": round(t_det, 3), "geometry_ms": round(t_geo, 3),
            "pattern_ms": round(t_pat, 3), "candidates": int(n_cand), "pairs": int(ctx.n_pairs),
            "candidates_per_s": round(n_cand / (t_det * 1e-3), 1)}


def time_graph(asm, torch, steps, warmup):
    asm.capture()
    for _ in range(warmup):
        asm.replay()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        asm.replay()
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall / steps * 1e3


def profile_stages(asm, ctx, torch, calls):
    ctx.profile_begin(calls)
    for _ in range(calls):
        asm.assemble()
    torch.cuda.synchronize()
    ms, n = ctx.profile_end()
    return ms / max(n, 1)


STAGES = ["pairs", "elements", "gather", "finish"]


def stage_kernels(tuning) -> dict:
    """Kernel names of each stage in the FP32-mixed form, with the tuned
    workgroup sizes as they appear in the rocprofv3 summaries."""
    return {"pairs": f"xk::pair_kernel<double, {tuning['pair_block']}, false>",
            "elements": "xk::beam_direct_kernel<float>",
            "gather": f"xk::gather_kernel<float, float, double, {tuning['gather_block']}>",
            "finish": "xk::finish_kernel"}
TRAFFIC_TABLE = os.path.join(HERE, "profiles", "assembly_kernels_91wire_hbm_traffic.md")


def beam_nnz(conn) -> int:
    """Entries of the beam part of the tangent pattern: distinct 3x3 blocks over
    the DOF groups (2n, 2n + 1) of every element's two nodes, times 9."""
    conn = np.asarray(conn, np.int64)
    g = np.stack([2 * conn[:, 0], 2 * conn[:, 0] + 1, 2 * conn[:, 1], 2 * conn[:, 1] + 1], axis=1)
    keys = (g[:, :, None] << 32) | g[:, None, :]
    return 9 * int(np.unique(keys.ravel()).size)


def stage_bytes(ctx, asm, n_pair, precision, nnz_beam):
    """Algorithmic bytes per launch of each kernel (compulsory inputs and outputs,
    nothing counted for staging records or re-reads):
    pairs     N_pair x 216 (read ids, previous frame, history, status 84 B;
              write s, t, gap, frame, p_n, history, status 132 B);
    elements  per node X 24 + u 48 read, f_int 48 written; per element conn 8;
              the beam values of K written (nnz_beam x s_K);
    gather    per node u, f_ext, f_int, M, u_pred, acc_old read and R written
              (7 x 48 B; with the attempt's GA state stored also v_pred read and
              acc, vel written, 10 x 48 B); the contact-only values of K written;
    finish    one partial sum per gather workgroup and one count per pair workgroup."""
    nn, ne = ctx.n_nodes, ctx.n_elems
    s_k = 8 if precision == "fp64" else 4
    nnz = asm.K.numel()
    return {
        "pairs": n_pair * 216,
        "elements": nn * (24 + 48 + 48) + ne * 8 + nnz_beam * s_k,
        "gather": nn * (10 if asm.ga_state else 7) * 48 + (nnz - nnz_beam) * s_k,
        "finish": (6 * nn + 255) // 256 * 8 + (n_pair + 255) // 256 * 4,
    }


FLOPS_TABLE = os.path.join(HERE, "profiles", "assembly_kernels_91wire_flops.md")


def kernel_mflop(kernel: str):
    """Floating-point work per dispatch (MFLOP, FP64 + FP32) of the kernel whose
    name starts with `kernel`, at the largest grid of the committed rocprofv3
    instruction-count summary, or None."""
    if not os.path.exists(FLOPS_TABLE):
        return None
    best = None
    for line in open(FLOPS_TABLE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 14 and cells[0].startswith(kernel) and cells[1].isdigit():
            grid, mf = int(cells[1]), float(cells[13])
            if best is None or grid > best[0]:
                best = (grid, mf)
    return None if best is None else best[1]


def traffic_bytes(kernel: str):
    """HBM bytes per dispatch of the kernel named `kernel`... at the largest grid in the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a copy kernel),
    or None if the summary has no row for it."""
    if not os.path.exists(TRAFFIC_TABLE):
        return None
    best = None
    for line in open(TRAFFIC_TABLE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 8 and cells[0].startswith(kernel) and cells[1].isdigit():
            grid, mb = int(cells[1]), float(cells[7])
            if best is None or grid > best[0]:
                best = (grid, mb)
    return None if best is None else int(best[1] * 1e6)


def host_cpu() -> str:
    """CPU model, logical CPUs, CPUs this process may run on, and the thread environment."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = os.cpu_count()
    env = ", ".join(f"{k}={os.environ.get(k, 'unset')}" for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS",
                                                                   "MKL_NUM_THREADS"))
    return f"{model}, {os.cpu_count()} logical CPUs ({allowed} usable), {env}"


def cpu_baseline(case, st, budget_s=15.0):
    """The FP64 NumPy oracle's full assembly on this host, one thread."""
    from threadpoolctl import threadpool_limits

    from oracle import assembly as oasm
    from oracle import beam as obeam
    from oracle import contact as ocon

    wl = importlib.import_module("xkep-cae_amd.workloads")
    m = st.mesh
    osec = obeam.circle_section(case.wire_radius, case.E, case.nu)
    cfg = ocon.ContactConfig(mu=case.mu, exclude_same_layer=True, elem_layer_map=wl.layer_map(st.elem_layer))
    prob = oasm.Problem(m.node_coords, m.connectivity, osec, case.k_pen, case.smoothing_delta, 0.5, case.mu, cfg)
    with threadpool_limits(1):
        op, _ = ocon.detect_candidates(ocon.Pairs(), m.node_coords, m.connectivity, m.radii, cfg)
        x = m.node_coords + st.u.reshape(-1, 6)[:, :3]
        op, _ = ocon.detect_candidates(op, x, m.connectivity, m.radii, cfg)
        op = ocon.update_geometry(op, x, cfg)
        ga = oasm.GeneralizedAlpha(obeam.lumped_mass_diag(m.node_coords, m.connectivity, case.rho, osec),
                                   case.rho_inf)
        dt = case.dt()
        ga.predict(np.zeros_like(st.u), dt)
        times = []
        t_start = time.perf_counter()
        while True:
            t0 = time.perf_counter()
            oasm.assemble_iteration(prob, op, st.u, st.u_ref, st.f_ext, st.fixed_dofs, ga, dt)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_start > budget_s and len(times) >= 3:
                break
    t = float(np.median(times[1:])) if len(times) > 2 else float(np.median(times))
    return (m.n_elems + len(op)) / t, len(times), t


def run(label, case, states, prec, torch, steps, warmup, tune=False, batch=None, mu_elem=None):
    log(f"[bench] {label}: setting up precision={prec}")
    ctx, asm, t_inc = setup(case, states, prec, torch, batch=batch, mu_elem=mu_elem)
    tuning = None
    if tune:
        T = importlib.import_module("xkep-cae_amd.autotune")
        tuning = T.autotune(asm, use_cache=False, calls=5)
        log(f"[bench] {label}: autotuned {{'cell_factor': {tuning['cell_factor']}, 'pair_block': "
            f"{tuning['pair_block']}, 'gather_block': {tuning['gather_block']}}}")
    n_pair = ctx.n_pairs
    asm.assemble()
    torch.cuda.synchronize()
    stages = profile_stages(asm, ctx, torch, 20)
    ms, ms_wall = time_graph(asm, torch, steps, warmup)
    n_eval = ctx.n_elems + n_pair
    inc = increment_breakdown(asm, ctx, torch)
    log(f"[bench] {label}: elems={ctx.n_elems} pairs={n_pair} nnzK={asm.K.numel()} first incr {t_inc * 1e3:.1f} ms, "
        f"steady {inc['steady_ms']:.2f} ms (detect {inc['detect_ms']:.2f}) | "
        f"{ms * 1e3:.1f} us/step (graph) wall {ms_wall * 1e3:.1f} | stages us {np.round(stages * 1e3, 1).tolist()} | "
        f"{n_eval / (ms * 1e-3):.3e} evals/s")
    out = dict(ctx=ctx, asm=asm, n_pair=n_pair, n_elem=ctx.n_elems, n_eval=n_eval, tuning=tuning, ms=ms, ms_wall=ms_wall, stages=stages,
               t_inc=t_inc, inc=inc, value=n_eval / (ms * 1e-3))
    return out


def coated_assembly(case, states, torch, calls=50):
    """The Kelvin-Voigt coating terms on (stiffness 0.1 k_pen, dashpot
    0.5 k_coat coat_dt, coating mu 0.2, core radii 0.9 of the contact radii)
    on the headline batch, FP32-mixed.  A coated assembly probes the
    compression first (a launch and a host read) and is not graph-captured,
    so both forms are timed eagerly, each call synchronised."""
    out = {}
    k = 0.1 * case.k_pen
    coat_dt = 0.05
    coat = (dict(coating_stiffness=k, coating_damping=0.5 * k * coat_dt, coating_mu=0.2, coating_k_t_ratio=0.5),
            0.9, coat_dt)
    for label, c in (("plain", None), ("coated", coat)):
        ctx, asm, _ = setup(case, states, "mixed", torch, coating=c)
        asm.assemble()
        torch.cuda.synchronize()
        t = []
        for _ in range(calls):
            t0 = time.perf_counter()
            asm.assemble()
            torch.cuda.synchronize()
            t.append(time.perf_counter() - t0)
        rec = {"us_per_assembly_eager": round(1e6 * float(np.median(t)), 1)}
        if c is not None:
            cc = ctx.get_coating()["coating_compression"]
            rec.update(compressed_pairs=int((cc > 0).sum()), pattern_rebuilds=int(asm.n_coat_rebuilds))
        out[label] = rec
        log(f"[bench] coated assembly x{len(states)} {label}: {rec}")
        del ctx, asm
        torch.cuda.empty_cache()
    out["coating_overhead"] = round(out["coated"]["us_per_assembly_eager"] / out["plain"]["us_per_assembly_eager"], 3)
    out["workload"] = (f"{len(states)} x 91-wire batch, FP32-mixed, Kelvin-Voigt coating (k 0.1 k_pen, "
                       "dashpot, coating mu 0.2, core 0.9 r) against the same batch without it, eager calls")
    return out


def mixed_vs_fp64(case, states, torch):
    """The FP32-mixed assembly against the FP64 plain one on the same batch and
    state (one fresh assembly each, the same pattern): relative residual and
    tangent errors, the pairs' statuses (the active set) and the active count."""
    res, geo = {}, {}
    for prec in ("fp64", "mixed", "fp32"):
        ctx, asm, _ = setup(case, states, prec, torch)
        R, K, _, nact = asm.assemble()
        torch.cuda.synchronize()
        res[prec] = (R.cpu().numpy().astype(np.float64), K.cpu().numpy().astype(np.float64),
                     ctx.get_pairs()["status"].copy(), int(nact.item()))
        pr = ctx.get_pairs()
        geo[prec] = (pr["gap"].copy(), pr["normal"].copy(), pr["tangent1"].copy())
        del ctx, asm, R, K
        torch.cuda.empty_cache()
    (R64, K64, s64, n64), (R32, K32, s32, n32) = res["fp64"], res["mixed"]
    Rf, Kf, sf, nf = res["fp32"]
    rad = float(np.max(states[0].mesh.radii))
    fp32 = {"R_rel_err": float(np.linalg.norm(Rf - R64) / np.linalg.norm(R64)),
            "K_rel_err_frobenius": float(np.linalg.norm(Kf - K64) / np.linalg.norm(K64)),
            "active_set_equal": bool(np.array_equal(s64, sf)), "n_active_fp32": nf,
            "gap_max_abs_err_over_r": float(np.max(np.abs(geo["fp32"][0] - geo["fp64"][0]))) / rad,
            "normal_max_abs_err": float(np.max(np.abs(geo["fp32"][1] - geo["fp64"][1]))),
            "tangent1_max_abs_err": float(np.max(np.abs(geo["fp32"][2] - geo["fp64"][2])))}
    out = {"R_rel_err": float(np.linalg.norm(R32 - R64) / np.linalg.norm(R64)),
           "K_rel_err_frobenius": float(np.linalg.norm(K32 - K64) / np.linalg.norm(K64)),
           "K_max_abs_err_over_max_K": float(np.max(np.abs(K32 - K64)) / np.abs(K64).max()),
           "active_set_equal": bool(np.array_equal(s64, s32)), "n_active_fp64": n64, "n_active_mixed": n32,
           "fp32_pairs": fp32,
           "note": "FP32-mixed vs FP64 plain, one assembly of the headline batch at the same state; "
                   "the FP64 form equals the FP64 oracle within 1e-10 (tests)"}
    log(f"[bench] mixed vs fp64: {out}")
    return out


def newton_attempt(case, torch, wl, reps=5):
    """One Newton attempt at 91 wires: the fused assembly plus the device
    linear solve (xk_linear_solve, refined to 1e-11) against the same assembly
    plus SciPy SuperLU on the host (the reference's LinearSolveProcess), FP64
    and FP32-mixed tangents."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla

    st = wl.make_state(case)
    out = {}
    for prec in ("fp64", "mixed"):
        ctx, asm, _ = setup(case, [st], prec, torch)
        du = torch.empty(ctx.ndof, dtype=torch.float64, device=ctx.dev)
        ta, ts, steps = [], [], []
        for rep in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            R, K, _, _ = asm.assemble()
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            steps.append(ctx.linear_solve(asm.u, asm.fixed_mask, R, du, asm._dyn, max_refine=16, tol=1e-11))
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            if rep:   # the first call also sets up the solve for the pattern
                ta.append(t1 - t0)
                ts.append(t2 - t1)
        rec = {"assembly_ms": round(1e3 * float(np.median(ta)), 3), "device_solve_ms": round(1e3 * float(np.median(ts)), 2),
               "refinement_steps": int(np.median(steps)), "contributing_pairs": int(ctx.pattern_counts()[2])}
        if prec == "fp64":
            Kc = asm.K_csr()
            n = Kc.shape[0]
            free = np.ones(n)
            free[st.fixed_dofs] = 0.0
            D = sp.diags(free)
            A = (D @ Kc @ D + sp.diags(1.0 - free)).tocsc()
            b = -R.cpu().numpy()
            b[st.fixed_dofs] = 0.0
            t0 = time.perf_counter()
            sla.spsolve(A, b)
            rec["host_superlu_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
            rec["solve_speedup_vs_superlu"] = round(rec["host_superlu_ms"] / rec["device_solve_ms"], 1)
        log(f"[bench] Newton attempt 91 wires {prec}: {rec}")
        out[prec] = rec
        del ctx, asm
        torch.cuda.empty_cache()
    out["workload"] = ("91-wire strand, one Newton attempt: fused assembly + device linear solve "
                       "(chain block LU + Woodbury over the pairs' tangent blocks) vs SciPy SuperLU on the host")
    return out


def newton_loop(torch, wl, force=0.1):
    """A whole dynamic solve of the 91-wire strand geometry (radial squeeze of
    the outer wires, workloads.squeeze_case, Coulomb mu 0.3, FP64): Newton
    attempts with host SuperLU, on the device (convergence and divergence
    tests on the device, attempts enqueued back to back), the same with the
    attempts after each increment's first replayed from HIP graphs, and on the
    device with the tests on the host (a host read after each attempt's
    stages).  The trajectories (attempts and active counts per call) must
    agree."""
    TP = importlib.import_module("xkep-cae_amd.train_prescreen")
    sq = wl.squeeze_case(force, n_wires=91, n_elems=150, length=80.0, radius=0.4, lay_angle_deg=12.0, mu=0.3)
    out, traj = {}, {}
    for mode in ("device_graph", "device", "device_hostctl", "host_superlu"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, b = TP.run_case(sq, precision="fp64", device_solve=mode != "host_superlu",
                             use_graph=mode == "device_graph", device_control=mode != "device_hostctl")
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        att = sum(c.n_attempts for c in res.calls)
        traj[mode] = [(c.n_attempts, c.n_active) for c in res.calls]
        out[mode] = {"seconds": round(t, 3), "attempts": att, "ms_per_attempt": round(1e3 * t / max(att, 1), 2)}
        if mode == "device_graph":
            out[mode]["graphs_captured"] = b.graph_attempts_captured
        log(f"[bench] Newton loop 91 wires {mode}: {out[mode]}")
    out["trajectories_equal"] = (traj["device_graph"] == traj["device"] == traj["device_hostctl"]
                                 == traj["host_superlu"])
    out["calls"] = traj["device"]
    out["workload"] = (f"91-wire strand (150 el/wire, r 0.4, L 80), dynamic radial squeeze {force} N per outer wire, "
                       "mu 0.3, 4 increments x up to 30 attempts (the reference's line search halves to 1/32)")
    return out


def coated_newton_loop(torch, wl):
    """A whole dynamic solve with the Kelvin-Voigt coating on: the 7-wire radial
    squeeze (workloads.squeeze_case(100), frictionless contact, coating
    stiffness 0.001 k_pen, coating mu 0.2, core radii 0.95 r), device attempts
    with the driver's tests on the device (a pair that starts pressing its
    coating without a pattern slot stops the attempt for a host rebuild), the
    same from graphs, and with the tests on the host (a host read per attempt
    and per assembly's pattern probe)."""
    TP = importlib.import_module("xkep-cae_amd.train_prescreen")
    sq = wl.squeeze_case(100.0)
    coat = dict(coating_stiffness=0.001 * sq["k_pen"], coating_damping=0.0, coating_mu=0.2, coating_k_t_ratio=0.5)
    out, traj = {}, {}
    for mode in ("device", "device_graph", "device_hostctl"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res, b = TP.run_case(sq, precision="fp64", device_solve=True, use_graph=mode == "device_graph",
                             device_control=mode != "device_hostctl", coating=coat, core_radii=0.95 * sq["radii"])
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        att = sum(c.n_attempts for c in res.calls)
        traj[mode] = [(c.n_attempts, c.n_active, c.converged) for c in res.calls]
        out[mode] = {"seconds": round(t, 3), "attempts": att, "ms_per_attempt": round(1e3 * t / max(att, 1), 2),
                     "pattern_rebuilds": b.asm.n_coat_rebuilds}
        log(f"[bench] coated Newton loop 7 wires {mode}: {out[mode]}")
    out["trajectories_equal"] = traj["device"] == traj["device_graph"] == traj["device_hostctl"]
    out["calls"] = len(traj["device"])
    out["workload"] = ("7-wire strand, dynamic radial squeeze 100 N per outer wire, Kelvin-Voigt coating "
                       "(k 0.001 k_pen, coating mu 0.2, core radii 0.95 r), FP64 device attempts")
    return out


def prescreen_leg(r, net, K, torch, steps, warmup, label, workload):
    """The learned prescreen on an assembled workload r (run()'s result): the
    watch list's step, and that step with the verification pass after a call of
    K assemblies and the per-increment features and selection spread over them."""
    asm, ctx = r["asm"], r["ctx"]
    asm.set_prescreen(net, elem_layer=asm.bench_layer)
    asm.begin_increment(detect=False)          # features + selection of the increment
    watched = asm.n_watch / max(ctx.n_pairs, 1)
    t_feat, _ = _timed(torch, lambda: ctx.lp_features(asm.u, asm.lp_layer, asm.lp_scale))
    t_sel, _ = _timed(torch, lambda: ctx.lp_select())
    stages = profile_stages(asm, ctx, torch, 20)
    ms, _ = time_graph(asm, torch, steps, warmup)
    # the verification pass after a call of K assemblies: replay of their geometry for the pruned pairs
    ctx.lp_select()
    for _ in range(K):
        asm.replay()
    t_ver, _ = _timed(torch, lambda: ctx.lp_verify(asm.u), reps=1)
    n_eval_watch = ctx.n_elems + asm.n_watch
    per_attempt_ms = ms + (t_ver + t_feat + t_sel) / K
    log(f"[bench] learned prescreen on {label}: watched {watched:.3f} of {ctx.n_pairs} pairs | "
        f"{ms * 1e3:.1f} us/step (graph) | stages us {np.round(stages * 1e3, 1).tolist()} | verify after {K} "
        f"assemblies {t_ver:.2f} ms, features {t_feat:.2f} ms, selection {t_sel:.2f} ms")
    rec = {
        "workload": workload,
        "watched_fraction": round(watched, 4), "watched_pairs": int(asm.n_watch),
        "us_per_step": round(ms * 1e3, 2), "stages_us": np.round(stages * 1e3, 2).tolist(),
        "verify_ms_per_call": round(t_ver, 3), "features_ms_per_increment": round(t_feat, 3),
        "select_ms_per_increment": round(t_sel, 3), "attempts_per_call": K,
        "us_per_attempt_amortized": round(per_attempt_ms * 1e3, 2),
        "value": round(n_eval_watch / (ms * 1e-3), 1),
        "value_amortized": round(n_eval_watch / (per_attempt_ms * 1e-3), 1),
        "speedup_vs_full_path": round(r["ms"] / ms, 3),
        "speedup_vs_full_path_amortized": round(r["ms"] / per_attempt_ms, 3),
        "value_note": "evals = elements + watched pairs (the narrowphase visits only those); amortized adds the "
                      "verification pass (geometry replay of the call's assemblies for the pruned pairs) per call "
                      "and the features and selection per increment, spread over the call's attempts",
        "exactness": "contributing pairs always visited, pruned pairs' geometry replayed per assembly: R, K_T, "
                     "n_active and every pair state identical to the full path"}
    asm.set_prescreen(None)
    ctx.lp_off()
    return rec


def side_measurements(batched, torch, wl, args):
    """Beside the headline: the learned prescreen on the batched 91-wire
    workload, the uneven 64-instance batch, and the 1000-wire strand (FP32-mixed,
    HIP-graph replay, per-step times), and one 91-wire Newton attempt with the
    device linear solve against host SuperLU."""
    out = {}
    steps = max(20, args.steps // 4)
    loop = newton_loop(torch, wl)
    out["newton_loop_91wire"] = loop
    out["coated_newton_loop_7wire"] = coated_newton_loop(torch, wl)
    att_per_call = loop["device"]["attempts"] / max(len(loop["calls"]), 1)
    P = importlib.import_module("xkep-cae_amd.prescreen")
    net = P.PrescreenNet.load() if os.path.exists(P.WEIGHTS) else None
    K = int(round(att_per_call))
    if net is not None:
        out["learned_prescreen"] = prescreen_leg(
            batched, net, K, torch, args.steps, args.warmup, f"mixed x{args.batch}",
            f"same batch of {args.batch} x 91 wires, pair kernel over the learned watch list")
    cases4, _, b4, mu4 = wl.uneven_batch()
    r4 = run("mixed uneven x64", cases4[0], None, "mixed", torch, steps, args.warmup, batch=b4, mu_elem=mu4)
    out["uneven_batch_64"] = {"n_elem": r4["n_elem"], "n_pair": r4["n_pair"], "us_per_step": round(r4["ms"] * 1e3, 2),
                              "value": round(r4["value"], 1), "stages_us": np.round(r4["stages"] * 1e3, 2).tolist()}
    del r4
    torch.cuda.empty_cache()
    st5 = wl.make_state(wl.C5)
    r5 = run("mixed 1000 wires", wl.C5, [st5], "mixed", torch, steps, args.warmup)
    out["wires_1000"] = {"n_elem": r5["n_elem"], "n_pair": r5["n_pair"], "us_per_step": round(r5["ms"] * 1e3, 2),
                         "value": round(r5["value"], 1), "stages_us": np.round(r5["stages"] * 1e3, 2).tolist(),
                         "increment_setup_ms": r5["inc"]["steady_ms"], "increment_first_ms": round(r5["t_inc"] * 1e3, 1),
                         "candidates_per_s": r5["inc"]["candidates_per_s"]}
    if net is not None:
        out["learned_prescreen_1000wire"] = prescreen_leg(
            r5, net, K, torch, steps, args.warmup, "mixed 1000 wires",
            "the 1000-wire strand, pair kernel over the learned watch list")
    del r5
    torch.cuda.empty_cache()
    out["coated_assembly"] = coated_assembly(wl.C3, batched["states"], torch)
    out["newton_attempt_91wire"] = newton_attempt(wl.C3, torch, wl)
    # the same batch with every candidate pair of the broadphase (no same-layer exclusion)
    ctx_u, asm_u, _ = setup(wl.C3, batched["states"], "mixed", torch, layer_filter=False)
    asm_u.assemble()
    torch.cuda.synchronize()
    st_u = profile_stages(asm_u, ctx_u, torch, 20)
    ms_u, _ = time_graph(asm_u, torch, steps, args.warmup)
    n_u = ctx_u.n_elems + ctx_u.n_pairs
    out["unfiltered_candidates"] = {"n_elem": ctx_u.n_elems, "n_pair": ctx_u.n_pairs,
                                    "us_per_step": round(ms_u * 1e3, 2), "value": round(n_u / (ms_u * 1e-3), 1),
                                    "stages_us": np.round(st_u * 1e3, 2).tolist(),
                                    "workload": f"the same {args.batch} x 91-wire batch without the same-layer "
                                                "exclusion: every broadphase candidate pair is a pair"}
    log(f"[bench] unfiltered candidates x{args.batch}: {out['unfiltered_candidates']}")
    del ctx_u, asm_u
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--batch", type=int, default=16, help="91-wire instances per batched launch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes and grid cell")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the side measurements (uneven 64-instance batch, 1000 wires, learned prescreen)")
    ap.add_argument("--newton-only", action="store_true",
                    help="only the 91-wire Newton attempt: device linear solve vs host SuperLU")
    ap.add_argument("--coated-only", action="store_true",
                    help="only the coated-assembly comparison on the batch (Kelvin-Voigt coating on and off)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {args.copy_calibration} MiB x 5")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(sharding by wire ownership is designed, not implemented).", file=sys.stderr)
        sys.exit(2)

    import torch

    wl = importlib.import_module("xkep-cae_amd.workloads")
    case = wl.C3
    if args.newton_only:
        print(json.dumps({"newton_attempt_91wire": newton_attempt(case, torch, wl),
                          "newton_loop_91wire": newton_loop(torch, wl),
                          "coated_newton_loop_7wire": coated_newton_loop(torch, wl)}), flush=True)
        return
    st = wl.make_state(case)
    if args.coated_only:
        states = [st] + [wl.make_state(case, seed=case.seed + 1000 * b) for b in range(1, args.batch)]
        print(json.dumps({"coated_assembly": coated_assembly(case, states, torch)}), flush=True)
        return
    plain = run("fp64 x1", case, [st], "fp64", torch, args.steps, args.warmup)
    single = run("mixed x1", case, [st], "mixed", torch, args.steps, args.warmup)
    del plain["asm"], plain["ctx"], single["asm"], single["ctx"]
    torch.cuda.empty_cache()
    states = [st] + [wl.make_state(case, seed=case.seed + 1000 * b) for b in range(1, args.batch)]
    batched = run(f"mixed x{args.batch}", case, states, "mixed", torch, args.steps, args.warmup,
                  tune=not args.no_autotune)
    batched["states"] = states
    # the FP64 plain form on the same batch (same tuned shapes through the cache)
    plain_b = run(f"fp64 x{args.batch}", case, states, "fp64", torch, args.steps, args.warmup,
                  tune=not args.no_autotune)
    del plain_b["asm"], plain_b["ctx"]
    torch.cuda.empty_cache()
    # the FP32 pair form (FP32 geometry of the non-contributing pairs) on the same batch
    f32_b = run(f"fp32 x{args.batch}", case, states, "fp32", torch, args.steps, args.warmup,
                tune=not args.no_autotune)
    del f32_b["asm"], f32_b["ctx"]
    torch.cuda.empty_cache()
    mixed_err = mixed_vs_fp64(case, states, torch)

    extra = {} if args.no_extra else side_measurements(batched, torch, wl, args)
    nnz_beam = beam_nnz(batched["ctx"].conn_np)
    sb = stage_bytes(batched["ctx"], batched["asm"], batched["n_pair"], "mixed", nnz_beam)
    dom = int(np.argmax(batched["stages"]))
    dom_ms = float(batched["stages"][dom])
    achieved = sb[STAGES[dom]] / (dom_ms * 1e-3) / 1e9
    kernels = stage_kernels(batched["ctx"].get_tuning())
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic_bytes(kernels[STAGES[dom]]),
            "kernel": kernels[STAGES[dom]], "algorithmic_bytes": int(sb[STAGES[dom]]),
            "kernel_us": round(dom_ms * 1e3, 2),
            "all_kernels": {k: {"us": round(float(t) * 1e3, 2), "algorithmic_bytes": int(sb[k]),
                                "GB_s": round(sb[k] / (float(t) * 1e-3) / 1e9, 1)}
                            for k, t in zip(STAGES, batched["stages"])}}
    # floating-point work per stage (measured instruction counts, committed summary)
    flop_base = {"pairs": "xk::pair_kernel<double", "elements": "xk::beam_direct_kernel<float>",
                 "gather": "xk::gather_kernel<float, float, double", "finish": "xk::finish_kernel"}
    for k, t in zip(STAGES, batched["stages"]):
        mf = kernel_mflop(flop_base[k])
        if mf is not None:
            roof["all_kernels"][k].update(mflop=mf, TFLOP_s=round(mf * 1e6 / (float(t) * 1e-3) / 1e12, 2),
                                          flop_per_byte=round(mf * 1e6 / max(sb[k], 1), 2))
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (FP64 NumPy oracle, 1 thread)")
        cv, n_runs, t_one = cpu_baseline(case, st, args.cpu_budget)
        cpu = {"value": round(cv, 1), "unit": "evals/s", "cores": 1, "kind": "port",
               "sample": f"{n_runs} full FP64 assemblies of one 91-wire step ({t_one:.3f} s each, median; "
                         f"evals = elements + candidate pairs), 1 thread (threadpoolctl) of {host_cpu()}"}
    value = batched["value"]
    out = {
        "metric": "assembled elem+contact-pair evals/s, FP32-mixed vs FP64 plain, 91 wires",
        "value": round(value, 1),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(batched["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp64 kinematics, forces, residual, norm; fp32 tangent)",
        "data": "synthetic",
        "config": {"workload": f"91-wire strand dynamic Newton assembly, batch of {args.batch} instances",
                   "instances": args.batch, "n_elem": batched["ctx"].n_elems, "n_pair": batched["n_pair"],
                   "n_nodes": batched["ctx"].n_nodes, "nnz_K": int(batched["asm"].K.numel()), "graph": True},
        "single_instance": {
            "fp64_plain": {"value": round(plain["value"], 1), "us_per_step": round(plain["ms"] * 1e3, 2),
                           "stages_us": np.round(plain["stages"] * 1e3, 2).tolist(),
                           "n_elem": plain["n_elem"], "n_pair": plain["n_pair"]},
            "fp32_mixed": {"value": round(single["value"], 1), "us_per_step": round(single["ms"] * 1e3, 2),
                           "stages_us": np.round(single["stages"] * 1e3, 2).tolist()},
        },
        "batched_stages_us": np.round(batched["stages"] * 1e3, 2).tolist(),
        "increment_setup_ms": round(batched["inc"]["steady_ms"], 2),
        "increment_setup": dict(batched["inc"], first_ms=round(batched["t_inc"] * 1e3, 2),
                                note="steady state: a later increment at the same displacement (buffers sized); "
                                     "first: the first increment after set-up, which sizes the pair-list and "
                                     "pattern buffers (allocations included)"),
        "tuning": None if batched["tuning"] is None else {k: batched["tuning"][k] for k in
                                                          ("cell_factor", "pair_block", "gather_block")},
        "roofline": roof,
        "cpu_baseline": cpu,
        "speedup_vs_fp64_gpu": round(value / plain_b["value"], 3),
        "speedup_vs_fp64_gpu_note": "FP32-mixed against FP64 plain on the same batch and launches",
        "fp64_batched": {"value": round(plain_b["value"], 1), "us_per_step": round(plain_b["ms"] * 1e3, 2),
                         "stages_us": np.round(plain_b["stages"] * 1e3, 2).tolist()},
        "mixed_vs_fp64": mixed_err,
        "fp32_pairs_batched": {"value": round(f32_b["value"], 1), "us_per_step": round(f32_b["ms"] * 1e3, 2),
                               "stages_us": np.round(f32_b["stages"] * 1e3, 2).tolist(),
                               "speedup_vs_fp64_gpu": round(f32_b["value"] / plain_b["value"], 3),
                               "vs_mixed": round(f32_b["value"] / value, 3),
                               "note": "FP32 pair form: the mixed form plus the FP32 narrowphase and frame of the "
                                       "non-contributing pairs, FP64 guard band at the hysteresis threshold; "
                                       "its errors are mixed_vs_fp64.fp32_pairs"},
        "speedup_vs_fp64_gpu_single_instance": round(value / plain["value"], 3),
        "speedup_vs_fp64_gpu_single_instance_note": "batched FP32-mixed against one unbatched FP64 instance "
                                                    "(batching and precision change together)",
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    loop = extra.get("newton_loop_91wire")
    if loop:
        # the increment's set-up (broadphase, geometry, pattern) spread over its Newton attempts
        att = loop["device"]["attempts"] / max(len(loop["calls"]), 1)
        out["increment_setup"]["attempts_per_increment"] = round(att, 2)
        out["increment_setup"]["amortized_us_per_attempt"] = round(1e3 * batched["inc"]["steady_ms"] / att, 2)
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
