"""Benchmark of the per-Newton-iteration assembly on one MI355X.

Workload: the 91-wire strand (layers 1+6+12+18+24+30, r = 0.4 mm, L = 80 mm,
150 elements per wire, lay angle 12 deg), Generalized-alpha (rho_inf = 0.8)
dynamic step, Huber contact k = 1e4 E r, Coulomb mu = 0.3, seeded
perturbation u ~ N(0, (1e-3 r)^2).  One step = one complete fused assembly
(narrowphase over every candidate pair, Huber, friction, element f_int and K,
R and K composition with inertia terms, ||R||^2) replayed from a HIP graph,
inputs resident in HBM.  Metric: (N_elem + N_pair) / t_step in the FP32-mixed
form; the FP64 plain form is timed beside it.

Prints ONE JSON line (rank 0, one GPU).
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup(case, st, precision, torch):
    C = importlib.import_module("xkep-cae_amd.context")
    A = importlib.import_module("xkep-cae_amd.assembly")
    T = importlib.import_module("xkep-cae_amd.time_integration")
    E = importlib.import_module("xkep-cae_amd.elements")
    m = st.mesh
    sec = C.circle_section(case.wire_radius, case.E, case.nu)
    ctx = C.DeviceContext(m.node_coords, m.connectivity, sec, radii=m.radii, precision=precision)
    Md = E.ULCRBeamAssembler(m.node_coords, m.connectivity, sec.E, sec.G, sec.A, sec.Iy, sec.Iz, sec.J,
                             sec.kappa_y, sec.kappa_z).lumped_mass_diag(case.rho)
    dt = case.dt()
    ga = T.GeneralizedAlphaProcess(Md, rho_inf=case.rho_inf)
    ga.predict_device(torch.zeros(ctx.ndof, dtype=torch.float64, device=ctx.dev), dt)
    asm = A.StrandAssembly(ctx, k_pen=case.k_pen, smoothing_delta=case.smoothing_delta, mu=case.mu,
                           f_ext=st.f_ext, fixed_dofs=st.fixed_dofs, time=ga, dt=dt)
    t0 = time.perf_counter()
    asm.begin_increment(np.zeros(ctx.ndof))   # set-up broadphase on the reference configuration
    asm.begin_increment(st.u)                  # increment: broadphase + hysteresis + pattern
    torch.cuda.synchronize()
    t_inc = time.perf_counter() - t0
    return ctx, asm, t_inc


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


def stage_bytes(ctx, asm, n_pair, precision):
    """Algorithmic (compulsory) bytes per launch of each stage."""
    nn, ne = ctx.n_nodes, ctx.n_elems
    s_k = 8 if precision == "fp64" else 4
    nnz = asm.K.numel()
    return {
        "pairs": n_pair * 216,
        "elements": nn * (24 + 48) + ne * 8,
        "residual": nn * (48 + 48 + 48),
        "tangent": nnz * s_k,
    }


def cpu_baseline(case, st, budget_s=15.0):
    """The FP64 NumPy oracle's full assembly on this host, one thread."""
    from threadpoolctl import threadpool_limits

    from oracle import assembly as oasm
    from oracle import beam as obeam
    from oracle import contact as ocon

    m = st.mesh
    osec = obeam.circle_section(case.wire_radius, case.E, case.nu)
    cfg = ocon.ContactConfig(mu=case.mu)
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
            res = oasm.assemble_iteration(prob, op, st.u, st.u_ref, st.f_ext, st.fixed_dofs, ga, dt)
            times.append(time.perf_counter() - t0)
            if time.perf_counter() - t_start > budget_s and len(times) >= 3:
                break
    t = float(np.median(times[1:])) if len(times) > 2 else float(np.median(times))
    return (m.n_elems + len(op)) / t, len(times), t


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(the multi-GPU sharding by wire ownership is a design, not an implementation).", file=sys.stderr)
        sys.exit(2)

    import torch

    wl = importlib.import_module("xkep-cae_amd.workloads")
    case = wl.C3
    st = wl.make_state(case)
    results = {}
    for prec in ("fp64", "mixed"):
        log(f"[bench] setting up {case.name} precision={prec}")
        ctx, asm, t_inc = setup(case, st, prec, torch)
        n_pair = ctx.n_pairs
        log(f"[bench] {prec}: nodes={ctx.n_nodes} elems={ctx.n_elems} pairs={n_pair} "
            f"nnzK={asm.K.numel()} increment setup {t_inc * 1e3:.1f} ms")
        asm.assemble()
        torch.cuda.synchronize()
        stages = profile_stages(asm, ctx, torch, 20)
        ms_graph, ms_wall = time_graph(asm, torch, args.steps, args.warmup)
        log(f"[bench] {prec}: {ms_graph * 1e3:.1f} us/step (graph, events), wall {ms_wall * 1e3:.1f} us/step, "
            f"stages us {np.round(stages * 1e3, 1).tolist()}")
        results[prec] = dict(ctx=ctx, asm=asm, n_pair=n_pair, ms=ms_graph, ms_wall=ms_wall, stages=stages,
                             t_inc=t_inc)
    mixed = results["mixed"]
    plain = results["fp64"]
    n_eval = mixed["ctx"].n_elems + mixed["n_pair"]
    value = n_eval / (mixed["ms"] * 1e-3)
    # roofline of the dominant stage of the mixed step
    names = ["pairs", "elements", "residual", "tangent"]
    sb = stage_bytes(mixed["ctx"], mixed["asm"], mixed["n_pair"], "mixed")
    dom = int(np.argmax(mixed["stages"]))
    dom_ms = float(mixed["stages"][dom])
    achieved = sb[names[dom]] / (dom_ms * 1e-3) / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None, "kernel": names[dom],
            "kernel_bytes": sb[names[dom]], "kernel_us": round(dom_ms * 1e3, 2)}
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] CPU baseline (FP64 NumPy oracle, 1 thread)")
        cv, n_runs, t_one = cpu_baseline(case, st, args.cpu_budget)
        cpu = {"value": round(cv, 1), "unit": "evals/s", "cores": 1, "kind": "port",
               "sample": f"{n_runs} full FP64 assemblies of the 91-wire step ({t_one:.2f} s each, median)"}
    out = {
        "metric": "assembled elem+contact-pair evals/s, FP32-mixed vs FP64 plain, 91 wires",
        "value": round(value, 1),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (fp64 kinematics/forces/residual, fp32 tangent)",
        "data": "synthetic",
        "config": {"workload": "91-wire strand dynamic Newton assembly (C3)", "n_elem": mixed["ctx"].n_elems,
                   "n_pair": mixed["n_pair"], "n_nodes": mixed["ctx"].n_nodes, "nnz_K": int(mixed["asm"].K.numel()),
                   "graph": True},
        "fp64_plain": {"value": round(n_eval / (plain["ms"] * 1e-3), 1), "ms_per_step": round(plain["ms"], 5),
                       "stages_us": np.round(plain["stages"] * 1e3, 2).tolist()},
        "mixed_stages_us": np.round(mixed["stages"] * 1e3, 2).tolist(),
        "increment_setup_ms": round(mixed["t_inc"] * 1e3, 1),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    out["speedup_vs_fp64_gpu"] = round(out["value"] / out["fp64_plain"]["value"], 3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
