"""Autotuner of the assembly path's launch shapes and broadphase grid cell.

No counterpart in the reference.  Tuned on the workload itself, timed with
HIP events (the library's per-stage profiler for the pair and gather kernels,
wall time around the synchronous candidate detection for the grid cell):

  pair_block    workgroup of the pair kernel        64, 128, 256, 512
  gather_block  workgroup of the gather kernel      128, 256, 512
  cell_factor   grid cell / mean box extent         0.5 .. 4  (the candidate set
                                                    does not depend on it)

Results are cached in a JSON file keyed by device, precision, element path and
the (power-of-two bucketed) element and pair counts, so a workload of the
same shape reuses them: ``$XKAMD_TUNE_CACHE`` or
``~/.cache/xkep-cae_amd/tuning.json``.
"""

from __future__ import annotations

import json
import math
import os
import time

PAIR_BLOCKS = (64, 128, 256, 512)
GATHER_BLOCKS = (128, 256, 512)
CELL_FACTORS = (0.5, 0.75, 1.0, 1.5, 2.0, 3.0, 4.0)


def cache_path() -> str:
    return os.environ.get("XKAMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "xkep-cae_amd", "tuning.json"))


def signature(ctx) -> str:
    torch = ctx.torch
    dev = torch.cuda.get_device_properties(0)
    arch = getattr(dev, "gcnArchName", dev.name).split(":")[0]
    b = lambda n: int(math.log2(max(n, 1)))
    return f"{arch}|{ctx.precision}|{'direct' if ctx.direct else 'staged'}|e{b(ctx.n_elems)}|p{b(ctx.n_pairs)}"


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(path, key, value):
    d = _load(path)
    d[key] = value
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _stage_us(asm, calls: int):
    ctx = asm.ctx
    asm.assemble()
    ctx.torch.cuda.synchronize()
    ctx.profile_begin(calls)
    for _ in range(calls):
        asm.assemble()
    ctx.torch.cuda.synchronize()
    ms, n = ctx.profile_end()
    return ms / max(n, 1) * 1e3


def tune_launch_shapes(asm, *, calls: int = 10) -> dict:
    """Pair and gather workgroup sizes minimising the event time of the pair,
    gather and finish kernels (the finish kernel reduces one partial per pair
    and per gather workgroup, so it depends on both)."""
    ctx = asm.ctx
    table = {"pair_block": {}, "gather_block": {}}
    for pb in PAIR_BLOCKS:
        ctx.set_tuning(pair_block=pb)
        st = _stage_us(asm, calls)
        table["pair_block"][pb] = float(st[0] + st[3])
    best_pb = min(table["pair_block"], key=table["pair_block"].get)
    ctx.set_tuning(pair_block=best_pb)
    for gb in GATHER_BLOCKS:
        ctx.set_tuning(gather_block=gb)
        st = _stage_us(asm, calls)
        table["gather_block"][gb] = float(st[2] + st[3])
    best_gb = min(table["gather_block"], key=table["gather_block"].get)
    ctx.set_tuning(gather_block=best_gb)
    asm.graph = None
    return {"pair_block": best_pb, "gather_block": best_gb, "table": table}


def tune_cell(asm, *, reps: int = 3) -> dict:
    """Grid cell factor minimising the candidate detection on the current
    configuration (the detection is repeated on the same coordinates, which
    leaves the pair list unchanged)."""
    ctx = asm.ctx
    x = asm.deformed_coords()
    table = {}
    for f in CELL_FACTORS:
        ctx.set_tuning(cell_factor=f)
        best = math.inf
        for _ in range(reps):
            ctx.torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.detect_candidates(x, elem_layer=asm._layer, midpoint_prescreening=asm.cfg.midpoint_prescreening,
                                  prescreening_margin=asm.cfg.prescreening_margin)
            best = min(best, time.perf_counter() - t0)
        table[f] = best * 1e3
    bf = min(table, key=table.get)
    ctx.set_tuning(cell_factor=bf)
    return {"cell_factor": bf, "table_ms": table}


def autotune(asm, *, use_cache: bool = True, calls: int = 10) -> dict:
    """Apply cached settings for this workload shape, or tune and cache them.
    The assembly must have had ``begin_increment`` (a pattern to time)."""
    ctx = asm.ctx
    key = signature(ctx)
    path = cache_path()
    if use_cache:
        hit = _load(path).get(key)
        if hit:
            ctx.set_tuning(cell_factor=hit["cell_factor"], pair_block=hit["pair_block"],
                           gather_block=hit["gather_block"])
            asm.graph = None
            return dict(hit, cached=True, key=key)
    cell = tune_cell(asm)
    asm.begin_increment(detect=False)          # pattern of the (unchanged) pair list
    shapes = tune_launch_shapes(asm, calls=calls)
    out = {"cell_factor": cell["cell_factor"], "pair_block": shapes["pair_block"],
           "gather_block": shapes["gather_block"]}
    _store(path, key, out)
    return dict(out, cached=False, key=key, tables={"cell_ms": cell["table_ms"], **shapes["table"]})
