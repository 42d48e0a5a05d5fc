"""Per-nw tile autotuner for the k-list R -> k transform (wbmi_r2k_klist_tile).

The phase-matrix GEMM out[k, e] = sum_R exp(2 pi i k.R) X[R, e] runs with one of
WBMI_KLIST_NTILES workgroup tiles (k-points x elements per workgroup).  Which is
fastest depends on nw (the elements per k-point are nch * nw^2, so the column
count and the share of ragged edge tiles change with nw) and on the output
precision.  `tune_klist` times every tile on the system's own channel stack and
a k-list of the given size, with HIP events on the context's stream after a
warm-up launch, and records the fastest in `ctx.tuning[("klist_tile", nw, prec)]`,
which `Context.r2k_klist` reads.  The table can be kept in a JSON file
(`save_table` / `load_table`) so that a process starts with the earlier choices.
"""
import json
import os

import numpy as np

from . import _lib
from .data_k import CHANNEL_NCOMP, device_system


def tile_shapes():
    """[(k-points, elements)] per workgroup for each tile index."""
    import ctypes
    out = []
    for t in range(64):   # until the library reports no such tile
        kr, el = ctypes.c_int(), ctypes.c_int()
        if _lib.lib().wbmi_r2k_klist_tile_shape(t, ctypes.byref(kr), ctypes.byref(el)) != _lib.WBMI_OK:
            break
        out.append((kr.value, el.value))
    return out


def tune_klist(system, channels, nk=4096, prec=_lib.WBMI_F64, ctx=None, reps=3, seed=0):
    """Time every k-list tile for `system` (its channel stack for `channels`, a list of
    (name, der)) on `nk` random k-points; record and return the fastest tile and the
    per-tile average times in ms."""
    import torch
    ctx = ctx or _lib.get_context(0)
    dev = device_system(system, ctx)
    keys = sorted(set(channels), key=list(CHANNEL_NCOMP).index)
    stack = dev.stack(keys)
    nw = system.num_wann
    nch = stack.shape[1]
    rng = np.random.default_rng(seed)
    kl = torch.from_numpy(rng.random((nk, 3))).to(ctx.device)
    cdt = torch.complex128 if prec == _lib.WBMI_F64 else torch.complex64
    out = torch.empty((nk, nch, nw, nw), dtype=cdt, device=ctx.device)
    times = {}
    saved, ctx.timer = ctx.timer, None
    try:
        for tile in range(len(tile_shapes())):
            ctx.r2k_klist(dev.iRvec_dev, stack, kl, out, prec, tile=tile)   # warm-up
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record(ctx.stream)
            for _ in range(reps):
                ctx.r2k_klist(dev.iRvec_dev, stack, kl, out, prec, tile=tile)
            e1.record(ctx.stream)
            e1.synchronize()
            times[tile] = e0.elapsed_time(e1) / reps
    finally:
        ctx.timer = saved
    best = min(times, key=times.get)
    ctx.tuning[("klist_tile", nw, int(prec))] = best
    return best, times


def table(ctx=None):
    """The context's tuning choices as a JSON-ready list."""
    ctx = ctx or _lib.get_context(0)
    return [{"knob": k, "nw": nw, "prec": p, "value": v} for (k, nw, p), v in sorted(ctx.tuning.items())]


def save_table(path, ctx=None):
    with open(path, "w") as f:
        json.dump(table(ctx), f, indent=1)


def load_table(path, ctx=None):
    """Merge choices saved by save_table into the context (a missing file is no error)."""
    ctx = ctx or _lib.get_context(0)
    if not os.path.exists(path):
        return 0
    with open(path) as f:
        rows = json.load(f)
    for r in rows:
        ctx.tuning[(r["knob"], int(r["nw"]), int(r["prec"]))] = int(r["value"])
    return len(rows)
