"""Benchmark: k-points/s of AHC + Berry-curvature dipole (Fermi sea) at nw=32 on one MI355X.

Workload (the bcc nw=32 synthetic model): num_wann=32, bcc lattice a=3 A,
R in {-4..4}^3 (729) with Wigner-Seitz weights, seed 2; AHC + BCD_sea with
external terms (AA_R present) on a 96^3 grid (NKFFT 16^3 x NKdiv 6^3) at 2001
Fermi levels spanning the bands, linspace(-22, 22, 2001) eV.

A step is one full run() over the 96^3 grid: R->k, eigh, U^+XU, formula
traces, Fermi scan and k-sum of all 216 K-blocks.  Both forms are timed:
the FP64 form and the mixed form (complex64 channels; FP64 Hamiltonian,
refined FP64 energies, FP64 denominators, traces and k-sum).  `value` is the
mixed form's throughput when its AHC error against the FP64 oracle on the CPU
sample meets 1e-6 (the metric's bar), else the FP64 form's.

`roofline.traffic` is the HBM bytes per launch of the dominant kernel from
rocprofv3 FETCH_SIZE + WRITE_SIZE counters (scaled by the counters' reading
for a plain copy, `--calib-copy`), read from profiles/hbm_traffic.json when
that file holds the kernel; null otherwise.

Prints one JSON line (rank 0).  `--gpus N` with N > 1 exits: this build runs
on one GPU.
"""
import argparse
import json
import multiprocessing as mp
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np

NW, RMAX, SEED = 32, 4, 2
NK = 96
EFERMI = (-22.0, 22.0, 2001)
# algorithmic work per k-point (AHC + BCD, external terms, nw=32, 37 channels)
FLOPS_ROTATE_PER_K_PER_CH = 2 * 8 * NW ** 3          # U^+ (X U): two complex nw^3 products
FLOPS_SCAN_BCD_PER_K = 18 * 8 * NW ** 3               # DerDcov V.D and D.V products (9 + 9)
FLOPS_SCAN_AHC_PER_K = 3 * 3 * 8 * NW ** 2            # Omega: 3 comps x 3 terms x nw^2 cmacs
FLOPS_EIGH_PER_K = int((16 / 3 + 8) * NW ** 3)        # LAPACK-equivalent Hermitian eigensolver
PEAK_HBM = 8000.0        # GB/s
PEAK_FP64 = 78.6         # TFLOP/s (dense MFMA and vector FP64)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_system():
    from wannierberri_amd import models
    return models.synthetic(num_wann=NW, Rmax=RMAX, seed=SEED)


def calculators(wb):
    Ef = np.linspace(*EFERMI)
    st = wb.calculators.static
    return {"ahc": st.AHC(Efermi=Ef), "bcd": st.BerryDipole_FermiSea(Efermi=Ef)}


# -- CPU baseline: the FP64 oracle, one worker process per host core ------------

def _cpu_worker(rank, dKs, nkfft, barrier, queue):
    """One process: imports, model and oracle system built first, then (after all workers are
    ready) the timed compute of its K-blocks: R->k, eigh, rotation, AHC and BCD traces, scan."""
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    from oracle import wb_oracle as O
    s = make_system()
    o = O.OracleSystem(s.real_lattice, s.rvec.iRvec, s.wannier_centers_cart,
                       {"Ham": s.get_R_mat("Ham"), "AA": s.get_R_mat("AA")})
    Ef = np.linspace(*EFERMI)
    barrier.wait()
    t0 = time.time()
    out = []
    for dK in dKs:
        d = O.OracleDataK(o, dK=dK, NKFFT=nkfft)
        out.append((d.nk, O.AHC(Ef)(d).data, O.BerryDipole_FermiSea(Ef)(d).data))
    queue.put((rank, t0, time.time(), out))


def host_cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=len(os.sched_getaffinity(0)), cpu_count=os.cpu_count(), model=model)


def cpu_baseline(nproc, nkfft, blocks_per_proc=2):
    """The oracle on `nproc` processes at once, `blocks_per_proc` K-blocks each.  The rate is the
    total k-points over the window from the first worker's start to the last one's end, after
    every worker has imported its modules and built its model (a barrier)."""
    rng = np.random.default_rng(7)
    dKs = [rng.uniform(0, 1, 3) / np.array(nkfft) for _ in range(nproc * blocks_per_proc)]
    ctx = mp.get_context("spawn")
    barrier, queue = ctx.Barrier(nproc), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=(r, dKs[r::nproc], nkfft, barrier, queue)) for r in range(nproc)]
    for p in procs:
        p.start()
    got = sorted(queue.get() for _ in procs)
    for p in procs:
        p.join()
    t_start, t_end = min(g[1] for g in got), max(g[2] for g in got)
    blocks = [b for g in got for b in g[3]]
    order = [i for r in range(nproc) for i in range(r, len(dKs), nproc)]
    nk = sum(b[0] for b in blocks)
    wall = t_end - t_start
    return dict(wall=wall, nk=nk, kps=nk / wall, dKs=[dKs[i] for i in order],
                ahc=[b[1] for b in blocks], bcd=[b[2] for b in blocks])


# -- GPU -------------------------------------------------------------------------

def gpu_step(wb, system, grid, calcs, precision):
    return wb.run(system, grid, calcs, parameters_K={"precision": precision}, save_files=False, verbose=False)


def stage_stats(timer):
    """{stage: (calls, total_ms, [units])} from the event timer."""
    out = {}
    for stage, lst in timer.items():
        ms = [e0.elapsed_time(e1) for e0, e1, _ in lst]
        out[stage] = (len(ms), float(np.sum(ms)), [u for _, _, u in lst])
    return out


def roofline(stage, calls, total_ms, units, prec):
    # rotation, formula products and the eigensolver's refinement run in FP64
    # arithmetic in both forms (FP64 MFMA); the forms differ in storage
    avg_s = total_ms / calls / 1e3
    peak_f = PEAK_FP64
    if stage == "rotate":
        nk, nw, nrot = units[0]
        work = nk * nrot * FLOPS_ROTATE_PER_K_PER_CH
        return dict(bound="mfma", achieved=work / avg_s / 1e12, peak=peak_f, unit="TFLOP/s")
    if stage.startswith("scan_f2"):
        nk = units[0][0]
        return dict(bound="mfma", achieved=nk * FLOPS_SCAN_BCD_PER_K / avg_s / 1e12, peak=PEAK_FP64, unit="TFLOP/s")
    if stage.startswith("scan_f1"):
        nk = units[0][0]
        return dict(bound="mfma", achieved=nk * FLOPS_SCAN_AHC_PER_K / avg_s / 1e12, peak=PEAK_FP64, unit="TFLOP/s")
    if stage == "eigh":
        nk = units[0][0]
        return dict(bound="mfma", achieved=nk * FLOPS_EIGH_PER_K / avg_s / 1e12, peak=peak_f, unit="TFLOP/s")
    if stage == "r2k":
        nk, nelem, box_bytes, work_bytes = units[0]
        esz = 16 if prec == "fp64" else 8
        # algorithmic: read the R-space box once, write the k-space channels once
        byt = box_bytes + nk * nelem * esz
        return dict(bound="hbm", achieved=byt / avg_s / 1e9, peak=PEAK_HBM, unit="GB/s")
    return None


def calib_copy(nbytes=1 << 30, reps=5):
    """Device-to-device copies of `nbytes` (the reference kernel for the HBM counters)."""
    import torch
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    log(f"[bench] calib-copy: {reps} copies of {nbytes} bytes")


def traffic_of(stage, prec):
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    return table.get(f"{stage}@{prec}", {}).get("bytes_per_launch")


KL_NW = (8, 16, 24, 36, 48)     # the five k-list models (seeds 40..44)


def bench_klists(args):
    """The k-list workload: 512 k-lists (sizes log-uniform in [8, 4096], points
    ~ N(centre, 0.02) around 16 centres, seed 4), list i on model i % 5
    (synthetic recipe, nw = 8, 16, 24, 36, 48, R cube |R_i| <= 4), AHC with
    external terms at 2001 levels, every point weighted 1 / (its model's
    points).  The k-list transform's workgroup tile is chosen per nw by the
    autotuner (untimed), then all lists of all five models are captured as one
    HIP graph (KlistGraph); a step is one replay of that graph."""
    import torch
    import wannierberri_amd as wb
    from wannierberri_amd import _lib, autotune, models

    lists = models.klist_batch()
    systems = [models.synthetic(num_wann=nw, Rmax=RMAX, seed=40 + i) for i, nw in enumerate(KL_NW)]
    groups = [[lists[i] for i in range(len(lists)) if i % len(KL_NW) == m] for m in range(len(KL_NW))]
    npts = sum(len(k) for k in lists)
    Ef = np.linspace(*EFERMI)
    calcs = {"ahc": wb.calculators.static.AHC(Efermi=Ef)}
    ctx = _lib.get_context(0)
    prec = args.precisions.split(",")[0]
    wprec = _lib.WBMI_F64 if prec == "fp64" else _lib.WBMI_F32
    tiles = {}
    for s in systems:
        need = sorted(calcs["ahc"].needed_channels(s), key=list(wb.data_k.CHANNEL_NCOMP).index)
        need = [k for k in need if not (prec == "fp32" and k == ("Ham", 0))]
        best, times = autotune.tune_klist(s, need, nk=8192, prec=wprec)
        tiles[s.num_wann] = {"tile": autotune.tile_shapes()[best], "ms": {str(autotune.tile_shapes()[t]): round(v, 3)
                                                                          for t, v in times.items()}}
        log(f"[bench] nw={s.num_wann}: tile {autotune.tile_shapes()[best]} ({times})")
    pk = {"precision": prec}
    t0 = time.perf_counter()
    g = wb.KlistGraph(list(zip(systems, groups)), calcs, parameters_K=pk)
    log(f"[bench] graph built in {time.perf_counter() - t0:.1f} s ({npts} points)")

    def eager_step():
        return [wb.run_klists(s, gl, calcs, parameters_K=pk) for s, gl in zip(systems, groups)]

    for _ in range(args.warmup):
        g.replay()
        eager_step()
    torch.cuda.synchronize()
    ctx.timer = {}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        eager_step()
    torch.cuda.synchronize()
    dt_eager = (time.perf_counter() - t0) / args.steps
    stats = stage_stats(ctx.timer)
    ctx.timer = None
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"[bench] klists: graph {dt * 1e3:.1f} ms/step, eager {dt_eager * 1e3:.1f} ms/step")
    for st, (n, ms, _) in sorted(stats.items(), key=lambda x: -x[1][1]):
        log(f"    {st:12s} calls/step={n // args.steps:5d} ms/step={ms / args.steps:9.2f} avg={ms / n:.3f} ms")
    n, ms, units = stats["r2k_klist"]
    flops = sum(8.0 * nk * nelem * nR for nk, nelem, nR in units)
    ach = flops / (ms / 1e3) / 1e12
    cb = None
    if not args.no_cpu:
        # oracle on (at most 200 points of) the first list of every model, one process
        from oracle import wb_oracle as O
        t = time.time()
        nk_cpu, err = 0, 0.0
        for s, gl, r in zip(systems, groups, res):
            o = O.OracleSystem(s.real_lattice, s.rvec.iRvec, s.wannier_centers_cart,
                               {"Ham": s.get_R_mat("Ham"), "AA": s.get_R_mat("AA")})
            sample = gl[0][:200]
            ref = O.AHC(Ef)(O.OracleDataK(o, k_list=sample)).data
            nk_cpu += len(sample)
            got = wb.run_klists(s, [sample], calcs, parameters_K=pk).results["ahc"].data
            err = max(err, float(np.abs(got - ref).max() / np.abs(ref).max()))
        wall = time.time() - t
        cb = {"value": nk_cpu / wall, "unit": "k-points/s", "cores": 1, "kind": "port",
              "sample": f"<= 200 points of the first list of each model ({nk_cpu} points), FP64 NumPy "
                        f"oracle, one process, "
                        f"wall {wall:.1f} s; GPU rel err on it {err:.2e}"}
    # the graph's totals against the eager run_klists totals
    eg = eager_step()
    gerr = max(float(np.abs(r.results["ahc"].data - e.results["ahc"].data).max() /
                     np.abs(e.results["ahc"].data).max()) for r, e in zip(g.replay(), eg))
    out = {
        "metric": "k-pts/s AHC over 512 weighted k-lists, 5 models nw 8-48, on 1 MI355X",
        "value": npts / dt, "unit": "k-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"512 k-lists (log-uniform 8..4096 points, {npts} total) on 5 synthetic models "
                               f"nw {KL_NW}, R cube |R_i|<={RMAX} (729), AHC at {EFERMI[2]} levels, one HIP graph "
                               f"replay per step, k-list tiles autotuned per nw",
                   "nk": npts, "precision": "mixed" if prec == "fp32" else "fp64"},
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_FP64, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64, "kernel": "r2k_klist", "traffic": traffic_of("r2k_klist", prec)},
        "eager_ms_per_step": dt_eager * 1e3,
        "graph_vs_eager_rel_err": gerr,
        "tiles": tiles,
        "stages_ms_per_step": {st: v[1] / args.steps for st, v in stats.items()},
    }
    if cb is not None:
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


def bench_config(args):
    """The other model configurations at their own sizes, as one more JSON line each (not the
    headline): `haldane` (2-band Haldane model, AHC at 1001 levels in [-3, 3] eV, 200 x 200 x 1 grid
    as NKFFT 50 x 50 x 1, FP64, the whole grid per step), `spinor64` (nw = 64 spinor model, R cube
    |R_i| <= 5 (1331): AHC with external terms, Morb and GME sea / surface with internal terms (the
    model has no BB, CC), 128^3 grid as NKFFT 16^3 K-blocks) and `large128` (nw = 128, R cube
    |R_i| <= 6 (2197): AHC, BCD sea with external terms and GME sea / surface with internal terms,
    4001 levels over the band range, 200^3 grid as NKFFT 8^3 K-blocks).  For the two large
    configurations a step is the first `--blocks` K-blocks of the grid's K-list (every K-block
    of a grid does the same work); value = k-points of those blocks / step time."""
    import torch
    import wannierberri_amd as wb
    from wannierberri_amd import models
    st = wb.calculators.static
    prec = args.precisions.split(",")[-1]
    ki = {"external_terms": False}
    if args.workload == "haldane":
        system = models.haldane()
        Ef = np.linspace(-3, 3, 1001)
        calcs = {"ahc": st.AHC(Efermi=Ef)}
        grid = wb.Grid(system=system, NKFFT=[50, 50, 1], NKdiv=[4, 4, 1])
        prec = "fp64"
        quant, nwdesc = "AHC", "2-band Haldane model, 1001 levels in [-3, 3] eV, 200x200x1 grid as NKFFT 50x50x1"
    elif args.workload == "spinor64":
        system = models.spinor(num_wann_half=32, Rmax=5, seed=3)
        E0 = np.linalg.eigvalsh(system.get_R_mat("Ham")[np.where((system.rvec.iRvec == 0).all(1))[0][0]])
        Ef = np.linspace(E0.min() - 1, E0.max() + 1, 2001)
        calcs = {"ahc": st.AHC(Efermi=Ef), "morb": st.Morb(Efermi=Ef, kwargs_formula=ki),
                 "gme": st.GME_orb_FermiSea(Efermi=Ef, kwargs_formula=ki),
                 "gmes": st.GME_orb_FermiSurf(Efermi=Ef, kwargs_formula=ki)}
        grid = wb.Grid(system=system, NKFFT=[16, 16, 16], NKdiv=[8, 8, 8])
        quant = "AHC+Morb+GME"
        nwdesc = (f"nw=64 spinor model (2x32 spin blocks, SOC block std 0.05 eV, seed 3), R cube |R_i|<=5 "
                  f"({system.rvec.nRvec}), 2001 levels over the onsite band range, 128^3 grid as NKFFT 16^3")
    else:
        system = models.synthetic(num_wann=128, Rmax=6, seed=5)
        Ef = np.linspace(-16.0, 16.0, 4001)
        calcs = {"ahc": st.AHC(Efermi=Ef), "bcd": st.BerryDipole_FermiSea(Efermi=Ef),
                 "gme": st.GME_orb_FermiSea(Efermi=Ef, kwargs_formula=ki),
                 "gmes": st.GME_orb_FermiSurf(Efermi=Ef, kwargs_formula=ki)}
        grid = wb.Grid(system=system, NKFFT=[8, 8, 8], NKdiv=[25, 25, 25])
        quant = "AHC+BCD+GME"
        nwdesc = (f"nw=128 synthetic model (seed 5), R cube |R_i|<=6 ({system.rvec.nRvec}), 4001 levels in "
                  f"[-16, 16] eV, 200^3 grid as NKFFT 8^3")
    K_list = grid.get_K_list(use_symmetry=False)
    blocks = K_list if args.workload == "haldane" else K_list[:args.blocks]
    need = set()
    for c in calcs.values():
        need |= set(c.needed_channels(system))
    pk = {"precision": prec, "channels": need}

    def step():
        tot = None
        for Kp in blocks:
            d = wb.Data_K(system, dK=Kp.Kp_fullBZ, grid=grid, Kpoint=Kp, **pk)
            r = wb.ResultDict({k: c(d) for k, c in calcs.items()}) * Kp.factor
            tot = r if tot is None else tot + r
            del d
        return tot

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    nk = len(blocks) * int(np.prod(grid.FFT))
    log(f"[bench] {args.workload}: {len(blocks)} of {len(K_list)} K-blocks, {nk} k-points, {dt * 1e3:.1f} ms/step")
    out = {"metric": f"k-pts/s {quant} at nw={system.num_wann} on 1 MI355X", "value": nk / dt, "unit": "k-points/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64" if prec == "fp64" else "f32+f64", "data": "synthetic",
           "config": {"workload": nwdesc, "nk_measured": nk, "k_blocks_measured": len(blocks),
                      "k_blocks_grid": len(K_list), "nk_grid": int(np.prod(grid.dense)),
                      "precision": "mixed" if prec == "fp32" else "fp64"}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["grid", "klists", "haldane", "spinor64", "large128"], default="grid",
                    help="grid: AHC+BCD on the 96^3 grid (default); klists: AHC over 512 k-lists; haldane, "
                         "spinor64, large128: the other model configurations (one JSON line each)")
    ap.add_argument("--blocks", type=int, default=8, help="K-blocks per step of spinor64 / large128")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nkfft", type=int, default=16)
    ap.add_argument("--precisions", default="fp64,fp32")
    ap.add_argument("--cpu-procs", type=int, default=0, help="0 = all host cores (up to 16)")
    ap.add_argument("--cpu-nkfft", default="6,6,4")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib-copy", action="store_true", help="only run the plain-copy calibration kernel")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        calib_copy()
        return
    if args.workload == "klists":
        bench_klists(args)
        return
    if args.workload in ("haldane", "spinor64", "large128"):
        bench_config(args)
        return

    import torch
    import wannierberri_amd as wb
    from wannierberri_amd import _lib

    system = make_system()
    nkfft = [args.nkfft] * 3
    assert NK % args.nkfft == 0
    grid = wb.Grid(system=system, NKFFT=nkfft, NKdiv=[NK // args.nkfft] * 3)
    calcs = calculators(wb)
    nk_total = NK ** 3
    ctx = _lib.get_context(0)
    precs = args.precisions.split(",")
    res = {}
    for prec in precs:
        log(f"[bench] {prec}: warmup {args.warmup}")
        for _ in range(args.warmup):
            r = gpu_step(wb, system, grid, calcs, prec)
            torch.cuda.synchronize()
        ctx.timer = {}
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            r = gpu_step(wb, system, grid, calcs, prec)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        timer, ctx.timer = ctx.timer, None
        stats = stage_stats(timer)
        res[prec] = dict(ms=dt * 1e3, kps=nk_total / dt, stats=stats,
                         ahc=r.results["ahc"].data.copy(), bcd=r.results["bcd"].data.copy())
        log(f"[bench] {prec}: {dt * 1e3:.1f} ms/step, {nk_total / dt:.4g} k-pts/s")
        for st, (n, ms, _) in sorted(stats.items(), key=lambda x: -x[1][1]):
            log(f"    {st:12s} calls/step={n // args.steps:5d} ms/step={ms / args.steps:9.2f} avg={ms / n:.3f} ms")

    out_err = {}
    cb = None
    if not args.no_cpu:
        hw = host_cpu_info()
        # one process per core of this process's CPU set, at most 16: a one-GPU share of the host
        nproc = args.cpu_procs or min(16, hw["nproc"])
        cnk = [int(x) for x in args.cpu_nkfft.split(",")]
        log(f"[bench] cpu baseline: oracle on {nproc} processes ({hw}), two {cnk} K-blocks each")
        cb = cpu_baseline(nproc, cnk)
        # GPU on the same K-blocks: error of each GPU form vs the FP64 oracle
        g = wb.Grid(system=system, NKFFT=cnk, NKdiv=[1, 1, 1])
        for prec in res:
            ea, eb = 0.0, 0.0
            for dK, a, b in zip(cb["dKs"], cb["ahc"], cb["bcd"]):
                d = wb.Data_K(system, dK=dK, grid=g, precision=prec)
                ra = calcs["ahc"](d).data
                rb = calcs["bcd"](d).data
                ea = max(ea, np.abs(ra - a).max() / np.abs(a).max())
                eb = max(eb, np.abs(rb - b).max() / np.abs(b).max())
            out_err[prec] = {"ahc": ea, "bcd": eb}

    # headline form: the mixed form when it meets the AHC bar against the oracle
    main_prec = "fp64" if "fp64" in res else precs[0]
    if "fp32" in res and out_err.get("fp32", {}).get("ahc", 1.0) <= 1e-6:
        main_prec = "fp32"
    m = res[main_prec]
    # the dominant kernel: the stage of one kernel launch per K-block (rotation, split rotation,
    # the two scans) with the largest time; R -> k (three DFT passes and their set-up, on a side
    # stream overlapped with the other stages by run(prefetch=True)) and eigh (Jacobi + refinement)
    # are several kernels each, every one of them shorter than the longest of these
    single = ("rotate", "rotate_split", "scan_f1", "scan_f2")
    dom = max(((k, v) for k, v in m["stats"].items() if k in single), key=lambda x: x[1][1])
    rl = roofline(dom[0], dom[1][0], dom[1][1], dom[1][2], main_prec)
    rl["frac"] = rl["achieved"] / rl["peak"]
    rl["kernel"] = dom[0]
    rl["traffic"] = traffic_of(dom[0], main_prec)

    out = {
        "metric": "k-pts/s AHC+BCD at nw=32: mixed-prec vs FP64 on 1 MI355X, rel err<=1e-6",
        "value": m["kps"], "unit": "k-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64", "data": "synthetic",
        "config": {"workload": f"bcc nw={NW} synthetic (seed {SEED}), R cube |R_i|<={RMAX} (729), AHC+BCD_sea, "
                               f"{NK}^3 grid as NKFFT {args.nkfft}^3 x NKdiv {NK // args.nkfft}^3, "
                               f"{EFERMI[2]} Fermi levels in [{EFERMI[0]}, {EFERMI[1]}] eV",
                   "nk": nk_total, "num_wann": NW, "nR": system.nRvec,
                   "precision": "mixed" if main_prec == "fp32" else "fp64"},
        "roofline": rl,
        "forms": {("mixed" if p == "fp32" else p): {"value": res[p]["kps"], "ms_per_step": res[p]["ms"],
                                                     "stages_ms_per_step": {st: v[1] / args.steps for st, v in
                                                                            res[p]["stats"].items()}}
                  for p in res},
    }
    if "fp64" in res and "fp32" in res:
        ref = res["fp64"]
        errs = {q: float(np.abs(res["fp32"][q] - ref[q]).max() / np.abs(ref[q]).max()) for q in ("ahc", "bcd")}
        out["mixed_vs_fp64"] = {"speedup": res["fp32"]["kps"] / ref["kps"], "rel_err_vs_fp64_gpu": errs}
    if cb is not None:
        out["rel_err_vs_oracle_on_cpu_sample"] = {("mixed" if p == "fp32" else p): e for p, e in out_err.items()}
        out["cpu_baseline"] = {"value": cb["kps"], "unit": "k-points/s", "cores": nproc, "kind": "port",
                               "sample": f"{len(cb['dKs'])} K-blocks of {cnk[0]}x{cnk[1]}x{cnk[2]} k-points ({cb['nk']} "
                                         f"k, same model, AHC+BCD, 2001 levels), FP64 NumPy oracle, one process per "
                                         f"core, OPENBLAS_NUM_THREADS=1, compute window {cb['wall']:.1f} s after "
                                         f"imports and model build; host: {hw['model']}, {hw['nproc']} CPUs in "
                                         f"this process's set (os.cpu_count {hw['cpu_count']})"}
        out["gpu_over_cpu"] = m["kps"] / cb["kps"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
