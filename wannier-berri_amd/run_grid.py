"""run(): integrate calculators over a Grid, K-block by K-block, on the MI355X.

Mirrors wannierberri/run_grid.py:32-403 for one process: the (irreducible)
K-list of the Grid, one data_k_class per K-block, point-group symmetrisation
of each K-block's results, the k-sum `sum_K result * Kp.factor`, optional
adaptive refinement (run_grid.py:352-389, Kpoint.divide,
exclude_equiv_points) and per-iteration output files
`{fout_name}-{calc}{-suffix}_iter-{i:04d}.npz/.dat`.
"""
import time
from collections.abc import Iterable

import numpy as np

from .data_k import Data_K
from . import restart as restart_files
from .grid import GridTetra, exclude_equiv_points
from .path import Path
from .result import ResultDict, TABresult


def run(system, grid, calculators, adpt_num_iter=0, use_irred_kpt=True, symmetrize=True, fout_name="result",
        suffix="", parameters_K=None, file_Klist_path=None, restart=False, allow_restart=False, dump_results=False,
        restart_iteration=-1, Klist_part=10, parallel=True, print_Kpoints=False, adpt_mesh=2, adpt_fac=1,
        print_progress_step_time=5, print_progress_step_percent=1, data_k_class=None, k_batch=50,
        save_files=True, verbose=True, prefetch=True):
    """Integrate `calculators` (dict name -> calculator) over `grid`.

    With `use_irred_kpt` the K-list holds one K-block per star of the grid's
    point group (and refinement merges equivalent sub-blocks), which implies
    `symmetrize`: every K-block's result is averaged over the group
    (run_grid.py:225-236,258-265,382-386).  `parallel` is accepted for
    compatibility: one process drives one GPU.  On a Path (tabulation only) the K-blocks are
    batches of `k_batch` path points, each a Data_K on that k-list, and the tables come back in
    path order.  With `prefetch` (and a data_k_class that has
    `prefetch`) the next K-block's R -> k runs on a side stream while the current one is
    evaluated.  `dump_results` / `allow_restart` keep a record in `file_Klist_path` (default
    `_tmp_wb`) from which `restart=True` continues (restart_iteration: -1 = the last
    iteration written); the files are npz / npy arrays (restart.py), regular grids only.
    """
    if data_k_class is None:
        data_k_class = Data_K
    parameters_K = dict(parameters_K or {})
    on_path = isinstance(grid, Path)
    if on_path:
        # only tabulation along a path; no symmetry, refinement or restart (run_grid.py:217-229)
        for key, calc in calculators.items():
            if not calc.allow_path:
                raise ValueError(f"Calculation along a Path is running, but calculator `{key}` is not compatible "
                                 "with a Path")
        symmetrize = False
        adpt_num_iter = 0
        restart = dump_results = allow_restart = False
    else:
        if use_irred_kpt:
            symmetrize = True
        for key, calc in calculators.items():
            if not calc.allow_grid:
                raise ValueError(f"Calculation on Grid is running, but calculator `{key}` is not compatible with a Grid")
    if (restart or allow_restart or dump_results) and isinstance(grid, GridTetra):
        raise NotImplementedError("restart files are written for regular grids only, not for GridTetra")
    if dump_results:
        allow_restart = True
    if restart:
        allow_restart = True
    rpath = file_Klist_path if file_Klist_path is not None else "_tmp_wb"
    if getattr(data_k_class, "accepts_channels", False):
        need = set()
        for calc in calculators.values():
            if hasattr(calc, "needed_channels"):
                need |= set(calc.needed_channels(system))
        parameters_K.setdefault("channels", need)

    if restart:
        # run_grid.py:272-286: the K-points, their weights and results as the record left them
        K_list = restart_files.read_klist(rpath, grid, calculators)
        start_iter, factors = restart_files.read_factors(rpath, restart_iteration)
        factors = np.concatenate([factors, np.zeros(len(K_list) - len(factors))])
        for Kp, fac in zip(K_list, factors):
            Kp.set_factor(float(fac))
        result_all = None
        for Kp in K_list:
            if Kp.was_evaluated_flag:
                r = Kp.get_result() * Kp.factor
                result_all = r if result_all is None else result_all + r
        if verbose:
            print(f"run(): {len(K_list)} K-points read from {rpath}, iteration {start_iter}", flush=True)
    else:
        K_list = grid.get_K_list(use_symmetry=use_irred_kpt and not on_path, k_batch=k_batch if on_path else None)
        start_iter = 0
        result_all = None
        if allow_restart:
            restart_files.prepare(rpath)
            restart_files.write_factors(rpath, [Kp.factor for Kp in K_list], 0)
    store_results = allow_restart or adpt_num_iter != 0
    if adpt_num_iter < 0:
        adpt_num_iter = -adpt_num_iter * np.prod(grid.div) / np.prod(adpt_mesh) / adpt_fac / 3
    adpt_num_iter = int(round(adpt_num_iter))
    if (adpt_mesh is None) or np.max(adpt_mesh) <= 1:
        adpt_num_iter = 0
    else:
        if not isinstance(adpt_mesh, Iterable):
            adpt_mesh = [adpt_mesh] * 3
        adpt_mesh = np.array(adpt_mesh)

    factors = np.array([Kp.factor for Kp in K_list])
    _clear_zcache(system)   # the resident R -> k z-pass outputs live for one run
    t0 = time.time()
    counter = 0
    side = None
    if prefetch and hasattr(data_k_class, "prefetch"):
        import torch
        side = torch.cuda.Stream()   # default priority
    for i_iter in range(adpt_num_iter + 1):
        result_iter = None
        todo = [(ik, Kp) for ik, Kp in enumerate(K_list) if not Kp.was_evaluated_flag]
        nxt = None
        for i, (ik, Kp) in enumerate(todo):
            data = nxt if nxt is not None else data_k_class(system, dK=Kp.Kp_fullBZ, grid=grid, Kpoint=Kp,
                                                            **parameters_K)
            nxt = None
            if side is not None and i + 1 < len(todo):
                nKp = todo[i + 1][1]
                nxt = data_k_class(system, dK=nKp.Kp_fullBZ, grid=grid, Kpoint=nKp, **parameters_K)
                if i > 0:   # the first block sets up the system's stacks and plans on the main stream
                    nxt.prefetch(side)
            res = ResultDict({k: v(data) for k, v in calculators.items()})
            del data
            res = system.pointgroup.symmetrize(res) if symmetrize else res
            if store_results:
                Kp.set_result(res)
            else:
                Kp.was_evaluated_flag = True
            if allow_restart:
                restart_files.dump_result(rpath, ik, res)
                if dump_results:   # kept in the file only (Kpoint.py:56-61)
                    Kp.result = None
                    Kp.result_loader = (lambda ik=ik: restart_files.load_result(rpath, ik, calculators))
            contrib = res * Kp.factor
            result_iter = contrib if result_iter is None else result_iter + contrib
            counter += 1
        if result_all is None:
            result_all = result_iter
        else:
            factors_old = factors
            factors = np.array([kp.factor for kp in K_list])
            diff = factors[:len(factors_old)] - factors_old
            if result_iter is not None:
                result_all = result_all + result_iter
            for i, fac in enumerate(diff):
                if abs(fac) > 1e-8:
                    result_all = result_all + K_list[i].get_result() * fac
        if allow_restart:
            restart_files.write_klist(rpath, K_list)
            restart_files.write_factors(rpath, [Kp.factor for Kp in K_list], i_iter + start_iter)
        if save_files and not (restart and i_iter == 0):
            result_all.savedata(prefix=fout_name, suffix=suffix, i_iter=i_iter + start_iter)
        if i_iter >= adpt_num_iter:
            break
        Kmax = np.array([K.max for K in K_list]).T
        select_points = set().union(*(np.argsort(Km)[-adpt_fac:] for Km in Kmax))
        l1 = len(K_list)
        for iK in select_points:
            K_list += K_list[iK].divide(ndiv=adpt_mesh, periodic=system.periodic, use_symmetry=use_irred_kpt)
        if use_irred_kpt:
            exclude_equiv_points(K_list, new_points=len(K_list) - l1)
    _clear_zcache(system)
    if verbose:
        print(f"run(): processed {counter} K-blocks in {time.time() - t0:.3f} s", flush=True)
    # k-resolved results onto the full grid, or in path order (run_grid.py:394-401)
    for val in result_all.results.values():
        if isinstance(val, TABresult):
            if on_path:
                val.self_to_path(grid)
            else:
                val.self_to_grid(grid.dense)
    return result_all


def _clear_zcache(system):
    for dev in getattr(system, "_device_cache", {}).values():
        dev.zcache_clear()


def _klist_batches(sizes, batch_points):
    """Consecutive lists packed into batches of up to `batch_points` points (a longer list is a
    batch of its own)."""
    batches, cur, npts = [], [], 0
    for i, n in enumerate(sizes):
        if cur and npts + n > batch_points:
            batches.append(cur)
            cur, npts = [], 0
        cur.append(i)
        npts += int(n)
    batches.append(cur)
    return batches


def run_klists(system, k_lists, calculators, weights=None, parameters_K=None, batch_points=1 << 15,
               per_list=False):
    """Static calculators over many k-lists, batched into few large launches.

    Each list i (reduced coordinates [n_i, 3]) is one "K-point" in the sense
    of run(): a calculator sees it as a Data_K with k_list = list i (so
    results carry the calculator's 1/n_i), and the total is
    sum_i weights[i] * result_i, as run() sums result * Kp.factor
    (run_grid.py:97-99).  `weights` defaults to n_i / sum_j n_j, i.e. every
    point weighted 1 / (total number of points).

    Consecutive lists are packed into batches of up to `batch_points` points
    (a longer list is a batch of its own): R -> k, eigh and rotation run once
    per batch.  With default weights and per_list=False the calculators run
    once per batch (weight n_batch / n_total); otherwise once per list on a
    slice of the batch.  Returns the total ResultDict, and with per_list=True
    also the list of per-list ResultDicts.
    """
    parameters_K = dict(parameters_K or {})
    need = set()
    for calc in calculators.values():
        if hasattr(calc, "needed_channels"):
            need |= set(calc.needed_channels(system))
    parameters_K.setdefault("channels", need)
    k_lists = [np.asarray(k, dtype=float).reshape(-1, 3) for k in k_lists]
    sizes = np.array([len(k) for k in k_lists], dtype=np.int64)
    if len(k_lists) == 0 or sizes.min() <= 0:
        raise ValueError("run_klists needs at least one list and no empty list")
    ntot = int(sizes.sum())
    uniform = weights is None
    weights = sizes / ntot if uniform else np.asarray(weights, dtype=float).reshape(-1)
    if weights.shape[0] != len(k_lists):
        raise ValueError(f"{weights.shape[0]} weights for {len(k_lists)} k-lists")
    batches = _klist_batches(sizes, batch_points)
    total, each = None, [None] * len(k_lists)
    for b in batches:
        data = Data_K(system, k_list=np.concatenate([k_lists[i] for i in b]), **parameters_K)
        if uniform and not per_list:
            res = ResultDict({k: c(data) for k, c in calculators.items()}) * (data.nk / ntot)
            total = res if total is None else total + res
        else:
            off = 0
            for i in b:
                sub = data.subset(off, off + int(sizes[i]))
                off += int(sizes[i])
                res = ResultDict({k: c(sub) for k, c in calculators.items()})
                each[i] = res
                contrib = res * float(weights[i])
                total = contrib if total is None else total + contrib
        del data
    return (total, each) if per_list else total


class KlistGraph:
    """Uniformly weighted k-lists of several systems (any mix of nw) as one HIP graph.

    `jobs` is a list of (system, k_lists); each system's total is the run_klists
    total over its lists (every point weighted 1 / its system's points).  The
    constructor packs each system's lists into batches (run_klists), puts every
    batch's k-points on the device once, runs the whole step eagerly (device
    systems, channel stacks, workspaces and Fermi grids are created there) and
    captures a second pass into one graph on a side stream: R -> k, eigh,
    rotation, scan and finalisation of every batch of every system, with the
    k-sum.  `replay()` launches that graph and returns one ResultDict per system,
    whose device arrays are the graph's outputs (overwritten by the next replay;
    read or copy them before).  The library calls inside the graph allocate
    nothing and clear accumulators with fill kernels, so replays are
    independent.
    """

    def __init__(self, jobs, calculators, parameters_K=None, batch_points=1 << 15, device=0):
        import torch
        from . import _lib
        self.torch = torch
        self.calculators = calculators
        self.parameters_K = dict(parameters_K or {})
        self.ctx = _lib.get_context(device)
        self.batches = []   # per system: [(k_list, k_list_dev, weight)]
        for system, k_lists in jobs:
            k_lists = [np.asarray(k, dtype=float).reshape(-1, 3) for k in k_lists]
            sizes = np.array([len(k) for k in k_lists], dtype=np.int64)
            if len(k_lists) == 0 or sizes.min() <= 0:
                raise ValueError("KlistGraph needs at least one list per system and no empty list")
            ntot = int(sizes.sum())
            need = set()
            for calc in calculators.values():
                if hasattr(calc, "needed_channels"):
                    need |= set(calc.needed_channels(system))
            rows = []
            for b in _klist_batches(sizes, batch_points):
                kl = np.concatenate([k_lists[i] for i in b])
                rows.append((kl, torch.from_numpy(kl).to(self.ctx.device), len(kl) / ntot))
            self.batches.append((system, need, rows))
        self.npoints = sum(len(kl) for _, _, rows in self.batches for kl, _, _ in rows)
        self._step()                       # eager: set-up and first-use allocations
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.ctx.device)
        side.wait_stream(torch.cuda.current_stream(self.ctx.device))
        prev = self.ctx.set_stream(side)
        try:
            with torch.cuda.graph(self.graph, stream=side):
                self.results = self._step()
        finally:
            self.ctx.set_stream(prev)
        torch.cuda.current_stream(self.ctx.device).wait_stream(side)

    def _step(self):
        out = []
        for system, need, rows in self.batches:
            params = dict(self.parameters_K)
            params.setdefault("channels", need)
            total = None
            for kl, kl_dev, w in rows:
                data = Data_K(system, k_list=kl, k_list_dev=kl_dev, **params)
                res = ResultDict({k: c(data) for k, c in self.calculators.items()}) * w
                total = res if total is None else total + res
                del data
            out.append(total)
        return out

    def replay(self):
        """Launch the graph on the current stream; returns the per-system ResultDicts."""
        self.graph.replay()
        for rd in self.results:
            for r in rd.results.values():
                # host copies of the previous replay's outputs are stale now
                r._np = None
                r._np_smooth = None
        return self.results
