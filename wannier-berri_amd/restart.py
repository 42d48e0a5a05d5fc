"""Restart files of run() (run_grid.py:242-348,416-440, grid/Kpoint.py:40-67), in npz / npy only.

Directory `file_Klist_path` (default `_tmp_wb`):

* `factors_iter-XXXXXXXX.npy`: the K-points' weights after each iteration (the
  reference's own file format, run_grid.py:420-423);
* `K_list.npz`: the K-points (K, dK, NKFFT, refinement level), whether each was
  evaluated and its refinement metric (`max`), rewritten after every iteration;
* `_Kp-<ik>.npz`: the (symmetrised) result of K-point ik, one group of keys per
  calculator: EnergyResult `<name>/Energies_0`, `<name>/data`, `<name>/rank`;
  TABresult `<name>/tab/kpoints`, `<name>/tab/recip_lattice`, `<name>/tab/mode`,
  `<name>/tab/save_mode` and per quantity q `<name>/tab/q/<q>/data`, `.../tr`, `.../inv`
  (the KBandResult's time-reversal and inversion parities, NaN for none).

The reference pickles its K-points and results; these files hold plain arrays
only, so this build neither reads nor writes the reference's restart files.
Only regular grids (KpointBZparallel) and EnergyResult / TABresult results are supported.
"""
import glob
import os

import numpy as np

from .grid import KpointBZparallel
from .result import EnergyResult, KBandResult, ResultDict, TABresult

KLIST = "K_list.npz"


def prepare(path):
    """Start a new record: create `path` and remove the files of an earlier one."""
    os.makedirs(path, exist_ok=True)
    for pattern in (KLIST, "factors_iter-*.npy", "_Kp-*.npz"):
        for f in glob.glob(os.path.join(path, pattern)):
            os.remove(f)


def write_factors(path, factors, it):
    np.save(os.path.join(path, f"factors_iter-{it:08d}.npy"), np.asarray(factors, dtype=float))


def read_factors(path, it=-1):
    """(iteration, factors) of iteration `it`; negative `it` counts from the last one written
    (run_grid.py:426-440)."""
    if it >= 0:
        return it, np.load(os.path.join(path, f"factors_iter-{it:08d}.npy"), allow_pickle=False)
    files = glob.glob(os.path.join(path, "factors_iter-*.npy"))
    if not files:
        raise FileNotFoundError(f"no factors_iter-*.npy in {path}")
    its = np.array(sorted(int(os.path.basename(f)[len("factors_iter-"):-len(".npy")]) for f in files))
    want = max(its[-1] + it + 1, 0)
    return read_factors(path, int(its[its <= want][-1]) if want not in its else int(want))


def result_file(path, ik):
    return os.path.join(path, f"_Kp-{ik}.npz")


def dump_result(path, ik, res):
    d = {}
    for name, r in res.results.items():
        if isinstance(r, TABresult):
            t = f"{name}/tab/"
            d[t + "kpoints"], d[t + "recip_lattice"] = r.kpoints, np.asarray(r.recip_lattice, dtype=float)
            d[t + "mode"], d[t + "save_mode"] = np.array(r.mode), np.array(r.save_mode)
            for q, kb in r.results.items():
                d[f"{t}q/{q}/data"] = kb.data
                d[f"{t}q/{q}/tr"] = np.array(np.nan if kb.transformTR is None else kb.transformTR, dtype=float)
                d[f"{t}q/{q}/inv"] = np.array(np.nan if kb.transformInv is None else kb.transformInv, dtype=float)
            continue
        if not isinstance(r, EnergyResult):
            raise NotImplementedError(f"restart files hold EnergyResult and TABresult results only, not "
                                      f"{type(r).__name__} ('{name}')")
        for key, val in r.as_dict().items():
            if key != "comment":
                d[f"{name}/{key}"] = val
    np.savez_compressed(result_file(path, ik), **d)


def _parity(x):
    x = float(x)
    return None if np.isnan(x) else int(x)


def _load_tab(f, name):
    t = f"{name}/tab/"
    qs = sorted({k[len(t) + 2:].rsplit("/", 1)[0] for k in f.files if k.startswith(t + "q/")})
    results = {q: KBandResult(f[f"{t}q/{q}/data"], transformTR=_parity(f[f"{t}q/{q}/tr"]),
                              transformInv=_parity(f[f"{t}q/{q}/inv"])) for q in qs}
    return TABresult(f[t + "kpoints"], f[t + "recip_lattice"], results=results, mode=str(f[t + "mode"]),
                     save_mode=str(f[t + "save_mode"]))


def load_result(path, ik, calculators):
    """The ResultDict of K-point ik, with each calculator's smoother and parities."""
    out = {}
    with np.load(result_file(path, ik), allow_pickle=False) as f:
        for name, calc in calculators.items():
            if f"{name}/tab/kpoints" in f.files:
                out[name] = _load_tab(f, name)
                continue
            keys = [k for k in f.files if k.startswith(name + "/")]
            E = [f[k] for k in sorted(keys) if k.split("/", 1)[1].startswith("Energies_")]
            rank = int(f[name + "/rank"])
            tr, inv = getattr(calc, "parity", (None, None))
            out[name] = EnergyResult(Energies=E, data=f[name + "/data"], rank=rank,
                                     smoothers=[getattr(calc, "smoother", None)], transformTR=tr, transformInv=inv)
    return ResultDict(out)


def write_klist(path, K_list):
    if any(type(Kp) is not KpointBZparallel for Kp in K_list):
        raise NotImplementedError("restart files are written for regular grids (KpointBZparallel) only")
    maxes = [np.asarray(Kp._max, dtype=float).reshape(-1) if Kp.was_evaluated_flag and Kp._max is not None
             else None for Kp in K_list]
    width = max([len(m) for m in maxes if m is not None], default=0)
    mx = np.full((len(K_list), width), np.nan)
    for i, m in enumerate(maxes):
        if m is not None:
            mx[i, :len(m)] = m
    np.savez_compressed(os.path.join(path, KLIST),
                        K=np.array([Kp.K for Kp in K_list]), dK=np.array([Kp.dK for Kp in K_list]),
                        NKFFT=np.array([Kp.NKFFT for Kp in K_list]),
                        refinement_level=np.array([Kp.refinement_level for Kp in K_list]),
                        evaluated=np.array([Kp.was_evaluated_flag for Kp in K_list]), max=mx)


def read_klist(path, grid, calculators):
    """K-points of the record, their results loaded from the files on demand."""
    with np.load(os.path.join(path, KLIST), allow_pickle=False) as f:
        K, dK, NKFFT, lev, ev, mx = (f[k] for k in ("K", "dK", "NKFFT", "refinement_level", "evaluated", "max"))
    out = []
    for i in range(len(K)):
        Kp = KpointBZparallel(K=K[i], dK=dK[i], NKFFT=NKFFT[i], factor=0.0, pointgroup=grid.pointgroup,
                              refinement_level=int(lev[i]))
        if ev[i]:
            Kp.was_evaluated_flag = True
            Kp._max = mx[i][~np.isnan(mx[i])]
            Kp.result_loader = (lambda ik=i: load_result(path, ik, calculators))
        out.append(Kp)
    return out
