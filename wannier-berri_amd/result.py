"""Result containers: EnergyResult, ResultDict, VoidResult.

Interface of wannierberri/result/{energyresult,resultdict,result}.py
(`+`, `*` by a number, `mul_array`, `.max`, `savedata`, `from_npz`).  The
data of a K-block result stays on the GPU (a torch tensor) so the k-sum over
K-blocks never synchronises; `.data` materialises a NumPy array.
"""
import os

import numpy as np

from .smoother import VoidSmoother


def _is_tensor(x):
    return hasattr(x, "device") and hasattr(x, "cpu") and not isinstance(x, np.ndarray)


class Result:
    def __radd__(self, other):
        return self + other

    def __rmul__(self, other):
        return self * other


class VoidResult(Result):
    """Identically zero result (result/result.py:73-103)."""

    def __add__(self, other):
        return other

    def __mul__(self, other):
        return self

    def __sub__(self, other):
        return (-1) * other

    @property
    def max(self):
        return np.zeros(3)

    def savedata(self, name, prefix, suffix, i_iter):
        pass

    def transform(self, sym):
        return self


class EnergyResult(Result):
    """Data on a Fermi-level grid: data[iEF, 3, 3, ...] (result/energyresult.py:11-278).

    Parameters follow the reference; `data` may be a NumPy array or a torch tensor.
    """

    def __init__(self, Energies=None, data=None, smoothers=None, transformTR=None, transformInv=None,
                 rank=None, E_titles=("Efermi", "Omega"), comment="undocumented", save_mode="bin+txt"):
        if not isinstance(Energies, (list, tuple)):
            Energies = [Energies]
        self.Energies = [np.asarray(E) for E in Energies]
        self.N_energies = len(self.Energies)
        E_titles = list(E_titles) if isinstance(E_titles, (list, tuple)) else [E_titles]
        self.E_titles = (E_titles + ["???"] * self.N_energies)[:self.N_energies]
        self._data = data
        self._np = None
        nd = len(data.shape)
        self.rank = nd - self.N_energies if rank is None else rank
        for i in range(self.N_energies):
            assert self.Energies[i].shape[0] == data.shape[i], "Energy grid does not match the data"
        self.set_smoother(smoothers)
        self.transformTR = transformTR
        self.transformInv = transformInv
        self.comment = comment
        self.save_mode = set(s for s in ("bin", "txt") if s in save_mode) if isinstance(save_mode, str) else set(save_mode)

    def set_smoother(self, smoothers):
        """One smoother per energy axis, None meaning none (energyresult.py:116-122)."""
        if smoothers is None:
            smoothers = (None,) * self.N_energies
        if not isinstance(smoothers, (list, tuple)):
            smoothers = [smoothers]
        assert len(smoothers) == self.N_energies
        self.smoothers = [VoidSmoother() if sm is None else sm for sm in smoothers]

    # -- data access ----------------------------------------------------------
    @property
    def data(self):
        if self._np is None:
            d = self._data
            self._np = d.detach().cpu().numpy() if _is_tensor(d) else np.asarray(d)
        return self._np

    @data.setter
    def data(self, value):
        self._data = value
        self._np = None

    @property
    def raw(self):
        """The underlying array (torch tensor on the GPU, or NumPy)."""
        return self._data

    def _new(self, data):
        return EnergyResult(self.Energies, data, smoothers=self.smoothers, transformTR=self.transformTR,
                            transformInv=self.transformInv, rank=self.rank, E_titles=self.E_titles,
                            comment=self.comment, save_mode=self.save_mode)

    # -- algebra (energyresult.py:131-200) ------------------------------------
    def __mul__(self, number):
        if isinstance(number, (int, float, np.floating, np.integer)):
            return self._new(self._data * float(number))
        raise TypeError("result can only be multiplied by a number")

    def __truediv__(self, number):
        return self * (1. / number)

    def __add__(self, other):
        if other is None or isinstance(other, VoidResult) or (isinstance(other, (int, float)) and other == 0):
            return self
        for i in range(self.N_energies):
            if np.linalg.norm(self.Energies[i] - other.Energies[i]) > 1e-8:
                raise RuntimeError(f"Adding results with different energies {i} ({self.E_titles[i]}) - not allowed")
            if self.smoothers[i] != other.smoothers[i]:
                raise RuntimeError(f"Adding results with different smoothers [{i}]: {self.smoothers[i]} and "
                                   f"{other.smoothers[i]}")
        a, b = self._data, other._data
        if _is_tensor(a) != _is_tensor(b):
            a, b = self.data, other.data
        return self._new(a + b)

    def __sub__(self, other):
        return self + (-1) * other

    def add(self, other):
        self.data = self._data + other._data

    def mul_array(self, other, axes=None):
        other = np.asarray(other)
        if isinstance(axes, int):
            axes = (axes,)
        if axes is None:
            axes = tuple(range(other.ndim))
        nd = len(self._data.shape)
        reshape = tuple((self._data.shape[i] if i in axes else 1) for i in range(nd))
        o = other.reshape(reshape)
        if _is_tensor(self._data):
            import torch
            o = torch.as_tensor(o, dtype=self._data.dtype, device=self._data.device)
        return self._new(self._data * o)

    def swapaxes(self, a, b):
        d = self._data
        return self._new(d.transpose(a, b).contiguous() if _is_tensor(d) else d.swapaxes(a, b))

    def transform(self, sym):
        """Image under a point operation (energyresult.py:266-278): the last `rank` axes
        rotated, times the TR / inversion parity (transformTR, transformInv: +1 or -1)
        where the operation contains time reversal / inversion."""
        for flag, f, what in ((sym.TR, self.transformTR, "time reversal"), (sym.Inv, self.transformInv, "inversion")):
            if flag and f is None:
                raise ValueError(f"the parity of '{self.comment}' under {what} is not known")
        fTR = self.transformTR if self.transformTR is not None else 1
        fI = self.transformInv if self.transformInv is not None else 1
        return self._new(sym.transform_tensor(self._data, self.rank, fTR, fI))

    # -- refinement criteria (energyresult.py:250-264) ------------------------
    @property
    def dataSmooth(self):
        """The data smoothed along the first energy axis by its smoother (energyresult.py:124-129:
        the reference's loop over the axes leaves the first axis' smoothing of the raw data)."""
        if getattr(self, "_np_smooth", None) is None or self._np_smooth[0] is not self._data:
            sm = self.smoothers[0](self._data, axis=0)
            sm = sm.detach().cpu().numpy() if _is_tensor(sm) else np.asarray(sm)
            self._np_smooth = (self._data, sm)
        return self._np_smooth[1]

    @property
    def max(self):
        d = self.dataSmooth
        return np.array([np.abs(d).max(), np.linalg.norm(d), np.linalg.norm(d[1:] - d[:-1])])

    # -- output (energyresult.py:202-248) -------------------------------------
    def as_dict(self):
        d = dict(E_titles=np.array(self.E_titles), data=self.data, rank=self.rank, comment=self.comment)
        for i, E in enumerate(self.Energies):
            d[f"Energies_{i}"] = E
        return d

    def save(self, name):
        with open(name + ".npz", "wb") as f:
            np.savez_compressed(f, **self.as_dict())

    def savetxt(self, name):
        """One line per point of the energy grid(s): the energies, then the data and the
        smoothed data, complex values as `re im` (energyresult.py:202-225)."""
        data = self.data
        smooth = self.dataSmooth
        cpx = np.iscomplexobj(data)
        frmt = "{0:^31s}" if cpx else "{0:^15s}"
        head = "".join("#### " + s + "\n" for s in self.comment.split("\n"))
        labels = ["".join("xyz"[i] for i in c) for c in np.ndindex(*([3] * self.rank))] if self.rank > 0 else ["value"]
        head += "#" + "    ".join(f"{s:^15s}" for s in self.E_titles) + " " * 8 + \
            "    ".join(frmt.format(b) for b in labels * 2) + "\n"
        lines = []
        for idx in np.ndindex(*data.shape[:self.N_energies]):
            row = list(data[idx].reshape(-1)) + list(smooth[idx].reshape(-1))
            Es = "    ".join(f"{self.Energies[i][j]:15.6e}" for i, j in enumerate(idx))
            vals = "    ".join((f"{x.real:15.6e} {x.imag:15.6e}" if cpx else f"{x:15.6e}") for x in row)
            lines.append(Es + "    " + vals)
        with open(name, "w") as f:
            f.write(head + "\n".join(lines))

    def savedata(self, name, prefix, suffix, i_iter):
        suffix = "-" + suffix if len(suffix) > 0 else ""
        prefix = prefix + "-" if len(prefix) > 0 else ""
        filename = prefix + name + suffix + f"_iter-{i_iter:04d}"
        if "bin" in self.save_mode:
            self.save(filename)
        if "txt" in self.save_mode:
            self.savetxt(filename + ".dat")

    @classmethod
    def from_npz(cls, file_npz):
        with np.load(file_npz, allow_pickle=False) as f:
            E = [f[k] for k in sorted(f.files) if k.startswith("Energies_")]
            return cls(Energies=E, data=f["data"], rank=int(f["rank"]) if "rank" in f.files else None)


class KBandResult(Result):
    """Per-(k, band) data [nk, nband, 3, ...] (result/kbandresult.py): K-blocks concatenate
    along k; K-point factors do not apply."""

    def __init__(self, data, transformTR=None, transformInv=None):
        self.data = np.asarray(data)
        self.transformTR = transformTR
        self.transformInv = transformInv

    @property
    def nk(self):
        return self.data.shape[0]

    @property
    def nband(self):
        return self.data.shape[1]

    @property
    def rank(self):
        return self.data.ndim - 2

    def __add__(self, other):
        if other is None or (isinstance(other, (int, float)) and other == 0):
            return self
        return KBandResult(np.concatenate((self.data, other.data), axis=0), self.transformTR, self.transformInv)

    def __mul__(self, other):
        return KBandResult(self.data * other, self.transformTR, self.transformInv)

    def __truediv__(self, other):
        return KBandResult(self.data / other, self.transformTR, self.transformInv)

    def mul_array(self, arr, axes=0):
        """Multiply along the axis after k (the Fermi levels of a k-resolved static result)."""
        a = np.asarray(arr)
        return KBandResult(self.data * a.reshape((1, -1) + (1,) * (self.data.ndim - 2)), self.transformTR,
                           self.transformInv)

    def add(self, other):
        """Elementwise, in place (kbandresult.py:66-67)."""
        self.data = self.data + other.data

    def swapaxes(self, a, b):
        """Swap tensor axes a, b counted as for EnergyResult (0 = the Fermi-level axis)."""
        return KBandResult(self.data.swapaxes(a + 1, b + 1), self.transformTR, self.transformInv)

    def transform(self, sym):
        """Image under a point operation (kbandresult.py:150-159): the Cartesian axes rotated,
        times the TR / inversion parity."""
        for flag, f, what in ((sym.TR, self.transformTR, "time reversal"), (sym.Inv, self.transformInv, "inversion")):
            if flag and f is None and self.rank > 0:
                raise ValueError(f"the parity of a tabulated quantity under {what} is not known")
        fTR = self.transformTR if self.transformTR is not None else 1
        fI = self.transformInv if self.transformInv is not None else 1
        return KBandResult(sym.transform_tensor(self.data, self.rank, fTR, fI), self.transformTR, self.transformInv)

    def to_grid(self, index, npoints):
        """Segment mean: grid point i gets the mean of the rows whose `index` is i (rows with
        index -1 are dropped; a point no row lands on is NaN, the mean of nothing)."""
        keep = index >= 0
        idx = index[keep]
        rows = self.data[keep].reshape(idx.shape[0], -1)
        count = np.bincount(idx, minlength=npoints).astype(float)
        total = np.zeros((npoints, rows.shape[1]), dtype=rows.dtype)
        np.add.at(total, idx, rows)
        with np.errstate(invalid="ignore", divide="ignore"):
            mean = total / count[:, None]
        return KBandResult(mean.reshape((npoints,) + self.data.shape[1:]), self.transformTR, self.transformInv)

    def to_path(self, k_map):
        """Rows reordered (and repeated) by `k_map` (kbandresult.py:133-141)."""
        return KBandResult(self.data[np.asarray(k_map, dtype=int)], self.transformTR, self.transformInv)

    def get_component(self, component=None):
        """'x', 'xy', 'trace', ... of the Cartesian indices (kbandresult.py)."""
        if component is None or component == "":
            return self.data
        xyz = {"x": 0, "y": 1, "z": 2}
        if component == "trace":
            assert self.rank == 2
            return self.data[:, :, 0, 0] + self.data[:, :, 1, 1] + self.data[:, :, 2, 2]
        assert len(component) == self.rank
        return self.data[(slice(None), slice(None)) + tuple(xyz[c] for c in component)]


class TABresult(Result):
    """k-resolved results of TabulatorAll (result/tabresult.py:11-212): k-points and a
    KBandResult per quantity; K-blocks concatenate, and run() puts the k-points on the
    grid at the end (`self_to_grid`)."""

    def __init__(self, kpoints, recip_lattice, results=None, mode="grid", save_mode="bin"):
        self.results = results or {}
        self.nband = self.results["Energy"].nband
        self.mode = mode
        self.grid = None
        self.gridorder = None
        self.recip_lattice = recip_lattice
        self.kpoints = np.array(kpoints, dtype=float) % 1
        self.save_mode = save_mode
        for k, r in self.results.items():
            assert len(self.kpoints) == r.nk, f"{len(self.kpoints)} k-points but {r.nk} in '{k}'"

    @property
    def Enk(self):
        return self.results["Energy"]

    def __mul__(self, other):
        return self   # K-point factors do not apply to tabulated quantities

    __truediv__ = __mul__

    def __add__(self, other):
        if other is None or (isinstance(other, (int, float)) and other == 0):
            return self
        results = {r: self.results[r] + other.results[r] for r in self.results if r in other.results}
        return TABresult(np.vstack((self.kpoints, other.kpoints)), self.recip_lattice, results, self.mode,
                         self.save_mode)

    @property
    def max(self):
        return np.zeros(0)

    def transform(self, sym):
        """The k-points and every quantity mapped by the operation (tabresult.py:128-132); the
        symmetrised sum over the group then holds every image, and to_grid averages those
        that land on the same grid point."""
        kp = sym.transform_reduced_vector(self.kpoints, self.recip_lattice)
        return TABresult(kp, self.recip_lattice, {r: v.transform(sym) for r, v in self.results.items()}, self.mode,
                         self.save_mode)

    def savedata(self, name, prefix, suffix, i_iter):
        pass   # as the reference (tabresult.py:56-60)

    def to_grid(self, grid, order="C"):
        """Results on the full grid (the format of tabresult.py:134-170): grid point
        (i0, i1, i2) at k = i / grid, listed in `order`, holding the mean of the k-points
        (symmetry images included) that fall on it within 1e-5; other k-points are dropped."""
        grid = np.array(grid, dtype=int)
        npoints = int(np.prod(grid))
        pos = self.kpoints * grid[None, :]
        nearest = np.rint(pos)
        on_grid = np.all(np.abs(pos - nearest) < 1e-5 * grid[None, :], axis=1)
        flat = np.ravel_multi_index(tuple((nearest.astype(int) % grid[None, :]).T), tuple(grid), order=order)
        flat = np.where(on_grid, flat, -1)
        k_new = np.stack(np.unravel_index(np.arange(npoints), tuple(grid), order=order), axis=1) / grid[None, :]
        res = TABresult(k_new, self.recip_lattice, {r: v.to_grid(flat, npoints) for r, v in self.results.items()},
                        self.mode, self.save_mode)
        res.grid = grid
        res.gridorder = order
        return res

    def self_to_grid(self, grid):
        self.__dict__.update(self.to_grid(grid).__dict__)

    def self_to_path(self, path):
        """Rows put in the order of the path's k-points (tabresult.py:171-185): each path point
        takes the computed k-point nearest to it modulo reciprocal lattice vectors, which must
        agree within 1e-5."""
        kp = path.get_kpoints()
        diff = self.kpoints[:, None, :] - kp[None, :, :]
        diff -= np.round(diff)
        k_map = np.argmin(np.linalg.norm(diff, axis=2), axis=0)
        d = self.kpoints[k_map] - kp
        d -= np.round(d)
        assert np.allclose(d, 0, atol=1e-5), f"kpoints from path and result differ by {np.max(abs(d))}"
        self.results = {r: v.to_path(k_map) for r, v in self.results.items()}
        self.kpoints = kp.copy()

    def get_data(self, quantity, iband=None, component=None):
        """Data of one quantity on the grid [N0, N1, N2, (bands), ...] (tabresult.py:186-230)."""
        if iband is None:
            iband = np.arange(self.nband)
        r = self.Enk if quantity == "Energy" else self.results[quantity]
        d = r.get_component(component)[:, iband]
        if self.grid is None:
            return d
        return d.reshape(tuple(self.grid) + d.shape[1:])


class ResultDict(Result):
    """Dictionary of results (result/resultdict.py:18-66)."""

    def __init__(self, results, save_mode="bin+txt"):
        self.results = results

    def __mul__(self, number):
        return ResultDict({k: v * number for k, v in self.results.items()})

    def __truediv__(self, number):
        return ResultDict({k: v / number for k, v in self.results.items()})

    def __add__(self, other):
        if other is None or (isinstance(other, (int, float)) and other == 0):
            return self
        return ResultDict({k: self.results[k] + other.results[k] for k in self.results if k in other.results})

    def __sub__(self, other):
        return self + (-1) * other

    def savedata(self, prefix, suffix, i_iter):
        for k, v in self.results.items():
            v.savedata(k, prefix, suffix, i_iter)

    def transform(self, sym):
        return ResultDict({k: v.transform(sym) for k, v in self.results.items()})

    @property
    def max(self):
        return np.array([x for v in self.results.values() for x in v.max])
