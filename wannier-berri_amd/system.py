"""System_R: an R-space Wannier / tight-binding model for the MI355X path.

Keeps the reference's input semantics (wannierberri/system/system_R.py):
matrices `_XX_R[key]` of shape `[nR, nw, nw, 3^m]`, `rvec.iRvec`, Wannier
centres used as the derivative shifts (system_R.py:776-780), `periodic`,
`force_internal_terms_only`, and the directory-of-npz layout of
`System_R.from_npz` (system_R.py:713-799, legacy `[m, n, R]` transposed at
795-796).  Device copies of the R-space data are made once per system and
device and cached here.
"""
import glob
import warnings
import os

import numpy as np

from .symmetry import PointGroup


class _Rvec:
    """Minimal `Rvectors` view: iRvec, nRvec, iR0 (fourier/rvectors.py:385-434)."""

    def __init__(self, iRvec):
        self.iRvec = np.array(iRvec, dtype=int)

    @property
    def nRvec(self):
        return self.iRvec.shape[0]

    @property
    def iR0(self):
        return self.iRvec.tolist().index([0, 0, 0])

    def iR(self, R):
        return self.iRvec.tolist().index(np.array(np.round(R), dtype=int).tolist())


class System_R:
    """R-space model (system/system_R.py:18-175)."""

    def __init__(self, real_lattice=None, iRvec=None, wannier_centers_cart=None, matrices=None,
                 periodic=(True, True, True), force_internal_terms_only=False, name="wberri", NKFFT=None):
        self.name = name
        self._XX_R = {}
        self.force_internal_terms_only = force_internal_terms_only
        self.periodic = np.zeros(3, dtype=bool)
        per = np.array(periodic, dtype=bool).ravel()
        self.periodic[:len(per)] = per
        self.is_phonon = False
        if NKFFT is not None:
            self._NKFFT_recommended = np.array(NKFFT)
        self._device_cache = {}
        if real_lattice is not None:
            self.real_lattice = np.array(real_lattice, dtype=float)
            self.pointgroup = PointGroup([], real_lattice=self.real_lattice)
        if iRvec is not None:
            self.rvec = _Rvec(iRvec)
        if matrices:
            for k, v in matrices.items():
                self.set_R_mat(k, v)
        if wannier_centers_cart is not None:
            self.wannier_centers_cart = np.array(wannier_centers_cart, dtype=float)
        elif self.has_R_mat("Ham"):
            self.wannier_centers_cart = np.zeros((self.num_wann, 3))

    def set_pointgroup(self, symmetry_gen=(), pointgroup=None, spacegroup=None):
        """Point group used for irreducible K-points and symmetrisation (system/system.py:140-170).

        `symmetry_gen`: generators (symmetry.PointSymmetry objects or names
        such as 'C3z', 'TimeReversal*C2x'); the group is closed under products.
        """
        if spacegroup is not None:
            raise NotImplementedError("point groups from a space group (irrep) are not available")
        if pointgroup is not None:
            assert np.allclose(pointgroup.recip_lattice, self.recip_lattice), \
                "the provided pointgroup has a different recip_lattice than the system"
            self.pointgroup = pointgroup
        else:
            self.pointgroup = PointGroup(symmetry_gen, real_lattice=self.real_lattice)

    def set_symmetry(self, symmetry_gen=(), pointgroup=None, spacegroup=None):
        """Deprecated name of set_pointgroup (system/system.py:130-137)."""
        self.set_pointgroup(symmetry_gen=symmetry_gen, pointgroup=pointgroup, spacegroup=spacegroup)

    # -- matrices -------------------------------------------------------------
    def get_R_mat(self, key):
        try:
            return self._XX_R[key]
        except KeyError:
            raise ValueError(f"The real-space matrix elements '{key}' are not set in the system, but are "
                             "required for the current calculation")

    def has_R_mat(self, key):
        return key in self._XX_R

    def set_R_mat(self, key, value, reset=False, Hermitian=False):
        """Set a full `[nR, nw, nw, ...]` matrix (system_R.py:117-175, R=None branch)."""
        value = np.array(value, dtype=complex)
        if hasattr(self, "rvec"):
            assert value.shape[0] == self.rvec.nRvec, f"the 0th dimension should be nR={self.rvec.nRvec}"
        if Hermitian:
            value = 0.5 * (value + self.conj_XX_R(value))
        if key in self._XX_R and not reset:
            raise RuntimeError(f"setting {key} for the second time without explicit permission")
        self._XX_R[key] = value
        self._device_cache.clear()
        self._herm_err = {}

    def set_spin_pairs(self, pairs):
        """SS_R for Wannier functions that are Sz eigenstates in pairs (up, down)
        (system/system_R.py:500-531): at R = 0, the Pauli matrices on each pair's 2x2 block,
        zero elsewhere (states in no pair carry no spin)."""
        pairs = [tuple(int(x) for x in p) for p in pairs]
        assert all(len(p) == 2 for p in pairs)
        states = [x for p in pairs for x in p]
        assert all(0 <= x < self.num_wann for x in states), "indices of states should be 0 <= i < num_wann"
        assert len(set(states)) == len(states), "some states appear more than once in pairs"
        if len(pairs) < self.num_wann / 2:
            warnings.warn(f"number of spin pairs {len(pairs)} is less than num_wann/2 = {self.num_wann / 2}; "
                          "the other states get no spin")
        # (sigma_x, sigma_y, sigma_z)[a, b] of the pair's (up, down) = (0, 1) indices
        pauli = np.array([[[0, 0, 1], [1, -1j, 0]], [[1, 1j, 0], [0, 0, -1]]], dtype=complex)
        S0 = np.zeros((self.num_wann, self.num_wann, 3), dtype=complex)
        for i, j in pairs:
            if np.linalg.norm(self.wannier_centers_cart[i] - self.wannier_centers_cart[j]) > 1e-3:
                warnings.warn(f"setting a spin pair for Wannier functions {i} and {j} with different centres")
            S0[i, i], S0[i, j], S0[j, i], S0[j, j] = pauli[0, 0], pauli[0, 1], pauli[1, 0], pauli[1, 1]
        SS = np.zeros((self.rvec.nRvec, self.num_wann, self.num_wann, 3), dtype=complex)
        SS[self.rvec.iR0] = S0
        self.set_R_mat("SS", SS, reset=True)

    def set_spin_interlaced(self):
        """set_spin_pairs with the spins interlaced, pairs (2i, 2i + 1) (system/system_R.py:533-545)."""
        assert self.num_wann % 2 == 0, f"odd number of Wannier functions {self.num_wann} cannot be grouped into spin pairs"
        self.set_spin_pairs([(2 * i, 2 * i + 1) for i in range(self.num_wann // 2)])

    def hermiticity_error(self, key="Ham"):
        """max |X(R) - X(-R)^+| / max |X|, an R-vector without its -R partner counting
        with its whole X(R) (the pairing of fourier/rvectors.py:424-430)."""
        cache = self.__dict__.setdefault("_herm_err", {})
        if key not in cache:
            X = self.get_R_mat(key)
            scale = np.abs(X).max()
            cache[key] = float(np.abs(X - self.conj_XX_R(X)).max() / scale) if scale > 0 else 0.
        return cache[key]

    def conj_XX_R(self, X):
        """X(-R)^dagger (fourier/rvectors.py:424-430)."""
        index = {tuple(R): i for i, R in enumerate(self.rvec.iRvec)}
        out = np.zeros_like(X)
        for i, R in enumerate(self.rvec.iRvec):
            j = index.get(tuple(-R))
            if j is not None:
                out[i] = X[j]
        return out.swapaxes(1, 2).conj()

    # -- geometry -------------------------------------------------------------
    @property
    def num_wann(self):
        return self._XX_R["Ham"].shape[1]

    @property
    def nRvec(self):
        return self.rvec.nRvec

    @property
    def cell_volume(self):
        return abs(np.linalg.det(self.real_lattice))   # system/system.py:173-181

    @property
    def recip_lattice(self):
        return 2 * np.pi * np.linalg.inv(self.real_lattice).T

    @property
    def wannier_centers_red(self):
        return self.wannier_centers_cart.dot(np.linalg.inv(self.real_lattice))   # system_R.py:608-610

    @property
    def cRvec_shifted(self):
        """R + t_n - t_m for element (m, n), shape [nR, nw, nw, 3] (rvectors.py:351-369)."""
        shifts = self.wannier_centers_red.dot(self.real_lattice)
        cR = self.rvec.iRvec.dot(self.real_lattice)
        return cR[:, None, None, :] + (-shifts[:, None, :] + shifts[None, :, :])[None]

    # -- Wigner-Seitz (MDRS) fold ----------------------------------------------
    def do_ws_dist(self, mp_grid, wannier_centers_cart=None, ws_dist_tol=1e-5):
        """Minimal-distance replica selection (system_R.py:572-588, rvectors.py:57-100,126-200,570-609).

        Every matrix is first folded onto the mp_grid cell (X_grid[R mod mp] = sum of X_R).
        Then, for the element (m, n) and each cell vector R0, the translates R0 + mp * T
        (T in {-3..3}^3) that minimise |R + t_n - t_m| (Cartesian; ties within ws_dist_tol)
        each get X_grid[R0][m, n] / (number of ties).  The new R-list is the union over all
        element shifts, R-vectors whose matrices all vanish (< 1e-8) are dropped, and
        NKFFT_recommended becomes mp_grid.  A negative ws_dist_tol is used as |ws_dist_tol|,
        with the shifts compared to 8 decimals (rvectors.py:78-86).
        """
        mp = np.array(mp_grid, dtype=int).reshape(-1)
        if mp.size == 1:
            mp = np.repeat(mp, 3)
        if mp.size != 3 or np.any(mp <= 0):
            raise ValueError(f"mp_grid should be one integer, or three integers. found {mp_grid}")
        wcc = self.wannier_centers_cart if wannier_centers_cart is None else np.asarray(wannier_centers_cart, float)
        tol = abs(ws_dist_tol)
        digits = 8 if ws_dist_tol < 0 else int(np.ceil(-np.log10(ws_dist_tol))) + 1
        L = self.real_lattice
        t_red = wcc.dot(np.linalg.inv(L))
        nw = t_red.shape[0]
        # element (m, n) needs R + t_n - t_m; elements with equal shifts share one selection
        shift = np.round(-t_red[:, None, :] + t_red[None, :, :], digits).reshape(-1, 3)
        shifts, shift_of = np.unique(shift, axis=0, return_inverse=True)
        shift_of = shift_of.reshape(nw, nw)
        cell = np.stack(np.meshgrid(*[np.arange(n) for n in mp], indexing="ij"), axis=-1).reshape(-1, 3)
        T = np.stack(np.meshgrid(*[np.arange(-3, 4)] * 3, indexing="ij"), axis=-1).reshape(-1, 3)
        cand = cell[:, None, :] + (T * mp)[None, :, :]            # [ncell, 343, 3]
        cand_cart = cand.dot(L)
        sel = []                                                    # per shift: (R [n, 3], cell index, 1/ndeg)
        for s in shifts:
            dist = np.linalg.norm(cand_cart + s.dot(L), axis=2)
            near = np.abs(dist - dist.min(axis=1, keepdims=True)) < tol
            ic, it = np.nonzero(near)
            nd = near.sum(axis=1)
            sel.append((cand[ic, it], ic, 1.0 / nd[ic]))
        allR = np.unique(np.concatenate([r for r, _, _ in sel]), axis=0)
        index = {tuple(r): i for i, r in enumerate(allR)}
        sel = [(np.array([index[tuple(r)] for r in R]), ic, w) for R, ic, w in sel]
        ncell = cell.shape[0]
        cell_of_R = np.ravel_multi_index(tuple((self.rvec.iRvec % mp).T), tuple(mp))
        new = {}
        for key, X in self._XX_R.items():
            grid = np.zeros((ncell,) + X.shape[1:], dtype=X.dtype)
            np.add.at(grid, cell_of_R, X)
            Y = np.zeros((allR.shape[0],) + X.shape[1:], dtype=X.dtype)
            for isft, (iR, ic, w) in enumerate(sel):
                m_idx, n_idx = np.nonzero(shift_of == isft)
                wx = w.reshape((-1,) + (1,) * (X.ndim - 3))
                for m, n in zip(m_idx, n_idx):
                    Y[iR, m, n] += grid[ic, m, n] * wx
            new[key] = Y
        keep = np.zeros(allR.shape[0], dtype=bool)
        for Y in new.values():
            keep |= np.abs(Y.reshape(Y.shape[0], -1)).max(axis=1) > 1e-8
        self.rvec = _Rvec(allR[keep])
        self._XX_R = {k: Y[keep] for k, Y in new.items()}
        self._device_cache.clear()
        self._herm_err = {}
        self._NKFFT_recommended = mp

    @property
    def NKFFT_recommended(self):
        """Smallest FFT grid on which the R-vectors do not overlap (rvectors.py:280-292)."""
        if hasattr(self, "_NKFFT_recommended"):
            return self._NKFFT_recommended
        nk = np.ones(3, dtype=int)
        for i in range(3):
            R = self.rvec.iRvec[:, i]
            if (R > 0).any():
                nk[i] += R.max()
            if (R < 0).any():
                nk[i] -= R.min()
        return nk

    # -- loaders --------------------------------------------------------------
    @classmethod
    def from_npz(cls, path, legacy=False, matrices=None, force_internal_terms_only=False):
        """Directory of `.npz` files as written by System_R.to_npz (system_R.py:713-799).

        Matrix files are `_XX_R_<key>.npz` (or `<key>.npz` in older fixtures).
        """
        if not os.path.exists(path):
            raise FileNotFoundError(f"directory {path} does not exist")

        def arr(name):
            with np.load(os.path.join(path, name + ".npz"), allow_pickle=False) as f:
                return f["arr_0"]
        real_lattice = arr("real_lattice")
        iRvec = arr("iRvec")
        if os.path.exists(os.path.join(path, "wannier_centers_cart.npz")):
            wcc = arr("wannier_centers_cart")
        elif os.path.exists(os.path.join(path, "wannier_centers_red.npz")):
            wcc = arr("wannier_centers_red").dot(real_lattice)
        else:
            wcc = None
        periodic = arr("periodic") if os.path.exists(os.path.join(path, "periodic.npz")) else (True, True, True)
        if matrices is None:
            names = [os.path.splitext(os.path.basename(f))[0] for f in glob.glob(os.path.join(path, "*.npz"))]
            matrices = sorted(set(n[6:] for n in names if n.startswith("_XX_R_")) |
                              set(n for n in names if n in ("Ham", "AA", "BB", "CC", "OO", "SS")))
        mats = {}
        for key in matrices:
            fn = "_XX_R_" + key if os.path.exists(os.path.join(path, "_XX_R_" + key + ".npz")) else key
            a = arr(fn)
            if legacy:
                a = np.transpose(a, (2, 0, 1) + tuple(range(3, a.ndim)))
            mats[key] = a
        system = cls(real_lattice, iRvec, wcc, mats, periodic=periodic,
                     force_internal_terms_only=force_internal_terms_only)
        pg = os.path.join(path, "pointgroup.npz")
        if os.path.exists(pg):
            with np.load(pg, allow_pickle=False) as f:
                system.set_pointgroup(pointgroup=PointGroup.from_dict({k: f[k] for k in f.files}))
        return system

    def to_npz(self, path, R_matrices=None, overwrite=True):
        """Write the directory layout that System_R.from_npz reads (system_R.py:658-711):
        real_lattice, iRvec, wannier_centers_cart, periodic, pointgroup and
        compressed `_XX_R_<key>.npz` matrices `[nR, nw, nw, ...]`."""
        os.makedirs(path, exist_ok=overwrite)
        for key, val in (("real_lattice", self.real_lattice), ("iRvec", self.rvec.iRvec),
                         ("wannier_centers_cart", self.wannier_centers_cart), ("periodic", self.periodic)):
            np.savez(os.path.join(path, key + ".npz"), val)
        np.savez(os.path.join(path, "pointgroup.npz"), **self.pointgroup.as_dict())
        for key in (self._XX_R if R_matrices is None else R_matrices):
            np.savez_compressed(os.path.join(path, "_XX_R_" + key + ".npz"), self.get_R_mat(key))

    @classmethod
    def from_arrays_npz(cls, file_npz):
        """Single `.npz` with keys real_lattice, iRvec, wannier_centers_cart, Ham, AA, ... (`[R, m, n, ...]`)."""
        with np.load(file_npz, allow_pickle=False) as f:
            mats = {k: f[k] for k in f.files if k in ("Ham", "AA", "BB", "CC", "OO", "SS", "SA", "SHA", "SR", "SH", "SHR")}
            periodic = f["periodic"] if "periodic" in f.files else (True, True, True)
            fito = bool(f["force_internal_terms_only"]) if "force_internal_terms_only" in f.files else False
            return cls(f["real_lattice"], f["iRvec"], f["wannier_centers_cart"], mats, periodic=periodic,
                       force_internal_terms_only=fito)

    def to_arrays_npz(self, file_npz):
        d = dict(real_lattice=self.real_lattice, iRvec=self.rvec.iRvec.astype(np.int32),
                 wannier_centers_cart=self.wannier_centers_cart, periodic=self.periodic,
                 force_internal_terms_only=self.force_internal_terms_only)
        d.update(self._XX_R)
        np.savez_compressed(file_npz, **d)
