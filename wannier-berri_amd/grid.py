"""K-grid: NK = NKdiv x NKFFT mixed decomposition and K-block objects.

Follows wannierberri/grid/grid.py:86-266 (Grid, determineNK, autoNK) and
grid/Kpoint.py:9-181 (KpointBZparallel, divide) for systems without
point-group symmetry (the identity group: every K-block is irreducible).
"""
from collections.abc import Iterable
import warnings

import numpy as np


def one2three(nk):
    if nk is None:
        return None
    if isinstance(nk, Iterable):
        nk = list(nk)
        assert len(nk) == 3
    else:
        nk = (nk,) * 3
    assert all(isinstance(n, (int, np.integer)) and n > 0 for n in nk)
    return np.array(nk, dtype=int)


def _iterate_vector(v1, v2):
    return ((x, y, z) for x in range(v1[0], v2[0]) for y in range(v1[1], v2[1]) for z in range(v1[2], v2[2]))


def autoNK(NK, NKFFTrec):
    """FFT grid between NKFFTrec and 2*NKFFTrec that best divides NK (grid.py:196-212)."""
    cand = np.array([f for f in _iterate_vector(NKFFTrec, NKFFTrec * 2)])
    NKdiv = np.array(np.round(NK[None, :] / cand), dtype=int)
    NKdiv[NKdiv <= 0] = 1
    change = NKdiv * cand / NK[None, :]
    sel = change > 1
    change[sel] = 1. / change[sel]
    FFT = cand[np.argmax(change.min(axis=1))]
    NKdiv = np.array(np.round(NK / FFT), dtype=int)
    NKdiv[NKdiv <= 0] = 1
    return NKdiv, FFT


def determineNK(periodic, NKdiv, NKFFT, NK, NKFFT_recommended, recip_lattice, length=None, length_FFT=None):
    """grid.py:215-266."""
    NKdiv, NKFFT, NK = one2three(NKdiv), one2three(NKFFT), one2three(NK)
    if length is not None and NK is None:
        NK = np.array(np.round(length / (2 * np.pi) * np.linalg.norm(recip_lattice, axis=1)), dtype=int)
    if length_FFT is not None and NKFFT is None:
        NKFFT = np.array(np.round(length_FFT / (2 * np.pi) * np.linalg.norm(recip_lattice, axis=1)), dtype=int)
    if (NKdiv is not None) and (NKFFT is not None):
        pass
    elif NK is not None:
        if NKFFT is not None:
            NKdiv = np.array(np.round(NK / NKFFT), dtype=int)
            NKdiv[NKdiv <= 0] = 1
        else:
            NKdiv, NKFFT = autoNK(NK, np.array(NKFFT_recommended))
    else:
        raise ValueError("you need to specify either NK or a pair (NKdiv,NKFFT) or (NK,NKFFT). "
                         f"found NK={NK}, NKdiv={NKdiv}, NKFFT={NKFFT} ")
    if NK is not None and not np.all(NK == NKFFT * NKdiv):
        warnings.warn(f" the requested k-grid {NK} was adjusted to {NKFFT * NKdiv}. ")
    notper = np.logical_not(periodic)
    NKdiv = np.array(NKdiv)
    NKFFT = np.array(NKFFT)
    NKdiv[notper] = 1
    NKFFT[notper] = 1
    return NKdiv, NKFFT


class KpointBZparallel:
    """A K-block: the NKFFT sub-grid shifted by K/NKFFT (Kpoint.py:9-181)."""

    def __init__(self, K=np.zeros(3), dK=np.ones(3), NKFFT=np.ones(3), factor=1., refinement_level=0):
        self.K = np.array(K, dtype=float)
        self.dK = np.array(dK, dtype=float)
        self.NKFFT = np.array(NKFFT, dtype=int)
        self.factor = factor
        self.refinement_level = refinement_level
        self.result = None
        self.was_evaluated_flag = False
        self._max = None

    @property
    def Kp_fullBZ(self):
        return self.K / self.NKFFT   # Kpoint.py:75-77

    @property
    def dK_fullBZ(self):
        return self.dK / self.NKFFT

    def set_result(self, res):
        self.result = res
        self._max = res.max if hasattr(res, "max") else None
        self.was_evaluated_flag = True

    def get_result(self):
        if self.result is None:
            raise RuntimeError("result for a K-point is called, which was never evaluated")
        return self.result

    def get_result_factor(self):
        return self.get_result() * self.factor

    def set_factor(self, factor):
        self.factor = factor

    @property
    def max(self):
        assert self.was_evaluated_flag, "max is called for a K-point which is not evaluated"
        return self._max * self.factor

    def divide(self, ndiv, periodic, use_symmetry=False):
        """Split into ndiv^3 sub-blocks (Kpoint.py:145-175)."""
        ndiv = np.array(ndiv, dtype=int)
        ndiv[np.logical_not(periodic)] = 1
        dK_adpt = self.dK / ndiv
        shift = (-self.dK + dK_adpt) / 2.
        newfac = self.factor / np.prod(ndiv)
        out = []
        for x in range(ndiv[0]):
            for y in range(ndiv[1]):
                for z in range(ndiv[2]):
                    out.append(KpointBZparallel(K=self.K + shift + dK_adpt * np.array([x, y, z]), dK=dK_adpt,
                                                NKFFT=self.NKFFT, factor=newfac,
                                                refinement_level=self.refinement_level + 1))
        self.set_factor(0)
        return out

    def __str__(self):
        return ("coord in rec.lattice = [ " + " , ".join(f"{x:10.6f}" for x in self.K) +
                f" ], refinement level:{self.refinement_level}, factor = {self.factor}")


class Grid:
    """k-grid NK = NKdiv * NKFFT (grid/grid.py:86-189).

    Exactly one of `system`, `real_lattice`, `recip_lattice` gives the lattice.
    """

    def __init__(self, system=None, real_lattice=None, recip_lattice=None, length=None, NKdiv=None,
                 NKFFT=None, NK=None, length_FFT=None, use_symmetry=False, pointgroup=None):
        found = [p is not None for p in (system, real_lattice, recip_lattice)]
        assert sum(found) == 1, "Exactly one of system, real_lattice, recip_lattice should be provided"
        if system is not None:
            real_lattice = system.real_lattice
            periodic = system.periodic
            NKFFT_rec = np.array(system.NKFFT_recommended)
        else:
            if real_lattice is None:
                real_lattice = 2 * np.pi * np.linalg.inv(np.array(recip_lattice)).T
            periodic = np.ones(3, dtype=bool)
            NKFFT_rec = np.ones(3, dtype=int)
        self.real_lattice = np.array(real_lattice)
        self.recip_lattice = 2 * np.pi * np.linalg.inv(self.real_lattice).T
        self.div, self.FFT = determineNK(periodic, NKdiv, NKFFT, NK, NKFFT_rec, self.recip_lattice,
                                         length=length, length_FFT=length_FFT)

    @property
    def dense(self):
        return self.div * self.FFT

    @property
    def str_short(self):
        return f"Grid() with NKdiv={self.div}, NKFFT={self.FFT}, NKtot={self.dense}"

    @property
    def points_FFT(self):
        N = self.FFT
        return np.array([[i / N[0], j / N[1], l / N[2]] for i in range(N[0]) for j in range(N[1])
                         for l in range(N[2])])

    def get_K_list(self, use_symmetry=False, k_batch=None):
        dK = 1. / self.div
        factor = 1. / np.prod(self.div)
        return [KpointBZparallel(K=np.array([x, y, z]) * dK, dK=dK, NKFFT=self.FFT, factor=factor, refinement_level=0)
                for x in range(self.div[0]) for y in range(self.div[1]) for z in range(self.div[2])]
