"""Smoothers of results on an energy grid (wannierberri/smoother.py:9-175).

A smoother convolves data along one energy axis with a broadening kernel
sampled on the grid step, cut at +-maxdE * smear and renormalised where the
window runs past either end of the grid.  Here the convolution is written as
one [NE, NE] banded matrix built once per smoother, applied with a single
matrix product (on the GPU when the data is a torch tensor there), instead of
a loop over energies.
"""
import numpy as np
from scipy.constants import Boltzmann, elementary_charge


class AbstractSmoother:
    """Common part (smoother.py:9-73): `smear` in eV, window of NE1 = int(maxdE * smear / dE)
    grid steps on each side, kernel values broaden(j dE) * dE."""
    _params = ("smear", "E", "maxdE", "NE1")

    def __init__(self, E, smear, maxdE):
        self.smear = smear
        self.E = np.array(E, dtype=float)
        self.maxdE = maxdE
        self.dE = self.E[1] - self.E[0]
        self.Emin, self.Emax = self.E[0], self.E[-1]
        self.NE = self.E.shape[0]
        self.NE1 = int(self.maxdE * self.smear / self.dE)
        self.smt = self._broaden(np.arange(-self.NE1, self.NE1 + 1) * self.dE) * self.dE
        self._mat = None
        self._dev = {}

    def _broaden(self, E):
        raise NotImplementedError

    @property
    def matrix(self):
        """M[i, j] = smt[j - i + NE1] / (sum of smt over the part of the window inside the grid)."""
        if self._mat is None:
            i = np.arange(self.NE)[:, None]
            off = np.arange(self.NE)[None, :] - i
            inside = np.abs(off) <= self.NE1
            M = np.where(inside, self.smt[np.clip(off + self.NE1, 0, 2 * self.NE1)], 0.)
            self._mat = M / M.sum(axis=1, keepdims=True)
        return self._mat

    def __call__(self, A, axis=0):
        """A smoothed along `axis` (NumPy array or torch tensor; the same kind is returned)."""
        assert A.shape[axis] == self.NE
        if hasattr(A, "device") and not isinstance(A, np.ndarray):
            import torch
            key = (A.device, A.dtype)
            if key not in self._dev:
                self._dev[key] = torch.as_tensor(self.matrix, dtype=A.dtype, device=A.device)
            B = A.movedim(axis, 0)
            out = (self._dev[key] @ B.reshape(self.NE, -1)).reshape(B.shape)
            return out.movedim(0, axis).contiguous()
        B = np.moveaxis(np.asarray(A), axis, 0)
        out = (self.matrix @ B.reshape(self.NE, -1)).reshape(B.shape)
        return np.moveaxis(out, 0, axis)

    def __eq__(self, other):
        if type(self) is not type(other):
            return False
        return all(np.allclose(getattr(self, p), getattr(other, p)) for p in self._params)

    def __str__(self):
        return f"<{type(self).__name__} smear={self.smear}, NE={self.NE}, NE1={self.NE1}, " \
               f"E={self.Emin}..{self.Emax}, step {self.dE}>"


class FermiDiracSmoother(AbstractSmoother):
    """-df/dE of the Fermi-Dirac function at temperature T_Kelvin (smoother.py:76-103)."""

    def __init__(self, E, T_Kelvin, maxdE=8):
        self.T_Kelvin = T_Kelvin
        super().__init__(E, T_Kelvin * Boltzmann / elementary_charge, maxdE)

    def _broaden(self, E):
        return 0.25 / self.smear / np.cosh(E / (2 * self.smear)) ** 2


class GaussianSmoother(AbstractSmoother):
    """Normalised Gaussian of width `smear` eV (smoother.py:106-127)."""

    def __init__(self, E, smear, maxdE=8):
        super().__init__(E, smear, maxdE)

    def _broaden(self, E):
        return np.exp(-(E / self.smear) ** 2) / self.smear / np.sqrt(np.pi)


class VoidSmoother(AbstractSmoother):
    """The identity (smoother.py:130-141)."""
    _params = ()

    def __init__(self):
        pass

    def __call__(self, A, axis=0):
        return A

    def __str__(self):
        return "<VoidSmoother>"


def get_smoother(energy, smear, mode=None):
    """Smoother for an evenly spaced energy grid (smoother.py:144-175): Fermi-Dirac (smear in K)
    or Gaussian (smear in eV); a void smoother without a grid, a width or with one level."""
    if energy is None or smear is None or smear <= 0 or len(energy) <= 1:
        return VoidSmoother()
    if mode == "Fermi-Dirac":
        return FermiDiracSmoother(energy, smear)
    if mode == "Gaussian":
        return GaussianSmoother(energy, smear)
    raise ValueError("Smoother mode not recognized.")
