"""Frequency-dependent calculators on the MI355X.

OpticalConductivity keeps the constructor of
wannierberri/calculators/dynamic.py:26-58,184-196 (Efermi, omega, kBT,
smr_fixed_width, smr_type, kwargs_formula, degen_thresh, degen_Kramers) and
its result: EnergyResult on (Efermi, omega) of complex [nEF, nomega, 3, 3]
tensors, times factor_opt / (nk V).  The per-k pair sums and the
(Fermi level, frequency) accumulation run in wbmi_optical_conductivity.
"""
import numpy as np

from .. import _lib
from .. import factors
from ..result import EnergyResult
from .static import Calculator, _external


class OpticalConductivity(Calculator):
    r"""Kubo interband optical conductivity (dynamic.py:184-196, Formula_OptCond 170-181)"""

    def __init__(self, Efermi=None, omega=None, kBT=0, smr_fixed_width=0.1, smr_type="Lorentzian",
                 kwargs_formula=None, dtype=complex, **kwargs):
        super().__init__(**kwargs)
        if smr_type not in ("Lorentzian", "Gaussian"):
            raise ValueError(f"Invalid smearing type {smr_type}")
        self.Efermi = np.array(Efermi, dtype=float).reshape(-1)
        self.omega = np.array(omega, dtype=float).reshape(-1)
        self.kBT = float(kBT)
        self.smr_fixed_width = float(smr_fixed_width)
        self.smr_type = smr_type
        self.kwargs_formula = dict(kwargs_formula or {})
        for k in self.kwargs_formula:
            if k != "external_terms":
                raise NotImplementedError(f"kwargs_formula option '{k}' is not available on the MI355X path")
        self.constant_factor = factors.factor_opt

    def needed_channels(self, system):
        need = {("Ham", 0), ("Ham", 1)}
        if _external(self.kwargs_formula, system):
            need.add(("AA", 0))
        return need

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        dev = data_K.ctx.device
        nEF, nw_ = len(self.Efermi), len(self.omega)
        ef = torch.from_numpy(self.Efermi).to(dev)
        om = torch.from_numpy(self.omega).to(dev)
        p = _lib.OptcondParams()
        p.nEF, p.nomega = nEF, nw_
        p.Efermi, p.omega = ef.data_ptr(), om.data_ptr()
        p.EFmin, p.EFmax = float(self.Efermi.min()), float(self.Efermi.max())
        p.kBT, p.eta = self.kBT, self.smr_fixed_width
        p.smr_gaussian = int(self.smr_type == "Gaussian")
        p.degen_thresh = float(self.degen_thresh)
        p.degen_Kramers = int(bool(self.degen_Kramers))
        p.external_terms = int(_external(self.kwargs_formula, data_K.system))
        for i, v in enumerate(data_K.role_table()):
            p.ch[i] = v
        out = torch.zeros((nEF, nw_, 3, 3, 2), dtype=torch.float64, device=dev)
        data_K.ctx.optical_conductivity(data_K.E_dev, data_K.Xk, data_K.prec, p, out)
        data = torch.view_as_complex(out) * (self.constant_factor / (data_K.nk * data_K.cell_volume))
        # transformTR of Formula_OptCond transposes (a, b): not a factor-only parity
        return EnergyResult([self.Efermi, self.omega], data, transformTR=None, transformInv=1,
                            E_titles=("Efermi", "Omega"), comment=self.comment, save_mode=self.save_mode)


__all__ = ["OpticalConductivity"]
