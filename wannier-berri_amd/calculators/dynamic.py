"""Frequency-dependent calculators on the MI355X.

OpticalConductivity, JDOS and SHC keep the constructors of
wannierberri/calculators/dynamic.py:26-58,145-241 (Efermi, omega, kBT,
smr_fixed_width, smr_type, kwargs_formula, degen_thresh, degen_Kramers;
SHC_type, shc_abc) and their results: EnergyResult on (Efermi, omega) of
complex [nEF, nomega, 3, 3] tensors times factor_opt / (nk V), real
[nEF, nomega] / (nk V), complex [nEF, nomega, 3, 3, 3] times
factor_shc / (nk V).  The per-k pair sums and the (Fermi level, frequency)
accumulation run in wbmi_optical_conductivity, wbmi_jdos and wbmi_spin_hall.
"""
import numpy as np

from .. import _lib
from .. import factors
from ..result import EnergyResult
from .static import Calculator, _external


def _params(calc, data_K):
    """wbmi_optcond_params of a dynamic calculator (device copies of Efermi and omega kept alive
    on the calculator)."""
    torch = data_K.torch
    dev = data_K.ctx.device
    key = str(dev)
    if key not in calc._dev_arrays:
        calc._dev_arrays[key] = (torch.from_numpy(calc.Efermi).to(dev), torch.from_numpy(calc.omega).to(dev))
    ef, om = calc._dev_arrays[key]
    p = _lib.OptcondParams()
    p.nEF, p.nomega = len(calc.Efermi), len(calc.omega)
    p.Efermi, p.omega = ef.data_ptr(), om.data_ptr()
    p.EFmin, p.EFmax = float(calc.Efermi.min()), float(calc.Efermi.max())
    p.omega_min, p.omega_max = float(calc.omega.min()), float(calc.omega.max())
    p.kBT, p.eta = calc.kBT, calc.smr_fixed_width
    p.smr_gaussian = int(calc.smr_type == "Gaussian")
    p.degen_thresh = float(calc.degen_thresh)
    p.degen_Kramers = int(bool(calc.degen_Kramers))
    return p


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
        self._dev_arrays = {}

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
        p = _params(self, data_K)
        p.external_terms = int(_external(self.kwargs_formula, data_K.system))
        for i, v in enumerate(data_K.role_table()):
            p.ch[i] = v
        out = torch.zeros((nEF, nw_, 3, 3, 2), dtype=torch.float64, device=dev)
        data_K.ctx.optical_conductivity(data_K.E_dev, data_K.Xk, data_K.prec, p, out)
        data = torch.view_as_complex(out) * (self.constant_factor / (data_K.nk * data_K.cell_volume))
        # transformTR of Formula_OptCond transposes (a, b): not a factor-only parity
        return EnergyResult([self.Efermi, self.omega], data, transformTR=None, transformInv=1,
                            E_titles=("Efermi", "Omega"), comment=self.comment, save_mode=self.save_mode)


class JDOS(Calculator):
    r"""Joint density of states (dynamic.py:145-167)

    :math:`\rho(\omega) = \sum_{\mathbf{k}} \sum_{m,n} \delta(E_{m\mathbf{k}} - E_{n\mathbf{k}} - \omega)
    \times \left(f(E_{n\mathbf{k}}) - f(E_{m\mathbf{k}})\right)`"""

    def __init__(self, Efermi=None, omega=None, kBT=0, smr_fixed_width=0.1, smr_type="Lorentzian",
                 kwargs_formula=None, dtype=float, **kwargs):
        super().__init__(**kwargs)
        if smr_type not in ("Lorentzian", "Gaussian"):
            raise ValueError(f"Invalid smearing type {smr_type}")
        self.Efermi = np.array(Efermi, dtype=float).reshape(-1)
        self.omega = np.array(omega, dtype=float).reshape(-1)
        self.kBT = float(kBT)
        self.smr_fixed_width = float(smr_fixed_width)
        self.smr_type = smr_type
        self._dev_arrays = {}

    def needed_channels(self, system):
        return {("Ham", 0)}

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        p = _params(self, data_K)
        out = torch.zeros((len(self.Efermi), len(self.omega)), dtype=torch.float64, device=data_K.ctx.device)
        data_K.ctx.jdos(data_K.E_dev, p, out)
        data = out * (1.0 / (data_K.nk * data_K.cell_volume))
        return EnergyResult([self.Efermi, self.omega], data, transformTR=1, transformInv=1,
                            E_titles=("Efermi", "Omega"), comment=self.comment, save_mode=self.save_mode)


class SHC(Calculator):
    r"""Kubo spin Hall conductivity (dynamic.py:224-241) with the Ryoo, Qiao or 'simple' spin
    current (formula/covariant.py:703-756); 'simple' also without external terms."""

    def __init__(self, SHC_type="ryoo", shc_abc=None, Efermi=None, omega=None, kBT=0, smr_fixed_width=0.1,
                 smr_type="Lorentzian", kwargs_formula=None, dtype=complex, **kwargs):
        super().__init__(**kwargs)
        if SHC_type not in ("ryoo", "qiao", "simple"):
            raise ValueError(f"spin_current_type must be `qiao` or `ryoo` or `simple`, not {SHC_type}")
        self.SHC_type = SHC_type
        if smr_type not in ("Lorentzian", "Gaussian"):
            raise ValueError(f"Invalid smearing type {smr_type}")
        self.kwargs_formula = dict(kwargs_formula or {})
        for k in self.kwargs_formula:
            if k != "external_terms":
                raise NotImplementedError(f"kwargs_formula option '{k}' is not available on the MI355X path")
        if shc_abc is not None:
            assert len(shc_abc) == 3
        self.shc_abc = None if shc_abc is None else tuple(int(x) - 1 for x in shc_abc)
        self.Efermi = np.array(Efermi, dtype=float).reshape(-1)
        self.omega = np.array(omega, dtype=float).reshape(-1)
        self.kBT = float(kBT)
        self.smr_fixed_width = float(smr_fixed_width)
        self.smr_type = smr_type
        self.constant_factor = factors.factor_shc
        self._dev_arrays = {}

    def needed_channels(self, system):
        if self.SHC_type == "simple":
            return {("Ham", 0), ("Ham", 1), ("SS", 0)} | ({("AA", 0)} if _external(self.kwargs_formula, system) else set())
        spin = [("SA", 0), ("SHA", 0)] if self.SHC_type == "ryoo" else [("SH", 0), ("SR", 0), ("SHR", 0)]
        return {("Ham", 0), ("Ham", 1), ("AA", 0), ("SS", 0), *spin}

    def __call__(self, data_K):
        torch = data_K.torch
        ext = _external(self.kwargs_formula, data_K.system)
        if not ext and self.SHC_type != "simple":
            raise NotImplementedError(f"spin Hall {self.SHC_type} without external terms is not implemented yet. "
                                      "Use `SHC_type='simple'`")
        data_K.ensure(self.needed_channels(data_K.system))
        p = _lib.ShcParams()
        p.base = _params(self, data_K)
        p.base.external_terms = int(ext)
        for i, v in enumerate(data_K.role_table()):
            p.base.ch[i] = v
        sl = data_K._slots
        p.kind = {"ryoo": _lib.WBMI_SHC_RYOO, "qiao": _lib.WBMI_SHC_QIAO, "simple": _lib.WBMI_SHC_SIMPLE}[self.SHC_type]
        for c in range(3):
            p.ch_S[c] = sl[("SS", 0)] + c
            p.ch_SH[c] = sl[("SH", 0)] + c if ("SH", 0) in sl else -1
        for c in range(9):
            p.ch_SA[c] = sl[("SA", 0)] + c if ("SA", 0) in sl else -1
            p.ch_SHA[c] = sl[("SHA", 0)] + c if ("SHA", 0) in sl else -1
            p.ch_SR[c] = sl[("SR", 0)] + c if ("SR", 0) in sl else -1
            p.ch_SHR[c] = sl[("SHR", 0)] + c if ("SHR", 0) in sl else -1
        nEF, nom = len(self.Efermi), len(self.omega)
        out = torch.zeros((nEF, nom, 3, 3, 3, 2), dtype=torch.float64, device=data_K.ctx.device)
        data_K.ctx.spin_hall(data_K.E_dev, data_K.Xk, data_K.prec, p, out)
        data = torch.view_as_complex(out) * (self.constant_factor / (data_K.nk * data_K.cell_volume))
        if self.shc_abc is not None:
            a, b, c = self.shc_abc
            data = data[:, :, a, b, c].contiguous()
        return EnergyResult([self.Efermi, self.omega], data, transformTR=1, transformInv=1,
                            E_titles=("Efermi", "Omega"), comment=self.comment, save_mode=self.save_mode)


class _PairCurrent(Calculator):
    """Shift and injection current: the pair walk of wbmi_shift_current / wbmi_injection_current."""
    external_ok = True

    def __init__(self, Efermi=None, omega=None, kBT=0, smr_fixed_width=0.1, smr_type="Lorentzian",
                 kwargs_formula=None, **kwargs):
        super().__init__(**kwargs)
        if smr_type not in ("Lorentzian", "Gaussian"):
            raise ValueError(f"Invalid smearing type {smr_type}")
        self.kwargs_formula = dict(kwargs_formula or {})
        for k in self.kwargs_formula:
            if k != "external_terms":
                raise NotImplementedError(f"kwargs_formula option '{k}' is not available on the MI355X path")
        self.Efermi = np.array(Efermi, dtype=float).reshape(-1)
        self.omega = np.array(omega, dtype=float).reshape(-1)
        self.kBT = float(kBT)
        self.smr_fixed_width = float(smr_fixed_width)
        self.smr_type = smr_type
        self._dev_arrays = {}

    def _params(self, data_K):
        p = _params(self, data_K)
        p.external_terms = int(_external(self.kwargs_formula, data_K.system))
        for i, v in enumerate(data_K.role_table()):
            p.ch[i] = v
        return p

    def _out(self, data_K):
        return data_K.torch.zeros((len(self.Efermi), len(self.omega), 3, 3, 3, 2), dtype=data_K.torch.float64,
                                  device=data_K.ctx.device)


class ShiftCurrent(_PairCurrent):
    r"""Shift current (dynamic.py:244-324), real [nEF, nomega, 3, 3, 3]"""

    def __init__(self, sc_eta, **kwargs):
        super().__init__(**kwargs)
        self.sc_eta = float(sc_eta)
        self.constant_factor = factors.factor_shift_current

    def needed_channels(self, system):
        need = {("Ham", 0), ("Ham", 1), ("Ham", 2)}
        if _external(self.kwargs_formula, system):
            need |= {("AA", 0), ("AA", 1)}
        return need

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        p = self._params(data_K)
        out = self._out(data_K)
        data_K.ctx.shift_current(data_K.E_dev, data_K.Xk, data_K.prec, p, self.sc_eta, out)
        data = out[..., 0] * (self.constant_factor / (data_K.nk * data_K.cell_volume))
        return EnergyResult([self.Efermi, self.omega], data, transformTR=1, transformInv=-1,
                            E_titles=("Efermi", "Omega"), comment=self.comment, save_mode=self.save_mode)


class InjectionCurrent(_PairCurrent):
    r"""Injection current (dynamic.py:327-361), complex [nEF, nomega, 3, 3, 3]"""

    def __init__(self, **kwargs):
        super().__init__(**kwargs)
        self.constant_factor = factors.factor_injection_current

    def needed_channels(self, system):
        need = {("Ham", 0), ("Ham", 1)}
        if _external(self.kwargs_formula, system):
            need.add(("AA", 0))
        return need

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        p = self._params(data_K)
        out = self._out(data_K)
        data_K.ctx.injection_current(data_K.E_dev, data_K.Xk, data_K.prec, p, out)
        data = torch.view_as_complex(out) * (self.constant_factor / (data_K.nk * data_K.cell_volume))
        # transformTR of the reference transposes (b, c): not a factor-only parity
        return EnergyResult([self.Efermi, self.omega], data, transformTR=None, transformInv=-1,
                            E_titles=("Efermi", "Omega"), comment=self.comment, save_mode=self.save_mode)


__all__ = ["OpticalConductivity", "JDOS", "SHC", "ShiftCurrent", "InjectionCurrent"]
