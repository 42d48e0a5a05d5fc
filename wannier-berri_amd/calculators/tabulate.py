"""k-resolved tabulation on the MI355X (wannierberri/calculators/tabulate.py:12-221).

Each tabulator returns, per k-point of the K-block and per band of `ibands`,
the trace of the band's group (degen_thresh) divided by the group size, from
one wbmi_tabulate launch per K-block; TabulatorAll packs them into a
TABresult, which run() puts on the full grid at the end.
"""
import numpy as np

from .. import _lib
from .. import factors
from ..result import KBandResult, TABresult
from .static import (Calculator, StaticCalculator, _PARITY, _external, der3e_bands, spin_current_channels,
                     spin_current_params)


class Tabulator(Calculator):
    """Tabulator (tabulate.py:12-48): formula id, rank and constant factor per subclass."""
    formula = None
    ndim = 0
    sign = 0
    factor = 1.

    def __init__(self, ibands=None, kwargs_formula=None, constant_factor=None, **kwargs):
        super().__init__(**kwargs)
        self.ibands = np.array(ibands) if ibands is not None else None
        self.kwargs_formula = dict(kwargs_formula or {})
        for k in self.kwargs_formula:
            if k not in ("internal_terms", "external_terms"):
                raise NotImplementedError(f"kwargs_formula option '{k}' is not available on the MI355X path")
        self.constant_factor = self.factor if constant_factor is None else constant_factor

    def needed_channels(self, system):
        return StaticCalculator.needed_channels(self, system)

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        p = _lib.ScanParams()
        p.formula = self.formula
        p.sign = self.sign
        p.internal_terms = int(bool(self.kwargs_formula.get("internal_terms", True)))
        p.external_terms = int(_external(self.kwargs_formula, data_K.system))
        p.fder, p.nEF_extra, p.EFmin, p.EFmax, p.dEF = 1, 1, 0.0, 0.0, 1.0
        p.degen_thresh = float(self.degen_thresh)
        p.degen_Kramers = int(bool(self.degen_Kramers))
        ncomp = 3 ** self.ndim
        out = torch.empty((data_K.nk, data_K.num_wann, ncomp), dtype=torch.float64, device=data_K.ctx.device)
        data_K.tabulate(p, out)
        ib = np.arange(data_K.num_wann) if self.ibands is None else self.ibands
        data = out.cpu().numpy()[:, ib].reshape((data_K.nk, len(ib)) + (3,) * self.ndim) * self.constant_factor
        tr, inv = _PARITY[self.formula]
        return KBandResult(data, transformTR=tr, transformInv=inv)


class Energy(Tabulator):
    """Band energies, the mean of each degenerate group (formula Eavln)"""
    formula = _lib.FORMULA_IDENTITY


class BerryCurvature(Tabulator):
    r""" Berry curvature :math:`\Omega_a` in units of angstrom^2"""
    formula = _lib.FORMULA_OMEGA
    ndim = 1


class DerBerryCurvature(Tabulator):
    r"Derivative of Berry curvature :math:`X_{ab}\partial_b\Omega_a` in units of angstrom^3"
    formula = _lib.FORMULA_DEROMEGA
    ndim = 2


class OrbitalMoment(Tabulator):
    r"Orbital moment :math:`m_{orb}` in units of Bohr magneton"
    formula = _lib.FORMULA_MORB_HPM
    ndim = 1
    sign = -1
    factor = factors.factor_morb_evA2_to_muB


class DerOrbitalMoment(Tabulator):
    r"Derivative of orbital moment :math:`\partial_a m_{orb}` in units of Bohr magneton*angstrom"
    formula = _lib.FORMULA_DERMORB
    ndim = 2
    sign = -1
    factor = factors.factor_morb_evA2_to_muB


class Der3E(Tabulator):
    r"Third derivative of the band energy :math:`\partial_a\partial_b\partial_c E` (tabulate.py:141-145), eV*angstrom^3"
    ndim = 3

    def needed_channels(self, system):
        return {("Ham", 0), ("Ham", 1), ("Ham", 2), ("Ham", 3)}

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        out = der3e_bands(data_K, self.degen_thresh, self.degen_Kramers, 0)
        ib = np.arange(data_K.num_wann) if self.ibands is None else self.ibands
        data = out.cpu().numpy()[:, ib].reshape((data_K.nk, len(ib), 3, 3, 3)) * self.constant_factor
        return KBandResult(data, transformTR=-1, transformInv=-1)


class DerSpin(Tabulator):
    r"Derivative of the spin :math:`\partial_b \langle u | \sigma_s | u \rangle` (tabulate.py:176-180), indices [s, b]"
    formula = _lib.FORMULA_DERSPIN
    ndim = 2


class _DiagTabulator(Tabulator):
    """Tabulators whose group traces are sums of diagonal elements of one channel triple
    (formula/covariant.py:321-336): one wbmi_tabulate_diag launch."""
    key = None
    ndim = 1

    def needed_channels(self, system):
        return {("Ham", 0), self.key}

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        s0 = data_K._slots[self.key]
        out = torch.empty((data_K.nk, data_K.num_wann, 3), dtype=torch.float64, device=data_K.ctx.device)
        data_K.ctx.tabulate_diag(data_K.E_dev, data_K.Xk, data_K.prec, [s0, s0 + 1, s0 + 2], self.degen_thresh,
                                 self.degen_Kramers, out)
        ib = np.arange(data_K.num_wann) if self.ibands is None else self.ibands
        data = out.cpu().numpy()[:, ib] * self.constant_factor
        return KBandResult(data, transformTR=self.parity[0], transformInv=self.parity[1])


class Velocity(_DiagTabulator):
    """Band velocity dE/dk (Re of the diagonal of Xbar('Ham', 1))"""
    key = ("Ham", 1)
    parity = (-1, -1)


class Spin(_DiagTabulator):
    r""" Spin expectation :math:` \langle u | \mathbf{\sigma} | u \rangle`"""
    key = ("SS", 0)
    parity = (-1, 1)


class InvMass(Tabulator):
    r""" second derivative of energy :math:`\partial_a\partial_b E` in units of eV*angstrom^2"""
    ndim = 2

    def needed_channels(self, system):
        return {("Ham", 0), ("Ham", 1), ("Ham", 2)}

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        v0, w0 = data_K._slots[("Ham", 1)], data_K._slots[("Ham", 2)]
        out = torch.empty((data_K.nk, data_K.num_wann, 9), dtype=torch.float64, device=data_K.ctx.device)
        data_K.ctx.tabulate_invmass(data_K.E_dev, data_K.Xk, data_K.prec, [v0 + i for i in range(3)],
                                    [w0 + i for i in range(9)], self.degen_thresh, self.degen_Kramers, out)
        ib = np.arange(data_K.num_wann) if self.ibands is None else self.ibands
        data = out.cpu().numpy()[:, ib].reshape((data_K.nk, len(ib), 3, 3)) * self.constant_factor
        return KBandResult(data, transformTR=1, transformInv=1)


class SpinBerry(Tabulator):
    r"""Spin Berry curvature per band (tabulate.py:218-221, formula/covariant.py:759-789):
    group mean of :math:`-2 \mathrm{Im} \sum_{l \notin G} J^{as}_{ml} (D^b_{lm} - i \bar A^b_{lm}) / (E_m - E_l)`,
    indices [a, b, s]; `kwargs_formula` may set spin_current_type ('ryoo', 'qiao', 'simple')."""
    ndim = 3

    def __init__(self, ibands=None, kwargs_formula=None, **kwargs):
        kf = dict(kwargs_formula or {})
        self.spin_current_type = kf.pop("spin_current_type", "ryoo")
        if self.spin_current_type not in ("ryoo", "qiao", "simple"):
            raise ValueError(f"spin_current_type must be `qiao` or `ryoo` or `simple`, not {self.spin_current_type}")
        super().__init__(ibands=ibands, kwargs_formula=kf, **kwargs)

    def needed_channels(self, system):
        return spin_current_channels(self.spin_current_type, _external(self.kwargs_formula, system))

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        ext = _external(self.kwargs_formula, data_K.system)
        p = spin_current_params(data_K, self.spin_current_type, ext, self.degen_thresh, self.degen_Kramers)
        out = torch.empty((data_K.nk, data_K.num_wann, 27), dtype=torch.float64, device=data_K.ctx.device)
        data_K.ctx.tabulate_spin_berry(data_K.E_dev, data_K.Xk, data_K.prec, p, out)
        ib = np.arange(data_K.num_wann) if self.ibands is None else self.ibands
        data = out.cpu().numpy()[:, ib].reshape((data_K.nk, len(ib), 3, 3, 3)) * self.constant_factor
        return KBandResult(data, transformTR=1, transformInv=1)


class TabulatorAll(Calculator):
    """A pack of tabulators (tabulate.py:55-105); "Energy" is always included.  mode 'grid'
    (run on a Grid, tables put on the full grid) or 'path' (run on a Path, tables in path order)."""

    def __init__(self, tabulators, ibands=None, mode="grid", save_mode="bin", print_comment=False):
        super().__init__(save_mode=save_mode, print_comment=print_comment)
        mode = mode.lower()
        if mode not in ("grid", "path"):
            raise ValueError(f"mode must be 'grid' or 'path', not '{mode}'")
        self.mode = mode
        self.tabulators = dict(tabulators)
        if "Energy" not in self.tabulators:
            self.tabulators["Energy"] = Energy()
        ibands = None if ibands is None else np.array(ibands)
        for k, v in self.tabulators.items():
            if not hasattr(v, "ibands"):
                continue   # k-resolved static calculators (static.py, k_resolved=True)
            if v.ibands is not None:
                if ibands is None or len(v.ibands) != len(ibands) or np.any(v.ibands != ibands):
                    raise ValueError(f"tabulator {k} has ibands={v.ibands} not equal to ibands={ibands} "
                                     "required in TabulatorAll")
            else:
                v.ibands = ibands

    def needed_channels(self, system):
        need = set()
        for v in self.tabulators.values():
            need |= set(v.needed_channels(system))
        return need

    @property
    def allow_path(self):
        return self.mode == "path"

    @property
    def allow_grid(self):
        return self.mode == "grid"

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        return TABresult(kpoints=np.array(data_K.kpoints_all, dtype=float), recip_lattice=data_K.system.recip_lattice,
                         results={k: v(data_K) for k, v in self.tabulators.items()}, mode=self.mode,
                         save_mode=self.save_mode)


__all__ = ["Tabulator", "TabulatorAll", "Energy", "Velocity", "Spin", "InvMass", "BerryCurvature", "DerBerryCurvature",
           "OrbitalMoment", "DerOrbitalMoment", "SpinBerry", "DerSpin", "Der3E"]
