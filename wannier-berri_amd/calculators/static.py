"""Static (Fermi-level scan) calculators on the MI355X.

Same constructor signatures and outputs as wannierberri/calculators/static.py;
each call runs the fused formula-trace + Fermi-scan kernel (wbmi_static_scan)
on the K-block and the prefix-sum / fder-stencil / scaling kernel
(wbmi_finalize); with tetra=True the tetrahedron-weighted scan
(wbmi_static_scan_tetra, corner energies from Data_K.corner_energies) and
wbmi_finalize_tetra.  The result stays on the GPU until read.
"""
import functools

import numpy as np

from .. import _lib
from .. import factors
from ..result import EnergyResult, KBandResult, VoidResult


class Calculator:
    """calculators/calculator.py:6-43."""

    def __init__(self, degen_thresh=1e-4, degen_Kramers=False, save_mode="bin+txt", print_comment=False):
        self.degen_thresh = degen_thresh
        self.degen_Kramers = degen_Kramers
        self.save_mode = save_mode
        if not hasattr(self, "comment"):
            self.comment = self.__doc__ if self.__doc__ is not None else "calculator not described"
        if print_comment:
            print(self.comment)

    @property
    def allow_path(self):
        return False

    @property
    def allow_grid(self):
        return True


# parity (+1 even, -1 odd) under time reversal and inversion of each formula's trace
# (formula/covariant.py: Identity 15-16, Omega 172-173, DerOmega 224-225,
# Morb_Hpm 435-436, DerMorb 535-536; VelOmega and VelHplus are products with the
# velocity, odd under both, formula/formula.py:127-128, data_K.py:16-46)
_PARITY = {_lib.FORMULA_IDENTITY: (1, 1), _lib.FORMULA_OMEGA: (-1, 1), _lib.FORMULA_DEROMEGA: (1, -1),
           _lib.FORMULA_VELOMEGA: (1, -1), _lib.FORMULA_MORB_HPM: (-1, 1), _lib.FORMULA_DERMORB: (1, -1),
           _lib.FORMULA_VELHPLUS: (1, -1), _lib.FORMULA_DERSPIN: (1, -1), _lib.FORMULA_VELSPIN: (1, -1),
           _lib.FORMULA_SPIN: (-1, 1), _lib.FORMULA_INVMASS: (1, 1), _lib.FORMULA_VELVEL: (1, 1),
           # products of velocities (odd, odd) and inverse masses (even, even) (formula.py:127-128);
           # the Hall tensors keep the parity of their rank-4 products (static.py:418-423)
           _lib.FORMULA_VELVELVEL: (-1, -1), _lib.FORMULA_MASSVEL: (-1, -1), _lib.FORMULA_HALL_SURF: (1, 1),
           _lib.FORMULA_HALL_SEA: (1, 1),
           # Omega (odd, even) times the spin (odd, even), Omega or Morb_Hpm (odd, even)
           _lib.FORMULA_OMEGASPIN: (1, 1), _lib.FORMULA_OMEGAOMEGA: (1, 1), _lib.FORMULA_OMEGAHPLUS: (1, 1)}
# formulas that read only the Hamiltonian's derivatives and the spin, never A, B, C
_SPIN = (_lib.FORMULA_DERSPIN, _lib.FORMULA_VELSPIN, _lib.FORMULA_SPIN)


# relative Hermiticity defect of Ham_R above which the k-derivative formulas refuse a system
HERMITIAN_TOL = 1e-8


def require_hermitian_ham(system, what):
    """DerOmega and DerMorb on the GPU take dH/dk and d2H/dk2 Hermitian (one product serves
    a term and its transpose, and W(y, x) is read as conj W(x, y)).  That holds when
    Ham(-R) = Ham(R)^+ for every R-vector; the reference does not force it
    (data_K_R.py:84-87), so a system without it is refused rather than given a
    different answer."""
    err = system.hermiticity_error("Ham")
    if err > HERMITIAN_TOL:
        raise NotImplementedError(
            f"{what} on the MI355X path needs Ham(-R) = Ham(R)^+ for every R-vector (an R-vector without its -R "
            f"partner breaks it); this system deviates by {err:.2e} relative to max|Ham|")


def _external(kwargs_formula, system):
    ext = kwargs_formula.get("external_terms", True)
    return bool(ext) and not system.force_internal_terms_only


def _with_band_weights(call):
    """Every __call__ of a static calculator (composites included) runs with the device
    select_bands weights of the call in place (StaticCalculator.scan_params reads them); a
    nested call reuses the outer call's."""
    @functools.wraps(call)
    def wrapped(self, data_K):
        if self.select_bands is None or getattr(self, "_bw_now", None) is not None:
            return call(self, data_K)
        self._bw_now = self._band_weights(data_K)
        try:
            return call(self, data_K)
        finally:
            self._bw_now = None
    return wrapped


class StaticCalculator(Calculator):
    """Fermi-level scan integrator (calculators/static.py:24-169).

    Subclasses set `formula` (a WBMI_FORMULA_* id), `ndim` and `fder`.
    select_bands (Fermi-level bins, fder >= 1): each band group's value is weighted by the
    share of its bands listed (static.py:133,136).  Emin / Emax bound the Fermi-sea (hole-like) group of the
    tetrahedron method (tetrahedron.py:246-258) and have no effect on Fermi-level bins, as in the
    reference (data_K.py:172-186 receives them but does not use them).
    """
    formula = None
    ndim = 0
    sign = 0

    def __init__(self, Efermi, tetra=False, smoother=None, constant_factor=1., use_factor=True, kwargs_formula=None,
                 Emin=-np.inf, Emax=np.inf, hole_like=False, k_resolved=False, fder=None, select_bands=None,
                 **kwargs):
        super().__init__(**kwargs)
        self.k_resolved = bool(k_resolved)
        self.select_bands = None if select_bands is None else np.array(select_bands, dtype=np.int64).reshape(-1)
        self._band_w = {}
        # Fermi-level bins: without effect, as in the reference (get_bands_in_range_groups_ik
        # takes them but selects groups by the Fermi range alone, data_K.py:172-186); tetra:
        # they bound the Fermi-sea / hole-like group (tetrahedron.py:246-258)
        self.Emin, self.Emax = float(Emin), float(Emax)
        self.kwargs_formula = dict(kwargs_formula or {})
        for k in self.kwargs_formula:
            if k not in ("internal_terms", "external_terms", "sign"):
                raise NotImplementedError(f"kwargs_formula option '{k}' is not available on the MI355X path")
        self.Efermi = np.array(Efermi, dtype=float)
        self.tetra = bool(tetra)
        if self.tetra and np.any(np.diff(self.Efermi) < 0):
            raise NotImplementedError("the tetrahedron method on the MI355X path needs ascending Fermi levels")
        self._Ef_dev = {}
        self.smoother = smoother
        self.use_factor = use_factor
        self.hole_like = hole_like
        if fder is not None:
            self.fder = fder
        self.constant_factor = constant_factor
        if hole_like and self.fder == 0:
            self.constant_factor *= -1
        self.extraEf = {0: 0, 1: 1, 2: 1, 3: 2}[self.fder]
        self.dEF = self.Efermi[1] - self.Efermi[0] if len(self.Efermi) > 1 else 0.001
        self.EFmin = self.Efermi[0] - self.extraEf * self.dEF
        self.EFmax = self.Efermi[-1] + self.extraEf * self.dEF
        self.nEF_extra = self.Efermi.shape[0] + 2 * self.extraEf

    @property
    def ncomp(self):
        """Components of one library call: rank-3 formulas run as three calls of 9 (comp_slice)."""
        return min(3 ** self.ndim, 9)

    @property
    def nslices(self):
        return 3 if self.ndim == 3 else 1

    @property
    def parity(self):
        """(time reversal, inversion) parity of the result (+1 even, -1 odd)."""
        return _PARITY[self.formula]

    def needed_channels(self, system):
        """Channels Xbar(name, der) the formula reads."""
        ext = _external(self.kwargs_formula, system)
        f = self.formula
        if f == _lib.FORMULA_IDENTITY:
            return {("Ham", 0)}
        if f == _lib.FORMULA_SPIN:
            return {("Ham", 0), ("SS", 0)}
        if f in _SPIN:
            return {("Ham", 0), ("Ham", 1), ("SS", 0)} | ({("SS", 1)} if f == _lib.FORMULA_DERSPIN else set())
        if f in (_lib.FORMULA_VELVEL, _lib.FORMULA_VELVELVEL):
            return {("Ham", 0), ("Ham", 1)}
        if f in (_lib.FORMULA_INVMASS, _lib.FORMULA_MASSVEL, _lib.FORMULA_HALL_SURF, _lib.FORMULA_HALL_SEA):
            return {("Ham", 0), ("Ham", 1), ("Ham", 2)}
        need = {("Ham", 0), ("Ham", 1)}
        if f in (_lib.FORMULA_DEROMEGA, _lib.FORMULA_DERMORB):
            require_hermitian_ham(system, type(self).__name__)
            need.add(("Ham", 2))
        if f == _lib.FORMULA_OMEGASPIN:
            need.add(("SS", 0))
        if ext:
            need |= {("AA", 0), ("rotAA", 0)}
            if f in (_lib.FORMULA_DEROMEGA, _lib.FORMULA_DERMORB):
                need |= {("AA", 1), ("rotAA", 1)}
            if f in (_lib.FORMULA_MORB_HPM, _lib.FORMULA_VELHPLUS, _lib.FORMULA_DERMORB, _lib.FORMULA_OMEGAHPLUS):
                need |= {("BB", 0), ("CC", 0)}
            if f == _lib.FORMULA_DERMORB:
                need |= {("BB", 1), ("CC", 1)}
        return need

    def scan_params(self, system):
        p = _lib.ScanParams()
        p.formula = self.formula
        p.sign = int(self.kwargs_formula.get("sign", self.sign))
        p.internal_terms = int(bool(self.kwargs_formula.get("internal_terms", True)))
        p.external_terms = int(_external(self.kwargs_formula, system)) if self.formula not in _SPIN else 0
        p.fder = self.fder
        p.nEF_extra = self.nEF_extra
        p.EFmin, p.EFmax, p.dEF = float(self.EFmin), float(self.EFmax), float(self.dEF)
        p.degen_thresh = float(self.degen_thresh)
        p.degen_Kramers = int(bool(self.degen_Kramers))
        p.comp_slice = 0
        bw = getattr(self, "_bw_now", None)
        p.band_weight = bw.data_ptr() if bw is not None else None
        p.erange = int(self.tetra and (np.isfinite(self.Emin) or np.isfinite(self.Emax)))
        p.Emin, p.Emax = self.Emin, self.Emax
        return p

    def _band_weights(self, data_K):
        """select_bands as a device [nw] float64 of listing counts (utility.py:398-403 counts
        every listed entry inside a group, so repeated bands count repeatedly); None without a
        selection.  The reference refuses a selection for the Fermi sea (data_K.py:179-180)."""
        if self.select_bands is None:
            return None
        if self.fder == 0:
            raise NotImplementedError("Selection of bands for Fermi sea is not implemented")
        if self.tetra and self.hole_like:
            raise NotImplementedError("Selection of bands for inverse Fermi sea is not implemented")
        nw = data_K.num_wann
        key = (data_K.ctx.device, nw)
        if key not in self._band_w:
            sel = self.select_bands[(self.select_bands >= 0) & (self.select_bands < nw)]
            w = np.bincount(sel, minlength=nw).astype(np.float64)
            self._band_w[key] = data_K.torch.from_numpy(w).to(data_K.ctx.device)
        return self._band_w[key]

    def _slices(self, data_K, run):
        """run(comp_slice) -> [..., nEF, 9] per slice, joined to [..., nEF, 27] (first index slowest)."""
        if self.nslices == 1:
            return run(0)
        outs = [run(a) for a in range(3)]
        return data_K.torch.stack(outs, dim=-2).reshape(outs[0].shape[:-1] + (27,))

    def _call_tetra(self, data_K, bands=None):
        """Tetrahedron weights (static.py:80-85,119-125): no extended Fermi grid, no stencil.
        Returns [nEF, ncomp], or [nk, nEF, ncomp] k-resolved (/ cell volume only)."""
        torch = data_K.torch
        dev = data_K.ctx.device
        Ef = self._Ef_dev.get(dev)
        if Ef is None:
            Ef = self._Ef_dev[dev] = torch.from_numpy(self.Efermi).to(dev)
        nEF = self.Efermi.shape[0]
        cf = self.constant_factor if self.use_factor else float(np.sign(self.constant_factor))

        def run(sl):
            p = self.scan_params(data_K.system)
            p.comp_slice = sl
            p.EFmin, p.EFmax = float(self.Efermi[0]), float(self.Efermi[-1])
            p.nEF_extra = nEF
            if self.hole_like:
                p.fder = -1   # 1 - occupation (static.py:81-82)
            if self.k_resolved:
                nk = data_K.nk
                hist = torch.zeros((nk, 2 * nEF, self.ncomp), dtype=torch.float64, device=dev)
                data_K.static_scan_tetra_k(p, Ef, hist, 1.0, bands=bands)
                out = torch.empty((nk, nEF, self.ncomp), dtype=torch.float64, device=dev)
                data_K.ctx.finalize_tetra_k(hist, cf / data_K.cell_volume, out)
                return out
            hist = torch.zeros((2 * nEF, self.ncomp), dtype=torch.float64, device=dev)
            data_K.static_scan_tetra(p, Ef, hist, 1.0, bands=bands)
            out = torch.empty((nEF, self.ncomp), dtype=torch.float64, device=dev)
            data_K.ctx.finalize_tetra(hist, cf / data_K.cell_volume / data_K.nk, out)
            return out

        return self._slices(data_K, run)

    def _call_k_resolved(self, data_K, bands=None):
        """Per k-point results [nk, nEF, 3^ndim] (static.py:60-101,156-164): each k its own
        histogram, / cell volume (no / nk), x factor."""
        torch = data_K.torch
        dev = data_K.ctx.device
        nk = data_K.nk
        nout = self.Efermi.shape[0]
        cf = self.constant_factor if self.use_factor else float(np.sign(self.constant_factor))

        def run(sl):
            p = self.scan_params(data_K.system)
            p.comp_slice = sl
            hist = torch.zeros((nk, self.nEF_extra, self.ncomp), dtype=torch.float64, device=dev)
            data_K.static_scan_k(p, hist, 1.0, bands=bands)
            out = torch.empty((nk, nout, self.ncomp), dtype=torch.float64, device=dev)
            data_K.ctx.finalize_k(hist, self.fder, self.dEF, cf / data_K.cell_volume, out)
            return out

        out = self._slices(data_K, run)
        tr, inv = self.parity
        return KBandResult(out.cpu().numpy().reshape((nk, nout) + (3,) * self.ndim), transformTR=tr, transformInv=inv)

    def band_values(self, data_K):
        """Per-band values [nk, nw, 3^ndim] (float64, device) for WBMI_FORMULA_BANDVALUES
        calculators, formed by the calculator before the scan; None for the others."""
        return None

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        if "__call__" in cls.__dict__:
            cls.__call__ = _with_band_weights(cls.__dict__["__call__"])

    def __call__(self, data_K):
        torch = data_K.torch
        data_K.ensure(self.needed_channels(data_K.system))
        bands = self.band_values(data_K)
        if self.k_resolved and not self.tetra:
            return self._call_k_resolved(data_K, bands)
        if self.tetra:
            out = self._call_tetra(data_K, bands)
            if self.k_resolved:
                tr, inv = self.parity
                return KBandResult(out.cpu().numpy().reshape((data_K.nk, self.Efermi.shape[0]) + (3,) * self.ndim),
                                   transformTR=tr, transformInv=inv)
            data = out.reshape((self.Efermi.shape[0],) + (3,) * self.ndim)
            tr, inv = self.parity
            return EnergyResult(self.Efermi, data, smoothers=[self.smoother], transformTR=tr, transformInv=inv,
                                comment=self.comment, save_mode=self.save_mode)
        dev = data_K.ctx.device
        nout = self.Efermi.shape[0]
        cf = self.constant_factor if self.use_factor else float(np.sign(self.constant_factor))
        scale = cf / data_K.cell_volume / data_K.nk

        def run(sl):
            p = self.scan_params(data_K.system)
            p.comp_slice = sl
            hist = torch.zeros((self.nEF_extra, self.ncomp), dtype=torch.float64, device=dev)
            data_K.static_scan(p, hist, 1.0, bands=bands)
            out = torch.empty((nout, self.ncomp), dtype=torch.float64, device=dev)
            data_K.ctx.finalize(hist, self.fder, self.dEF, scale, out)
            return out

        out = self._slices(data_K, run)
        data = out.reshape((nout,) + (3,) * self.ndim)
        tr, inv = self.parity
        return EnergyResult(self.Efermi, data, smoothers=[self.smoother], transformTR=tr, transformInv=inv,
                            comment=self.comment, save_mode=self.save_mode)


SPIN_CURRENT_KINDS = {"ryoo": _lib.WBMI_SHC_RYOO, "qiao": _lib.WBMI_SHC_QIAO, "simple": _lib.WBMI_SHC_SIMPLE}


def spin_current_channels(kind, ext):
    """Channels of the spin current J of `kind` (formula/covariant.py:703-756) and of the velocity
    and Berry connection next to it."""
    need = {("Ham", 0), ("Ham", 1), ("SS", 0)} | ({("AA", 0)} if ext else set())
    if kind == "ryoo":
        need |= {("SA", 0), ("SHA", 0)}
    elif kind == "qiao":
        need |= {("SH", 0), ("SR", 0), ("SHR", 0)}
    return need


def spin_current_params(data_K, kind, ext, degen_thresh, degen_Kramers):
    """wbmi_shc_params for the spin Berry curvature of the Data_K's channels (no Fermi levels)."""
    if kind not in SPIN_CURRENT_KINDS:
        raise ValueError(f"spin_current_type must be `qiao` or `ryoo` or `simple`, not {kind}")
    if not ext and kind != "simple":
        raise NotImplementedError(f"the {kind} spin current needs the external terms")
    p = _lib.ShcParams()
    p.base.degen_thresh = float(degen_thresh)
    p.base.degen_Kramers = int(bool(degen_Kramers))
    p.base.external_terms = int(ext)
    for i, v in enumerate(data_K.role_table()):
        p.base.ch[i] = v
    sl = data_K._slots
    p.kind = SPIN_CURRENT_KINDS[kind]
    for c in range(3):
        p.ch_S[c] = sl[("SS", 0)] + c
        p.ch_SH[c] = sl[("SH", 0)] + c if ("SH", 0) in sl else -1
    for c in range(9):
        for name, arr in (("SA", p.ch_SA), ("SHA", p.ch_SHA), ("SR", p.ch_SR), ("SHR", p.ch_SHR)):
            arr[c] = sl[(name, 0)] + c if (name, 0) in sl else -1
    return p


# -- implemented calculators (calculators/static.py:192-498) --------------------


StaticCalculator.__call__ = _with_band_weights(StaticCalculator.__call__)

class _DOS(StaticCalculator):
    formula = _lib.FORMULA_IDENTITY
    ndim = 0

    def __init__(self, fder, **kwargs):
        self.fder = fder
        super().__init__(**kwargs)

    def __call__(self, data_K):
        return super().__call__(data_K) * data_K.cell_volume


class DOS(_DOS):
    r"""Density of states"""

    def __init__(self, **kwargs):
        super().__init__(fder=1, **kwargs)


class CumDOS(_DOS):
    r"""Cumulative density of states"""

    def __init__(self, **kwargs):
        super().__init__(fder=0, **kwargs)


class AHC(StaticCalculator):
    r"""Anomalous Hall conductivity (:math:`s^3 \cdot A^2 / (kg \cdot m^3) = S/m`)

        | Output: :math:`O = -e^2/\hbar \int [dk] \Omega f`"""
    formula = _lib.FORMULA_OMEGA
    ndim = 1
    fder = 0

    def __init__(self, constant_factor=factors.factor_ahc, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class Ohmic_FermiSea(StaticCalculator):
    r"""Ohmic conductivity (:math:`S/m`), Fermi sea (static.py:379-390)

        | Output: :math:`\sigma_{\alpha\beta} = e^2/\hbar \tau \int [dk] \partial_\beta v_\alpha f` for tau = 1 fs
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta} E_\beta`"""
    formula = _lib.FORMULA_INVMASS
    ndim = 2
    fder = 0

    def __init__(self, constant_factor=factors.factor_ohmic, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class Ohmic_FermiSurf(StaticCalculator):
    r"""Ohmic conductivity (:math:`S/m`), Fermi surface (static.py:393-404)

        | Output: :math:`\sigma_{\alpha\beta} = -e^2/\hbar \tau \int [dk] v_\alpha v_\beta f'` for tau = 1 fs
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta} E_\beta`"""
    formula = _lib.FORMULA_VELVEL
    ndim = 2
    fder = 1

    def __init__(self, constant_factor=factors.factor_ohmic, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class SHC(StaticCalculator):
    r"""Spin Hall conductivity, dc (:math:`S/m`), Fermi sea (static.py:647-660)

        | Output: :math:`\sigma_{\alpha\beta}^s = -e^2/\hbar \int [dk] Im(j_{nm,\alpha} v_{mn,\beta})/(\epsilon_n - \epsilon_m)^2 f`
        | kwargs_formula={'spin_current_type': 'ryoo' | 'qiao' | 'simple'}

    The spin Berry curvature of every band (wbmi_spin_berry_bands: the spin current, then each
    band's share of its group's trace) is binned by the Fermi scan as caller values
    (WBMI_FORMULA_BANDVALUES), 9 components per launch."""
    formula = _lib.FORMULA_BANDVALUES
    ndim = 3
    fder = 0

    def __init__(self, constant_factor=-factors.factor_ahc / 2, kwargs_formula=None, **kwargs):
        kf = dict(kwargs_formula or {})
        self.spin_current_type = kf.pop("spin_current_type", "ryoo")
        if self.spin_current_type not in SPIN_CURRENT_KINDS:
            raise ValueError(f"spin_current_type must be `qiao` or `ryoo` or `simple`, not {self.spin_current_type}")
        super().__init__(constant_factor=constant_factor, kwargs_formula=kf, **kwargs)

    @property
    def parity(self):
        return (1, 1)   # SpinOmega: transform_ident for both (covariant.py:770-771)

    def needed_channels(self, system):
        return spin_current_channels(self.spin_current_type, _external(self.kwargs_formula, system))

    def band_values(self, data_K):
        ext = _external(self.kwargs_formula, data_K.system)
        p = spin_current_params(data_K, self.spin_current_type, ext, self.degen_thresh, self.degen_Kramers)
        out = data_K.torch.empty((data_K.nk, data_K.num_wann, 27), dtype=data_K.torch.float64,
                                 device=data_K.ctx.device)
        data_K.ctx.tabulate_spin_berry(data_K.E_dev, data_K.Xk, data_K.prec, p, out, per_band=True)
        return out


class Hall_classic_FermiSurf(StaticCalculator):
    r"""Classic Hall conductivity (:math:`S/m/T`), Fermi surface (static.py:407-424)

        | Output: :math:`\sigma_{\alpha\beta :\mu} = e^3/\hbar^2 \tau^2 \epsilon_{\gamma\mu\rho} \int [dk] v_\alpha \partial_\rho v_\beta v_\gamma f'`
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta :\mu} E_\beta B_\mu`"""
    formula = _lib.FORMULA_HALL_SURF
    ndim = 2
    fder = 1

    def __init__(self, constant_factor=factors.factor_hall_classic, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class Hall_classic_FermiSea(StaticCalculator):
    r"""Classic Hall conductivity (:math:`S/m/T`), Fermi sea (static.py:427-446)

        | Output: :math:`\sigma_{\alpha\beta :\mu} = -e^3/\hbar^2 \tau^2 \epsilon_{\gamma\mu\rho} \int [dk] \partial_\gamma v_\alpha \partial_\rho v_\beta f`
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta :\mu} E_\beta B_\mu`"""
    formula = _lib.FORMULA_HALL_SEA
    ndim = 2
    fder = 0

    def __init__(self, constant_factor=factors.factor_hall_classic, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class AHC_Zeeman_spin(StaticCalculator):
    r"""Zeeman (spin) correction to the AHC (:math:`S/m/T`), Fermi surface (static.py:605-615)

        | Output: :math:`ZAHC^{spin}_{\alpha\beta :\mu} = e^2/\hbar \int [dk] \Omega_\delta s_\mu f'`"""
    formula = _lib.FORMULA_OMEGASPIN
    ndim = 2
    fder = 1

    def __init__(self, constant_factor=factors.fac_spin_Z * factors.factor_ahc, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class OmegaOmega(StaticCalculator):
    """Fermi-surface integral of Omega_a Omega_b (static.py:618-623)."""
    formula = _lib.FORMULA_OMEGAOMEGA
    ndim = 2
    fder = 1


class AHC_Zeeman_orb(StaticCalculator):
    r"""Zeeman (orbital) correction to the AHC (:math:`S/m/T`), Fermi surface (static.py:626-644)

        | Output: :math:`ZAHC^{orb}_{\alpha\beta :\mu} = e^2/\hbar \int [dk] \Omega_\delta m_\mu f'`,
        | :math:`m = H + G - 2E_f \Omega`"""
    formula = _lib.FORMULA_OMEGAHPLUS
    ndim = 2
    fder = 1

    def __init__(self, constant_factor=factors.fac_orb_Z * factors.factor_ahc, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)
        self.OmegaOmega = OmegaOmega(constant_factor=constant_factor, print_comment=False, **kwargs)

    def needed_channels(self, system):
        return super().needed_channels(system) | self.OmegaOmega.needed_channels(system)

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        Hplus = super().__call__(data_K)
        Omega = self.OmegaOmega(data_K).mul_array(self.Efermi)
        if self.k_resolved:
            Hplus.add(-2 * Omega)
            return Hplus
        return Hplus - 2 * Omega


def der3e_bands(data_K, degen_thresh, degen_Kramers, mode):
    """Der3E per band [nk, nw, 27] (wbmi_der3e_bands; mode 0: group means, 1: per-band shares)."""
    torch = data_K.torch
    sl = data_K._slots
    v0, w0, w3 = sl[("Ham", 1)], sl[("Ham", 2)], sl[("Ham", 3)]
    out = torch.empty((data_K.nk, data_K.num_wann, 27), dtype=torch.float64, device=data_K.ctx.device)
    data_K.ctx.der3e_bands(data_K.E_dev, data_K.Xk, data_K.prec, [v0 + i for i in range(3)], [w0 + i for i in range(9)],
                           [w3 + i for i in range(27)], degen_thresh, degen_Kramers, mode, out)
    return out


class NLDrude_FermiSea(StaticCalculator):
    r"""Nonlinear Drude conductivity (:math:`S^2/A`), Fermi sea (static.py:519-529)

        | Output: :math:`\sigma_{\alpha\beta\gamma} = -e^3/\hbar^2 \tau^2 \int [dk] \partial_{\beta\gamma} v_\alpha f`
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta\gamma} E_\beta E\gamma`

    The third derivative of the band energies (Der3E) per band (wbmi_der3e_bands) is binned by the
    Fermi scan as caller values (WBMI_FORMULA_BANDVALUES), 9 components per launch."""
    formula = _lib.FORMULA_BANDVALUES
    ndim = 3
    fder = 0

    def __init__(self, constant_factor=factors.factor_nldrude, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)

    @property
    def parity(self):
        return (-1, -1)   # Der3E: transform_odd for both (covariant.py:135-136)

    def needed_channels(self, system):
        return {("Ham", 0), ("Ham", 1), ("Ham", 2), ("Ham", 3)}

    def band_values(self, data_K):
        return der3e_bands(data_K, self.degen_thresh, self.degen_Kramers, 1)


class NLDrude_FermiSurf(StaticCalculator):
    r"""Nonlinear Drude conductivity (:math:`S^2/A`), Fermi surface (static.py:532-542)

        | Output: :math:`\sigma_{\alpha\beta\gamma} = e^3/\hbar^2 \tau^2 \int [dk] \partial_\beta v_\alpha v_\gamma f'`
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta\gamma} E_\beta E\gamma`"""
    formula = _lib.FORMULA_MASSVEL
    ndim = 3
    fder = 1

    def __init__(self, constant_factor=factors.factor_nldrude, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class NLDrude_Fermider2(StaticCalculator):
    r"""Nonlinear Drude conductivity (:math:`S^2/A`), second derivative of the occupation
    (static.py:545-557)

        | Output: :math:`\sigma_{\alpha\beta\gamma} = -e^3/\hbar^2 \tau^2 \int [dk] v_\beta v_\alpha v_\gamma f''`
        | Instruction: :math:`j_\alpha = \sigma_{\alpha\beta\gamma} E_\beta E\gamma`"""
    formula = _lib.FORMULA_VELVELVEL
    ndim = 3
    fder = 2

    def __init__(self, constant_factor=factors.factor_nldrude / 2, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class BerryDipole_FermiSurf(StaticCalculator):
    r"""Berry curvature dipole (dimensionless), Fermi surface

        | Output: :math:`D_{\beta\delta} = -\int [dk] v_\beta \Omega_\delta f'`"""
    formula = _lib.FORMULA_VELOMEGA
    ndim = 2
    fder = 1


class NLAHC_FermiSurf(BerryDipole_FermiSurf):
    r"""Nonlinear anomalous Hall conductivity (:math:`S^2/A`), Fermi surface"""

    def __init__(self, constant_factor=factors.factor_nlahc, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class BerryDipole_FermiSea(StaticCalculator):
    r"""Berry curvature dipole (dimensionless), Fermi sea

        | Output: :math:`D_{\beta\delta} = \int [dk] \partial_\beta \Omega_\delta f`"""
    formula = _lib.FORMULA_DEROMEGA
    ndim = 2
    fder = 0

    def __call__(self, data_K):
        # swap axes to match eq. (29) of DOI:10.1038/s41524-021-00498-5 (static.py:483-487)
        return super().__call__(data_K).swapaxes(1, 2)


class NLAHC_FermiSea(BerryDipole_FermiSea):
    r"""Nonlinear anomalous Hall conductivity (:math:`S^2/A`), Fermi sea"""

    def __init__(self, constant_factor=factors.factor_nlahc, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


class Morb(StaticCalculator):
    r"""Orbital magnetic moment per unit cell (:math:`\mu_B`)

        | Output: :math:`M = -\int [dk] (H + G - 2E_f \cdot \Omega) f`"""
    formula = _lib.FORMULA_MORB_HPM
    ndim = 1
    fder = 0
    sign = 1

    def __init__(self, constant_factor=-factors.eV_au / factors.bohr ** 2, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)
        self.AHC = AHC(constant_factor=constant_factor, print_comment=False, **kwargs)

    def needed_channels(self, system):
        return super().needed_channels(system) | self.AHC.needed_channels(system)

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        Hplus = super().__call__(data_K)
        Omega = self.AHC(data_K).mul_array(self.Efermi, axes=0)
        Hplus.add(-2 * Omega)
        return Hplus * data_K.cell_volume


class GME_orb_FermiSurf(StaticCalculator):
    r"""Gyrotropic tensor, orbital part, as a Fermi-surface integral (calculators/static.py:266-281)

        | Output: :math:`K^{orb}_{\alpha :\mu} = \int [dk] v_\alpha m^{orb}_\mu f'`,
        | :math:`m^{orb} = H + G - 2E_f \cdot \Omega`"""
    formula = _lib.FORMULA_VELHPLUS
    ndim = 2
    fder = 1
    sign = 1

    def __init__(self, constant_factor=factors.factor_gme_orb, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)
        self.BerryDipole = BerryDipole_FermiSurf(constant_factor=constant_factor, print_comment=False, **kwargs)

    def needed_channels(self, system):
        return super().needed_channels(system) | self.BerryDipole.needed_channels(system)

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        Hplus = super().__call__(data_K)
        Omega = self.BerryDipole(data_K).mul_array(self.Efermi)
        if self.k_resolved:
            Hplus.add(-2 * Omega)
            return Hplus
        return Hplus - 2 * Omega


class GME_orb_FermiSea(StaticCalculator):
    r"""Gyrotropic tensor, orbital part, as a Fermi-sea integral (calculators/static.py:284-300)

        | Output: :math:`K^{orb}_{\alpha :\mu} = -\int [dk] \partial_\alpha m_\mu f`,
        | :math:`m = H + G - 2E_f \cdot \Omega`"""
    formula = _lib.FORMULA_DERMORB
    ndim = 2
    fder = 0
    sign = 1

    def __init__(self, constant_factor=factors.factor_gme_orb, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)
        self.BerryDipole = BerryDipole_FermiSea(constant_factor=constant_factor, print_comment=False, **kwargs)

    def needed_channels(self, system):
        return super().needed_channels(system) | self.BerryDipole.needed_channels(system)

    def __call__(self, data_K):
        data_K.ensure(self.needed_channels(data_K.system))
        Hplus = super().__call__(data_K).swapaxes(1, 2)
        Omega = self.BerryDipole(data_K).mul_array(self.Efermi)
        if self.k_resolved:
            Hplus.add(-2 * Omega)
            return Hplus
        return Hplus - 2 * Omega


class Spin(StaticCalculator):
    r"""Spin per unit cell (dimensionless) (calculators/static.py:217-228)

       | Output: :math:`\int [dk] s f`"""
    formula = _lib.FORMULA_SPIN
    ndim = 1
    fder = 0

    def __call__(self, data_K):
        return super().__call__(data_K) * data_K.cell_volume


class GME_spin_FermiSea(StaticCalculator):
    r"""Gyrotropic tensor, spin part, as a Fermi-sea integral (calculators/static.py:322-335)

        | Output: :math:`K^{spin}_{\alpha :\mu} = -\int [dk] \partial_\alpha s_\mu f`"""
    formula = _lib.FORMULA_DERSPIN
    ndim = 2
    fder = 0

    def __init__(self, constant_factor=factors.factor_gme_spin, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)

    def __call__(self, data_K):
        # swap axes as eq. (29) of DOI:10.1038/s41524-021-00498-5 (static.py:331-335)
        return super().__call__(data_K).swapaxes(1, 2)


class GME_spin_FermiSurf(StaticCalculator):
    r"""Gyrotropic tensor, spin part, as a Fermi-surface integral (calculators/static.py:338-349)

        | Output: :math:`K^{spin}_{\alpha :\mu} = \tau \int [dk] v_\alpha s_\mu f'`"""
    formula = _lib.FORMULA_VELSPIN
    ndim = 2
    fder = 1

    def __init__(self, constant_factor=factors.factor_gme_spin, **kwargs):
        super().__init__(constant_factor=constant_factor, **kwargs)


__all__ = ["Calculator", "StaticCalculator", "DOS", "CumDOS", "AHC", "BerryDipole_FermiSurf", "NLAHC_FermiSurf",
           "BerryDipole_FermiSea", "NLAHC_FermiSea", "Morb", "GME_orb_FermiSurf", "GME_orb_FermiSea",
           "GME_spin_FermiSea", "GME_spin_FermiSurf", "Spin", "Ohmic_FermiSea", "Ohmic_FermiSurf",
           "Hall_classic_FermiSurf", "Hall_classic_FermiSea", "NLDrude_FermiSurf", "NLDrude_Fermider2", "NLDrude_FermiSea", "SHC",
           "AHC_Zeeman_spin", "OmegaOmega", "AHC_Zeeman_orb",
           "VoidResult"]
