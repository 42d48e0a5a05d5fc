"""Binding for the reference package: its own run() and calculators on this engine's Data_K.

The reference builds one `data_k_class(system, dK=..., grid=..., Kpoint=..., **parameters_K)`
per K-block (run_grid.py:139-141,259) and its formulas read `E_K`, `UU_K` and
`Xbar(name, der)` from it (data_K/data_K.py:130-334, data_K/data_K_R.py:69-97).
`data_k_class_for(Data_K_R)` returns a subclass of the reference's class whose
three heavy members come from the GPU: R->k, eigh and the rotation run here,
everything else (D_H, dEig_inv, the formulas, the Fermi scan) stays the reference's.

    from wannierberri.data_K.data_K_R import Data_K_R
    from wannierberri_amd.reference_binding import data_k_class_for
    wberri.run(system, grid, calculators, data_k_class=data_k_class_for(Data_K_R))
"""
import numpy as np

from .data_k import CHANNEL_NCOMP, Data_K
from .system import System_R

# R-space matrices the engine knows; the reference system may hold any subset
MATRICES = ("Ham", "AA", "BB", "CC", "SS", "SA", "SHA", "SR", "SH", "SHR")


def system_from_reference(system):
    """This engine's System_R with every R-space matrix of a reference System_R
    (`_XX_R` entries `[nR, nw, nw, ...]`, system/system_R.py:117-175), its lattice,
    Wannier centres, periodicity and force_internal_terms_only."""
    mats = {k: np.asarray(system.get_R_mat(k)) for k in MATRICES if system.has_R_mat(k)}
    own = System_R(np.asarray(system.real_lattice), np.asarray(system.rvec.iRvec), np.asarray(system.wannier_centers_cart),
                   mats, periodic=np.asarray(system.periodic),
                   force_internal_terms_only=bool(getattr(system, "force_internal_terms_only", False)))
    return own


def data_k_class_for(base, precision="fp64"):
    """Subclass of the reference's Data_K class `base` whose E_K, UU_K and Xbar come
    from the GPU (`precision` 'fp64' or 'fp32')."""

    class Data_K_MI355X(base):
        def __init__(self, system, **parameters):
            super().__init__(system, **parameters)
            # the base class has settled dK / k_list (a path K-point gives a k-list, data_K.py:84-92)
            cache = system.__dict__.setdefault("_mi355x_system", {})
            if "own" not in cache:
                cache["own"] = system_from_reference(system)
            self._gpu = Data_K(cache["own"], dK=self.dK, grid=parameters.get("grid"), k_list=self.k_list,
                               precision=precision)

        @property
        def E_K(self):                      # np.linalg.eigh of data_K.py:211-218
            return self._gpu.E_K

        @property
        def UU_K(self):                     # data_K.py:309-322
            return self._gpu.UU_K

        def Xbar(self, name, der=0):        # data_K_R.py:69-89
            if (name, der) not in CHANNEL_NCOMP:
                return super().Xbar(name, der)
            return self._gpu.Xbar(name, der)

    return Data_K_MI355X
