"""wannierberri_amd: MI355X-native engine for the Berry-quantity path of wannier-berri.

The API surface of the path is the reference's: System_R -> Data_K ->
calculators / run().  All numerics run in the HIP library libwbmi.so
(include/wbmi.h); see `wannierberri_amd.build_lib` to build it.
"""
from . import autotune, calculators, factors, smoother, symmetry
from .calculators import static
from .data_k import Data_K
from .grid import Grid, KpointBZparallel
from .path import KpointBZpath, Path
from .result import EnergyResult, ResultDict, VoidResult
from .run_grid import KlistGraph, run, run_klists
from .smoother import get_smoother
from .system import System_R

__version__ = "0.1.0"
__all__ = ["System_R", "Grid", "Path", "Data_K", "run", "run_klists", "calculators", "EnergyResult", "ResultDict", "VoidResult",
           "KpointBZparallel", "factors", "symmetry", "smoother", "get_smoother", "KlistGraph"]
