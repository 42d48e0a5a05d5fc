"""Calculators of the MI355X path (static Fermi-level scans, dynamic calculators, tabulation)."""
from . import dynamic, static, tabulate
from .tabulate import TabulatorAll
from .static import Calculator, StaticCalculator
