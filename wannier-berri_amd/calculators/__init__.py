"""Calculators of the MI355X path (static Fermi-level scans)."""
from . import static
from .static import Calculator, StaticCalculator
