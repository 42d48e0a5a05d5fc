"""Calculators of the MI355X path (static Fermi-level scans, optical conductivity)."""
from . import dynamic, static
from .static import Calculator, StaticCalculator
