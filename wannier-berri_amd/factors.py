"""Physical unit factors of the Berry-quantity calculators.

Same definitions as wannierberri/factors.py:6-28 (SciPy CODATA constants), so
that results agree with the reference to the last bit of the constants.
"""
import numpy as np
from scipy.constants import elementary_charge, hbar, electron_mass, physical_constants, angstrom

TAU_UNIT = 1E-15
bohr_magneton = elementary_charge * hbar / (2 * electron_mass)
bohr = physical_constants['Bohr radius'][0] / angstrom
eV_au = physical_constants['electron volt-hartree relationship'][0]
factor_morb_evA2_to_muB = -(elementary_charge * angstrom ** 2) * elementary_charge / (2 * hbar) / bohr_magneton
factor_gme_orb = factor_morb_evA2_to_muB * bohr_magneton / angstrom ** 2
factor_gme_spin = -bohr_magneton / angstrom ** 2   # factors.py:22
factor_ahc = -(elementary_charge ** 2 / hbar / angstrom)
factor_opt = -factor_ahc   # factors.py:36
factor_shc = -factor_ahc   # factors.py:37
factor_shift_current = hbar / elementary_charge * np.pi * elementary_charge ** 3 / (4 * hbar ** 2)  # factors.py:50
factor_injection_current = -np.pi * elementary_charge ** 3 / (hbar ** 2) * TAU_UNIT               # factors.py:51
factor_nlahc = elementary_charge ** 3 / hbar ** 2 * TAU_UNIT
# tau^1 E^1, velocities converted from eV*m to m/s (factors.py:25-26)
factor_ohmic = elementary_charge ** 2 / hbar / angstrom * TAU_UNIT * elementary_charge / hbar
fac_orb_Z = elementary_charge / 2 / hbar * angstrom ** 2   # m_orb * B in eV (factors.py:16)
fac_spin_Z = hbar / (2 * electron_mass)                      # m_spin * B in eV (factors.py:17)
# tau^2 E^1 B^1 and tau^2 E^2 B^0, velocities in m/s (factors.py:28-31)
factor_hall_classic = -(elementary_charge ** 3 / hbar ** 2 * angstrom * TAU_UNIT ** 2 * elementary_charge ** 2 / hbar ** 2)
factor_nldrude = -(elementary_charge ** 3 / hbar ** 2 * TAU_UNIT ** 2 * elementary_charge / hbar)
