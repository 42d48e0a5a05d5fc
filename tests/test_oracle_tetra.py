"""Pins of the oracle's tetrahedron method (no GPU).

* `tetra_occupation` der=0 against the occupied volume fraction of a tetrahedron,
  counted by sampling (uniform barycentric points), and der=1..3 against central
  differences of the order below;
* `tetra_band_weight` (12 tetrahedra around the centre of the parallelepiped) against
  the exact fraction of a cube below a plane, for a linear band, where the linear
  interpolation is exact;
* hole_like: cumulative DOS of electrons minus that of holes = number of bands;
* the whole chain (corner energies, group weights, weighted sum, symmetrisation,
  refinement) through the package's `run()` with the oracle evaluating each K-block,
  against the reference's fixtures `Chiral_tetra-*` (C3z, tests/test_run.py:835-845)
  and `Haldane-*-PythTB-adpt3-tetra-sym_iter-000*` (tests/test_run.py:730-735).
"""
import os

import numpy as np
import pytest

from oracle import wb_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def wb():
    import wannierberri_amd
    return wannierberri_amd


@pytest.fixture(scope="module")
def results_run():
    with np.load(os.path.join(GOLDEN, "results_run.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("corners", [(0.3, -0.2, 1.1, 0.45), (0.0, 0.0, 1.0, 2.0), (-1.0, 0.5, 0.5, 0.5)])
def test_tetra_occupation_volume(corners):
    rng = np.random.default_rng(7)
    lam = rng.dirichlet(np.ones(4), size=400000)
    f = lam @ np.array(corners)
    ef = np.linspace(min(corners) - 0.1, max(corners) + 0.1, 41)
    sampled = (f[None, :] < ef[:, None]).mean(axis=1)
    assert np.abs(O.tetra_occupation(ef, corners, 0) - sampled).max() < 3e-3


@pytest.mark.parametrize("der", [1, 2, 3])
def test_tetra_occupation_derivatives(der):
    corners = (0.3, -0.2, 1.1, 0.45)
    h = 1e-5
    # points away from the corner energies, where each piece is smooth
    ef = np.array([-0.1, 0.0, 0.2, 0.35, 0.4, 0.6, 0.9, 1.0])
    fd = (O.tetra_occupation(ef + h, corners, der - 1) - O.tetra_occupation(ef - h, corners, der - 1)) / (2 * h)
    assert np.abs(O.tetra_occupation(ef, corners, der) - fd).max() < 1e-5 * max(1, np.abs(fd).max())


def test_tetra_band_weight_linear_band():
    """E(k) = E0 + v.k on the cube [-1/2, 1/2]^3: the 12 tetrahedra cover the cube once
    and the interpolation is exact, so the weight is the cube's volume fraction below EF."""
    v = np.array([0.7, -0.4, 1.3])
    E0 = 0.2
    s = np.array([-0.5, 0.5])
    corners = E0 + v[0] * s[:, None, None] + v[1] * s[None, :, None] + v[2] * s[None, None, :]
    n = 120
    g = (np.arange(n) + 0.5) / n - 0.5
    E = E0 + v[0] * g[:, None, None] + v[1] * g[None, :, None] + v[2] * g[None, None, :]
    ef = np.linspace(-1.2, 1.6, 29)
    exact = np.array([(E < e).mean() for e in ef])
    w = O.tetra_band_weight(ef, E0, corners, 0)
    assert np.abs(w - exact).max() < 2e-3
    # density of states: the derivative integrates to one over the band
    efd = np.linspace(-1.5, 1.9, 20001)
    dos = O.tetra_band_weight(efd, E0, corners, 1)
    assert abs(np.trapezoid(dos, efd) - 1) < 1e-4
    assert np.array_equal(O.tetra_band_weight(ef, E0, corners, -1), 1 - w)


class _OracleBlock:
    """data_k_class for run(): the oracle evaluates the K-block, corners included."""
    osys = None

    def __init__(self, system, dK=None, grid=None, Kpoint=None, **kw):
        self.d = O.OracleDataK(self.osys, dK=dK, NKFFT=grid.FFT, dK_block=Kpoint.dK)


class _OracleCalc:
    allow_grid = True

    def __init__(self, wb, calc, tr, inv):
        self.wb, self.calc, self.tr, self.inv = wb, calc, tr, inv

    def __call__(self, data):
        r = self.calc(data.d)
        return self.wb.EnergyResult(r.Efermi, r.data, transformTR=self.tr, transformInv=self.inv)


def test_chiral_tetra(wb, results_run):
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    s.set_pointgroup(["C3z"])
    _OracleBlock.osys = O.load_golden_system(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    Ef = np.linspace(-5, 8, 27)
    calcs = {"ahc": _OracleCalc(wb, O.AHC(Ef, tetra=True), -1, 1),
             "berry_dipole": _OracleCalc(wb, O.BerryDipole_FermiSea(Ef, tetra=True, use_factor=False), 1, -1),
             "berry_dipole_fsurf": _OracleCalc(wb, O.BerryDipole_FermiSurf(Ef, tetra=True, use_factor=False), 1, -1),
             "dos": _OracleCalc(wb, O.DOS(Ef, tetra=True), 1, 1),
             "cumdos": _OracleCalc(wb, O.CumDOS(Ef, tetra=True), 1, 1),
             "ohmic": _OracleCalc(wb, O.Ohmic_FermiSea(Ef, tetra=True), 1, 1),
             "ohmic_fsurf": _OracleCalc(wb, O.Ohmic_FermiSurf(Ef, tetra=True), 1, 1),
             "hall_classic": _OracleCalc(wb, O.Hall_classic_FermiSea(Ef, tetra=True), 1, 1),
             "hall_classic_fsurf": _OracleCalc(wb, O.Hall_classic_FermiSurf(Ef, tetra=True), 1, 1),
             "der3e": _OracleCalc(wb, O.NLDrude_FermiSea(Ef, tetra=True), -1, -1)}
    grid = wb.Grid(system=s, NK=[10, 10, 4], NKFFT=[5, 5, 2])
    r = wb.run(s, grid, calcs, use_irred_kpt=True, symmetrize=True, data_k_class=_OracleBlock,
               save_files=False, verbose=False)
    for k in calcs:
        assert rel_err(r.results[k].data, results_run["chiral_tetra_" + k]) < 1e-8, k


def test_haldane_tetra_refinement(wb, results_run):
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_haldane.npz"))
    s.set_pointgroup(["C3z"])
    _OracleBlock.osys = O.load_golden_system(os.path.join(GOLDEN, "sys_haldane.npz"))
    Ef = np.linspace(-3, 3, 11)
    calcs = {"ahc": _OracleCalc(wb, O.AHC(Ef, tetra=True), -1, 1),
             "dos": _OracleCalc(wb, O.DOS(Ef, tetra=True), 1, 1)}
    grid = wb.Grid(system=s, NK=[10, 10, 1], NKFFT=[5, 5, 1])
    for it in (0, 1):
        r = wb.run(s, grid, calcs, adpt_num_iter=it, adpt_mesh=3, use_irred_kpt=True, symmetrize=True,
                   data_k_class=_OracleBlock, save_files=False, verbose=False)
        for k in calcs:
            assert rel_err(r.results[k].data, results_run[f"haldane_adpt3_tetra_{k}_iter{it}"]) < 1e-8, (k, it)


def test_tetra_hole_like_counts(wb):
    """hole_like (weights 1 - occupation, the group of bands above the range added with
    weight 1, constant factor negated): the cumulative DOS of holes H and of electrons C
    satisfy C - H = number of bands at every Fermi level (any level, bands inside or
    outside the window)."""
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    _OracleBlock.osys = O.load_golden_system(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    Ef = np.linspace(-3, 1, 9)     # the upper band stays above the window at most k
    calcs = {"c": _OracleCalc(wb, O.CumDOS(Ef, tetra=True), 1, 1),
             "h": _OracleCalc(wb, O.CumDOS(Ef, tetra=True, hole_like=True), 1, 1)}
    grid = wb.Grid(system=s, NK=[6, 6, 4], NKFFT=[3, 3, 2])
    r = wb.run(s, grid, calcs, use_irred_kpt=False, symmetrize=False, data_k_class=_OracleBlock,
               save_files=False, verbose=False)
    assert np.allclose(r.results["c"].data - r.results["h"].data, s.num_wann, rtol=0, atol=1e-12)


class _OracleBlockAny:
    """data_k_class for run() on parallelepiped or tetrahedral K-blocks."""
    osys = None

    def __init__(self, system, dK=None, grid=None, Kpoint=None, **kw):
        if hasattr(Kpoint, "vertices_fullBZ"):
            self.d = O.OracleDataK(self.osys, dK=dK, NKFFT=grid.FFT, tetra_vertices=Kpoint.vertices_fullBZ)
        else:
            self.d = O.OracleDataK(self.osys, dK=dK, NKFFT=grid.FFT, dK_block=Kpoint.dK)


def test_chiral_tetragrid(wb, results_run):
    """The tetrahedral grid (grid_tetra.py, Kpoint_tetra.py; one tetrahedron per k-point,
    tetrahedron.py:185-209) against the reference's Chiral_tetragrid fixtures
    (tests/test_run.py:958-968; C3z symmetrisation of every block's result)."""
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    s.set_pointgroup(["C3z"])
    _OracleBlockAny.osys = O.load_golden_system(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    Ef = np.linspace(-5, 8, 27)
    calcs = {"ahc": _OracleCalc(wb, O.AHC(Ef, tetra=True), -1, 1),
             "berry_dipole": _OracleCalc(wb, O.BerryDipole_FermiSea(Ef, tetra=True, use_factor=False), 1, -1),
             "berry_dipole_fsurf": _OracleCalc(wb, O.BerryDipole_FermiSurf(Ef, tetra=True, use_factor=False), 1, -1),
             "dos": _OracleCalc(wb, O.DOS(Ef, tetra=True), 1, 1),
             "cumdos": _OracleCalc(wb, O.CumDOS(Ef, tetra=True), 1, 1)}
    grid = wb.grid.GridTetra(s, length=8, NKFFT=[5, 5, 2])
    r = wb.run(s, grid, calcs, use_irred_kpt=True, symmetrize=True, data_k_class=_OracleBlockAny,
               save_files=False, verbose=False)
    for k in calcs:
        assert rel_err(r.results[k].data, results_run["chiral_tetragrid_" + k]) < 1e-8, k


TE_CALCS = {
    # name: (oracle calculator, TR parity, inversion parity) -- tests/test_run.py:154-170
    "dos": (lambda Ef: O.DOS(Ef, tetra=True), 1, 1),
    "cumdos": (lambda Ef: O.CumDOS(Ef, tetra=True), 1, 1),
    "berry_dipole": (lambda Ef: O.NLAHC_FermiSea(Ef, tetra=True), 1, -1),
    "BerryDipole_FermiSea": (lambda Ef: O.BerryDipole_FermiSea(Ef, tetra=True), 1, -1),
    "GME_orb_FermiSea": (lambda Ef: O.GME_orb_FermiSea(Ef, tetra=True), 1, -1),
    "GME_orb_FermiSurf": (lambda Ef: O.GME_orb_FermiSurf(Ef, tetra=True), 1, -1),
    "AHC_Zeeman_spin": (lambda Ef: O.AHC_Zeeman_spin(Ef, tetra=True), 1, 1),
}


def test_te_qe_gme(wb):
    """Tellurium (QE, 24 WF, 65 R, external terms from AA, BB, CC) with the tetrahedron
    method, the point group generated by C3z, C2x and time reversal, irreducible
    K-points and symmetrisation, NK [3, 3, 4], NKFFT [1, 1, 4], Efermi = linspace(4, 8, 11)
    (tests/test_run.py:1102-1119) against the reference's Te_QE fixtures.

    This pins the composition of the orbital gyrotropic tensor: GME_orb_FermiSea =
    swapaxes(DerMorb) - 2 E_F BCD_sea and GME_orb_FermiSurf = VelHplus - 2 E_F BCD_surf,
    both times factor_gme_orb (static.py:266-300), and NLAHC = BCD_sea * factor_nlahc
    (static.py:490-498): the fixtures of BerryDipole_FermiSea (factor 1) and of
    berry_dipole (NLAHC_FermiSea) differ by that factor alone."""
    path = os.path.join(GOLDEN, "sys_te_qe.npz")
    s = wb.System_R.from_arrays_npz(path)
    s.set_pointgroup(["C3z", "C2x", "TimeReversal"])
    assert s.pointgroup.size == 12
    _OracleBlock.osys = O.load_golden_system(path)
    with np.load(os.path.join(GOLDEN, "results_te_qe.npz"), allow_pickle=False) as f:
        ref = {k: f[k] for k in f.files}
    Ef = np.linspace(4, 8, 11)
    calcs = {k: _OracleCalc(wb, c(Ef), tr, inv) for k, (c, tr, inv) in TE_CALCS.items()}
    grid = wb.Grid(system=s, NK=[3, 3, 4], NKFFT=[1, 1, 4])
    r = wb.run(s, grid, calcs, use_irred_kpt=True, symmetrize=True, data_k_class=_OracleBlock,
               save_files=False, verbose=False)
    for k in calcs:
        assert np.allclose(ref["te_" + k + "__Efermi"], Ef)
        assert np.abs(ref["te_" + k]).max() > 0, k
        assert rel_err(r.results[k].data, ref["te_" + k]) < 1e-8, (k, rel_err(r.results[k].data, ref["te_" + k]))
    # NLAHC_FermiSea = BerryDipole_FermiSea x factor_nlahc (elements allowed by symmetry;
    # the forbidden ones are rounding noise in both files)
    bcd = ref["te_BerryDipole_FermiSea"]
    big = np.abs(bcd) > 1e-6 * np.abs(bcd).max()
    assert big.sum() > 20
    assert np.abs(ref["te_berry_dipole"][big] / bcd[big] / O.FACTOR_NLAHC - 1).max() < 1e-8


def test_tetra_erange_cumdos(wb):
    """Emin / Emax of the tetrahedron method (tetrahedron.py:246-258) on the Haldane model
    (bands in [-3.01, -0.58] and [0.58, 3.01]), Fermi levels in the gap: the CumDOS Fermi sea
    is the valence band, one state per k; Emin = -0.3 (above the valence band) leaves it out
    (CumDOS 0), Emin = -3.5 keeps it; the hole-like CumDOS loses the conduction band with
    Emax = 0.3 and keeps it with Emax = 3.5; infinite bounds change nothing."""
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_haldane.npz"))
    _OracleBlock.osys = O.load_golden_system(os.path.join(GOLDEN, "sys_haldane.npz"))
    Ef = np.linspace(-0.4, 0.4, 9)
    grid = wb.Grid(system=s, NK=[6, 6, 1], NKFFT=[3, 3, 1])

    def cum(**c):
        calcs = {"c": _OracleCalc(wb, O.CumDOS(Ef, tetra=True, **c), 1, 1)}
        r = wb.run(s, grid, calcs, data_k_class=_OracleBlock, save_files=False, verbose=False)
        return r.results["c"].data
    base = cum()
    assert np.allclose(base, 1.0, rtol=0, atol=1e-12)
    assert np.allclose(cum(Emin=-np.inf, Emax=np.inf), base, rtol=0, atol=1e-12)
    assert np.allclose(cum(Emin=-3.5), base, rtol=0, atol=1e-12)
    assert np.allclose(cum(Emin=-0.3), 0.0, rtol=0, atol=1e-12)
    hb = cum(hole_like=True)
    assert np.allclose(np.abs(hb), 1.0, rtol=0, atol=1e-12)
    assert np.allclose(cum(hole_like=True, Emax=3.5), hb, rtol=0, atol=1e-12)
    assert np.allclose(cum(hole_like=True, Emax=0.3), 0.0, rtol=0, atol=1e-12)


def test_tetra_select_bands_partition(wb):
    """select_bands with the tetrahedron method (tetrahedron.py:147-149,238-241) on the Haldane
    model with its two bands as one group (degen_Kramers): DOS with band 0 listed is half the
    DOS, both bands listed the whole; the Fermi sea refuses a selection."""
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_haldane.npz"))
    _OracleBlock.osys = O.load_golden_system(os.path.join(GOLDEN, "sys_haldane.npz"))
    Ef = np.linspace(-2.5, 2.5, 11)
    grid = wb.Grid(system=s, NK=[6, 6, 1], NKFFT=[3, 3, 1])

    def dos(sel, calc=O.DOS):
        c = {"c": _OracleCalc(wb, calc(Ef, tetra=True, degen_Kramers=True, select_bands=sel), 1, 1)}
        return wb.run(s, grid, c, data_k_class=_OracleBlock, save_files=False, verbose=False).results["c"].data
    full = dos(None)
    assert np.abs(full).max() > 0
    assert np.allclose(dos([0]), 0.5 * full, rtol=0, atol=1e-12 * np.abs(full).max())
    assert np.allclose(dos([0, 1]), full, rtol=0, atol=1e-12 * np.abs(full).max())
    with pytest.raises(NotImplementedError, match="Fermi sea"):
        dos([0], calc=O.CumDOS)
