"""GPU parity of the tetrahedron method (tetra=True): corner energies, the
tetrahedron-weighted scan (wbmi_static_scan_tetra) and its finalisation.

* the reference's fixtures: Chiral model with C3z, irreducible K-points and
  symmetrisation (`Chiral_tetra-*`, tests/test_run.py:227-238,835-845), and the
  Haldane C3z refinement with adpt_mesh 3 (`Haldane-*-PythTB-adpt3-tetra-sym`,
  tests/test_run.py:730-735), at the reference's bar 1e-8 * max|ref|;
* the oracle on a random Hermitian 8-band model with external terms: non-additive
  formulas (Morb, GME sea), the Fermi-surface form (fder 1) and fder 2, 3 on
  unevenly spaced Fermi levels, at 1e-10 * max (FP64) and 2e-6 (mixed form).
"""
import os
import sys

import numpy as np
import pytest

from oracle import wb_oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def wb():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import wannierberri_amd
    return wannierberri_amd


@pytest.fixture(scope="module")
def results_run():
    with np.load(os.path.join(GOLDEN, "results_run.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("fp32", 1e-5)])
def test_chiral_tetra(wb, results_run, precision, tol):
    say(f"test_chiral_tetra {precision}")
    st = wb.calculators.static
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    s.set_pointgroup(["C3z"])
    Ef = np.linspace(-5, 8, 27)
    calcs = {"ahc": st.AHC(Efermi=Ef, tetra=True),
             "berry_dipole": st.BerryDipole_FermiSea(Efermi=Ef, tetra=True, use_factor=False),
             "berry_dipole_fsurf": st.BerryDipole_FermiSurf(Efermi=Ef, tetra=True, use_factor=False),
             "dos": st.DOS(Efermi=Ef, tetra=True),
             "cumdos": st.CumDOS(Efermi=Ef, tetra=True),
             "ohmic": st.Ohmic_FermiSea(Efermi=Ef, tetra=True),
             "ohmic_fsurf": st.Ohmic_FermiSurf(Efermi=Ef, tetra=True),
             "hall_classic": st.Hall_classic_FermiSea(Efermi=Ef, tetra=True),
             "hall_classic_fsurf": st.Hall_classic_FermiSurf(Efermi=Ef, tetra=True),
             "der3e": st.NLDrude_FermiSea(Efermi=Ef, tetra=True)}
    grid = wb.Grid(system=s, NK=[10, 10, 4], NKFFT=[5, 5, 2])
    r = wb.run(s, grid, calcs, use_irred_kpt=True, symmetrize=True, parameters_K={"precision": precision},
               save_files=False, verbose=False)
    for k in calcs:
        err = rel_err(r.results[k].data, results_run["chiral_tetra_" + k])
        say(f"  {k}: {err:.2e}")
        assert err < tol, k


def test_haldane_tetra_refinement(wb, results_run):
    say("test_haldane_tetra_refinement")
    st = wb.calculators.static
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_haldane.npz"))
    s.set_pointgroup(["C3z"])
    Ef = np.linspace(-3, 3, 11)
    grid = wb.Grid(system=s, NK=[10, 10, 1], NKFFT=[5, 5, 1])
    for it in (0, 1):
        calcs = {"ahc": st.AHC(Efermi=Ef, tetra=True), "dos": st.DOS(Efermi=Ef, tetra=True)}
        r = wb.run(s, grid, calcs, adpt_num_iter=it, adpt_mesh=3, use_irred_kpt=True, symmetrize=True,
                   save_files=False, verbose=False)
        for k in calcs:
            assert rel_err(r.results[k].data, results_run[f"haldane_adpt3_tetra_{k}_iter{it}"]) < 1e-8, (k, it)


def _oracle_tetra(osys, calcs, NK, NKFFT):
    NKdiv, NKFFT, blocks = O.grid_blocks(NK, NKFFT)
    tot = {}
    for dK, factor in blocks:
        d = O.OracleDataK(osys, dK=dK, NKFFT=NKFFT, dK_block=1. / NKdiv)
        for k, c in calcs.items():
            r = c(d) * factor
            tot[k] = r if k not in tot else tot[k] + r
    return tot


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 2e-6)])
def test_tetra_random_vs_oracle(wb, precision, tol):
    say(f"test_tetra_random_vs_oracle {precision}")
    st = wb.calculators.static
    from wannierberri_amd import models
    s = models.synthetic(num_wann=8, Rmax=2, seed=11, with_BBCC=True)
    o = O.OracleSystem(s.real_lattice, s.rvec.iRvec, s.wannier_centers_cart,
                       {k: s.get_R_mat(k) for k in ("Ham", "AA", "BB", "CC")})
    E0 = np.linalg.eigvalsh(s.get_R_mat("Ham")[np.where((s.rvec.iRvec == 0).all(1))[0][0]])
    E = np.linspace(E0.min() - 1, E0.max() + 1, 23)
    Ef = np.sort(np.concatenate([E, E[:-1] + 0.37 * (E[1] - E[0])]))   # uneven spacing
    gpu = {"morb": st.Morb(Efermi=Ef, tetra=True),
           "gme_sea": st.GME_orb_FermiSea(Efermi=Ef, tetra=True),
           "bcd_surf": st.BerryDipole_FermiSurf(Efermi=Ef, tetra=True),
           "d2": st._DOS(fder=2, Efermi=Ef, tetra=True),
           "d3": st._DOS(fder=3, Efermi=Ef, tetra=True)}
    ora = {"morb": O.Morb(Ef, tetra=True),
           "gme_sea": O.GME_orb_FermiSea(Ef, tetra=True),
           "bcd_surf": O.BerryDipole_FermiSurf(Ef, tetra=True),
           "d2": O.CumDOS(Ef, fder=2, tetra=True),
           "d3": O.CumDOS(Ef, fder=3, tetra=True)}
    NK, NKFFT = [4, 4, 4], [2, 2, 2]
    grid = wb.Grid(system=s, NK=NK, NKFFT=NKFFT)
    r = wb.run(s, grid, gpu, parameters_K={"precision": precision}, save_files=False, verbose=False)
    ref = _oracle_tetra(o, ora, NK, NKFFT)
    for k in gpu:
        err = rel_err(r.results[k].data, ref[k].data)
        say(f"  {k}: {err:.2e}")
        assert err < tol, k


def test_tetra_hole_like_vs_oracle(wb):
    """hole_like with the tetrahedron method (weights 1 - occupation and the group of bands
    above the range, tetrahedron.py:165-172,246-253): AHC (additive), Morb (non-additive) and
    CumDOS on a random Hermitian model vs the oracle, with a window that leaves bands above."""
    say("test_tetra_hole_like_vs_oracle")
    st = wb.calculators.static
    from wannierberri_amd import models
    s = models.synthetic(num_wann=8, Rmax=2, seed=11, with_BBCC=True)
    o = O.OracleSystem(s.real_lattice, s.rvec.iRvec, s.wannier_centers_cart,
                       {k: s.get_R_mat(k) for k in ("Ham", "AA", "BB", "CC")})
    E0 = np.linalg.eigvalsh(s.get_R_mat("Ham")[np.where((s.rvec.iRvec == 0).all(1))[0][0]])
    Ef = np.linspace(E0.min() - 0.5, 0.5 * (E0.min() + E0.max()), 17)
    gpu = {"ahc": st.AHC(Efermi=Ef, tetra=True, hole_like=True),
           "morb": st.Morb(Efermi=Ef, tetra=True, hole_like=True),
           "cumdos": st.CumDOS(Efermi=Ef, tetra=True, hole_like=True)}
    ora = {"ahc": O.AHC(Ef, tetra=True, hole_like=True), "morb": O.Morb(Ef, tetra=True, hole_like=True),
           "cumdos": O.CumDOS(Ef, tetra=True, hole_like=True)}
    NK, NKFFT = [4, 4, 4], [2, 2, 2]
    r = wb.run(s, wb.Grid(system=s, NK=NK, NKFFT=NKFFT), gpu, save_files=False, verbose=False)
    ref = _oracle_tetra(o, ora, NK, NKFFT)
    for k in gpu:
        err = rel_err(r.results[k].data, ref[k].data)
        say(f"  {k}: {err:.2e}")
        assert err < 1e-10, k


def test_tetra_select_bands_vs_oracle(wb):
    """select_bands with the tetrahedron method (tetrahedron.py:147-149,238-241): DOS and the
    Fermi-surface Berry curvature dipole with pairs of bands as groups (degen_Kramers), the even
    bands and the odd bands listed twice, vs the oracle (FP64, 1e-10); the Fermi sea refuses it."""
    say("test_tetra_select_bands_vs_oracle")
    st = wb.calculators.static
    from wannierberri_amd import models
    s = models.synthetic(num_wann=8, Rmax=2, seed=11, with_BBCC=True)
    o = O.OracleSystem(s.real_lattice, s.rvec.iRvec, s.wannier_centers_cart,
                       {k: s.get_R_mat(k) for k in ("Ham", "AA", "BB", "CC")})
    E0 = np.linalg.eigvalsh(s.get_R_mat("Ham")[np.where((s.rvec.iRvec == 0).all(1))[0][0]])
    Ef = np.linspace(E0.min() - 1, E0.max() + 1, 23)
    NK, NKFFT = [4, 4, 4], [2, 2, 2]
    for sel in (list(range(0, 8, 2)), list(range(1, 8, 2)) * 2):
        kw = dict(tetra=True, degen_Kramers=True, select_bands=sel)
        gpu = {"dos": st.DOS(Efermi=Ef, **kw), "bcd": st.BerryDipole_FermiSurf(Efermi=Ef, **kw)}
        ora = {"dos": O.DOS(Ef, **kw), "bcd": O.BerryDipole_FermiSurf(Ef, **kw)}
        r = wb.run(s, wb.Grid(system=s, NK=NK, NKFFT=NKFFT), gpu, save_files=False, verbose=False)
        ref = _oracle_tetra(o, ora, NK, NKFFT)
        for k in gpu:
            err = rel_err(r.results[k].data, ref[k].data)
            say(f"  {sel[:3]} {k}: {err:.2e}")
            assert np.abs(ref[k].data).max() > 0 and err < 1e-10, k
    with pytest.raises(NotImplementedError, match="Fermi sea"):
        wb.run(s, wb.Grid(system=s, NK=NK, NKFFT=NKFFT), {"c": st.CumDOS(Efermi=Ef, tetra=True, select_bands=[0])},
               save_files=False, verbose=False)


def test_tetra_erange_vs_oracle(wb):
    """Emin / Emax with the tetrahedron method (tetrahedron.py:246-258): the Fermi sea leaves out
    the bands entirely below Emin (so its group has "out" bands on both sides), the hole-like
    group those entirely above Emax.  AHC (additive), Morb (non-additive) and CumDOS on a random
    Hermitian model vs the oracle, FP64 at 1e-10."""
    say("test_tetra_erange_vs_oracle")
    st = wb.calculators.static
    from wannierberri_amd import models
    s = models.synthetic(num_wann=8, Rmax=2, seed=11, with_BBCC=True)
    o = O.OracleSystem(s.real_lattice, s.rvec.iRvec, s.wannier_centers_cart,
                       {k: s.get_R_mat(k) for k in ("Ham", "AA", "BB", "CC")})
    E0 = np.linalg.eigvalsh(s.get_R_mat("Ham")[np.where((s.rvec.iRvec == 0).all(1))[0][0]])
    lo, hi = E0.min(), E0.max()
    NK, NKFFT = [4, 4, 4], [2, 2, 2]
    for name, Ef, kw in (("sea", np.linspace(0.5 * (lo + hi), hi + 0.5, 17), dict(Emin=lo + 0.3 * (hi - lo))),
                         ("hole", np.linspace(lo - 0.5, 0.5 * (lo + hi), 17),
                          dict(hole_like=True, Emax=hi - 0.3 * (hi - lo)))):
        say(f"  {name}: {kw}")
        gpu = {"ahc": st.AHC(Efermi=Ef, tetra=True, **kw), "morb": st.Morb(Efermi=Ef, tetra=True, **kw),
               "cumdos": st.CumDOS(Efermi=Ef, tetra=True, **kw)}
        ora = {"ahc": O.AHC(Ef, tetra=True, **kw), "morb": O.Morb(Ef, tetra=True, **kw),
               "cumdos": O.CumDOS(Ef, tetra=True, **kw)}
        r = wb.run(s, wb.Grid(system=s, NK=NK, NKFFT=NKFFT), gpu, save_files=False, verbose=False)
        ref = _oracle_tetra(o, ora, NK, NKFFT)
        full = _oracle_tetra(o, {"cumdos": O.CumDOS(Ef, tetra=True, hole_like=kw.get("hole_like", False))}, NK, NKFFT)
        assert np.abs(full["cumdos"].data - ref["cumdos"].data).max() > 0.1   # the bound removes bands
        for k in gpu:
            err = rel_err(r.results[k].data, ref[k].data)
            say(f"  {k}: {err:.2e}")
            assert err < 1e-10, (name, k)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("fp32", 1e-5)])
def test_chiral_tetragrid(wb, results_run, precision, tol):
    """The tetrahedral grid (GridTetra(length=8, NKFFT=[5, 5, 2]), one tetrahedron per
    k-point, 4 corner energies) against the reference's Chiral_tetragrid fixtures
    (tests/test_run.py:958-968)."""
    say(f"test_chiral_tetragrid {precision}")
    st = wb.calculators.static
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_chiral_left.npz"))
    s.set_pointgroup(["C3z"])
    Ef = np.linspace(-5, 8, 27)
    calcs = {"ahc": st.AHC(Efermi=Ef, tetra=True),
             "berry_dipole": st.BerryDipole_FermiSea(Efermi=Ef, tetra=True, use_factor=False),
             "berry_dipole_fsurf": st.BerryDipole_FermiSurf(Efermi=Ef, tetra=True, use_factor=False),
             "dos": st.DOS(Efermi=Ef, tetra=True),
             "cumdos": st.CumDOS(Efermi=Ef, tetra=True)}
    grid = wb.grid.GridTetra(s, length=8, NKFFT=[5, 5, 2])
    r = wb.run(s, grid, calcs, use_irred_kpt=True, symmetrize=True, parameters_K={"precision": precision},
               save_files=False, verbose=False)
    for k in calcs:
        err = rel_err(r.results[k].data, results_run["chiral_tetragrid_" + k])
        say(f"  {k}: {err:.2e}")
        assert err < tol, k


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("fp32", 1e-5)])
def test_te_qe_gme(wb, precision, tol):
    """Tellurium (QE): tetra=True, point group C3z x C2x x TR, irreducible K-points,
    NK [3, 3, 4], NKFFT [1, 1, 4] (tests/test_run.py:1102-1119) against the reference's
    Te_QE fixtures: GME_orb sea and surface, BerryDipole_FermiSea, NLAHC_FermiSea, DOS,
    cumulative DOS, AHC_Zeeman_spin; NLDrude (Fermi surface, f'') vanish by time reversal.  Bar 1e-8 x max|ref| (FP64); the mixed form's error is printed."""
    say(f"test_te_qe_gme {precision}")
    st = wb.calculators.static
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_te_qe.npz"))
    s.set_pointgroup(["C3z", "C2x", "TimeReversal"])
    with np.load(os.path.join(GOLDEN, "results_te_qe.npz"), allow_pickle=False) as f:
        ref = {k: f[k] for k in f.files}
    Ef = np.linspace(4, 8, 11)
    calcs = {"dos": st.DOS(Efermi=Ef, tetra=True),
             "cumdos": st.CumDOS(Efermi=Ef, tetra=True),
             "berry_dipole": st.NLAHC_FermiSea(Efermi=Ef, tetra=True),
             "BerryDipole_FermiSea": st.BerryDipole_FermiSea(Efermi=Ef, tetra=True),
             "GME_orb_FermiSea": st.GME_orb_FermiSea(Efermi=Ef, tetra=True),
             "GME_orb_FermiSurf": st.GME_orb_FermiSurf(Efermi=Ef, tetra=True),
             "NLDrude_FermiSurf": st.NLDrude_FermiSurf(Efermi=Ef, tetra=True),
             "NLDrude_Fermider2": st.NLDrude_Fermider2(Efermi=Ef, tetra=True),
             "AHC_Zeeman_spin": st.AHC_Zeeman_spin(Efermi=Ef, tetra=True)}
    grid = wb.Grid(system=s, NK=[3, 3, 4], NKFFT=[1, 1, 4])
    r = wb.run(s, grid, calcs, use_irred_kpt=True, symmetrize=True, parameters_K={"precision": precision},
               save_files=False, verbose=False)
    for k in calcs:
        if k.startswith("NLDrude"):
            continue
        err = rel_err(r.results[k].data, ref["te_" + k])
        say(f"  {k}: {err:.2e}")
        assert err < tol, k
    # the nonlinear Drude tensors are odd under time reversal, which is in the group: the
    # symmetrised results vanish; the reference's files hold its rounding noise (~1e-19)
    # (relative to the same tensors on 64 random k-points, a set with no symmetry)
    nl = {"NLDrude_FermiSurf": st.NLDrude_FermiSurf(Efermi=Ef), "NLDrude_Fermider2": st.NLDrude_Fermider2(Efermi=Ef)}
    r_ns = wb.run_klists(s, [np.random.default_rng(3).random((64, 3))], nl, parameters_K={"precision": precision})
    for k in nl:
        got, scale = r.results[k].data, np.abs(r_ns.results[k].data).max()
        say(f"  {k}: max {np.abs(got).max():.1e}, unsymmetrised {scale:.1e}, "
            f"reference file {np.abs(ref['te_' + k]).max():.1e}")
        assert got.shape == ref["te_" + k].shape and scale > 0
        assert np.abs(got).max() < 1e-12 * scale and np.abs(ref["te_" + k]).max() < 1e-12 * scale
