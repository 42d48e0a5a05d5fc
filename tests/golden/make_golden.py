"""Extract the golden fixtures used by the parity tests.

Run in the development container, where the wannier-berri source tree is at
/root/reference.  It reads the reference's own committed test data with
`np.load(allow_pickle=False)` (numeric arrays only) and writes compact `.npz`
files here: model inputs in the `[R, m, n, ...]` layout and the expected
outputs.  Nothing from the reference is imported or executed.

    python tests/golden/make_golden.py
"""
import os
import numpy as np

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def _arr(path, name):
    with np.load(os.path.join(path, name + ".npz"), allow_pickle=False) as f:
        return f["arr_0"]


def system(dirname, out, keys, legacy, extra=None):
    path = os.path.join(REF, dirname)
    d = {"real_lattice": _arr(path, "real_lattice"), "iRvec": _arr(path, "iRvec").astype(np.int32)}
    if os.path.exists(os.path.join(path, "wannier_centers_cart.npz")):
        d["wannier_centers_cart"] = _arr(path, "wannier_centers_cart")
    else:
        d["wannier_centers_cart"] = _arr(path, "wannier_centers_red").dot(d["real_lattice"])
    if os.path.exists(os.path.join(path, "periodic.npz")):
        d["periodic"] = _arr(path, "periodic")
    for k in keys:
        fn = "_XX_R_" + k if os.path.exists(os.path.join(path, "_XX_R_" + k + ".npz")) else k
        a = _arr(path, fn)
        if legacy:
            a = np.transpose(a, (2, 0, 1) + tuple(range(3, a.ndim)))
        d[k] = a
    if extra:
        d.update(extra)
    np.savez_compressed(os.path.join(HERE, out), **d)


def results(pairs, out):
    d = {}
    for key, rel in pairs.items():
        with np.load(os.path.join(REF, rel), allow_pickle=False) as f:
            d[key] = f["data"]
            d[key + "__Efermi"] = f["Energies_0"]
    np.savez_compressed(os.path.join(HERE, out), **d)


def dat_columns(rel):
    """Numeric rows of an EnergyResult .dat file: energy, data..., smoothed data..."""
    rows = [line.split() for line in open(os.path.join(REF, rel)) if line.strip() and not line.startswith("#")]
    return np.array(rows, dtype=float)


def main():
    I = "reference/integrate_files/"
    system("reference/systems/Haldane", "sys_haldane.npz", ["Ham"], legacy=True,
           extra={"force_internal_terms_only": True})
    system("reference/systems/Chiral_left", "sys_chiral_left.npz", ["Ham"], legacy=True,
           extra={"force_internal_terms_only": True})
    system("data/random", "sys_random.npz", ["Ham", "AA", "BB", "CC", "SS", "SA", "SHA", "SR", "SH", "SHR"],
           legacy=True)
    system("reference/systems/GaAs_W90", "sys_gaas_w90.npz", ["Ham", "AA"], legacy=True)
    system("reference/systems/Fe_W90", "sys_fe_w90.npz", ["Ham", "AA", "BB", "CC", "SS", "SA", "SHA", "SR", "SH", "SHR"],
           legacy=True)
    system("reference/systems/system_Fe_sym_W90_OSD", "sys_fe_sym_osd.npz",
           ["Ham", "AA", "BB", "CC", "SS", "SA", "SHA", "SH", "SR", "SHR"], legacy=False)
    # Fe_sym (Wannier90, projections) after set_spin_interlaced (tests/common_systems.py:276-279):
    # Ham and AA of Fe_W90_proj and the SS that set_spin_interlaced wrote
    system("reference/systems/Fe_W90_proj_set_spin", "sys_fe_w90_proj_set_spin.npz", ["Ham", "AA", "SS"], legacy=True)
    # tellurium from Quantum ESPRESSO (24 WF, 65 R, legacy [m, n, R]); the tests set its point group
    # from the generators C3z, C2x, TimeReversal (tests/common_systems.py:29,1022-1026)
    system("data/Te_qe/system", "sys_te_qe.npz", ["Ham", "AA", "BB", "CC", "SS"], legacy=True)
    results({"haldane_ahc": I + "Haldane-ahc_iter-0000.npz",
             "chiral_ahc": I + "Chiral-ahc_iter-0000.npz",
             "chiral_berry_dipole": I + "Chiral-berry_dipole_iter-0000.npz",
             "chiral_berry_dipole_fsurf": I + "Chiral-berry_dipole_fsurf_iter-0000.npz",
             "random_ahc": I + "random-ahc_iter-0000.npz",
             "random_Morb": I + "random-Morb_iter-0000.npz",
             "fe_ident_fder0": "reference/calculators/calculator-Fe-ident-fder=0.npz",
             "fe_ident_fder1": "reference/calculators/calculator-Fe-ident-fder=1.npz",
             "fe_ident_fder2": "reference/calculators/calculator-Fe-ident-fder=2.npz",
             "fe_ident_fder3": "reference/calculators/calculator-Fe-ident-fder=3.npz",
             # Haldane with the C3z point group, irreducible K-points, symmetrised results and
             # one refinement iteration (adpt_mesh 2, resp. 3): tests/test_run.py:691-745
             "haldane_sym_ahc_iter1": I + "Haldane-ahc_iter-0001.npz",
             "haldane_sym_dos_iter1": I + "Haldane-dos_iter-0001.npz",
             "haldane_adpt3_ahc_iter0": I + "Haldane-ahc-PythTB-adpt3-sym_iter-0000.npz",
             "haldane_adpt3_ahc_iter1": I + "Haldane-ahc-PythTB-adpt3-sym_iter-0001.npz",
             "haldane_adpt3_dos_iter1": I + "Haldane-dos-PythTB-adpt3-sym_iter-0001.npz",
             # Kubo optical conductivity, Chiral model (tests/test_run.py:183-199,799-832)
             "chiral_opt_conductivity": I + "Chiral-opt_conductivity_iter-0000.npz",
             "chiral_jdos": I + "Chiral-jdos_iter-0000.npz",
             # Kubo optical and spin Hall (Ryoo) conductivity on the random model
             # (tests/test_run.py:653-667)
             "random_opt_conductivity": I + "random-opt_conductivity_iter-0000.npz",
             "random_opt_SHCryoo": I + "random-opt_SHCryoo_iter-0000.npz",
             "random_opt_SHCqiao": I + "random-opt_SHCqiao_iter-0000.npz",
             # the same on Fe (Wannier90), grid 4^3 / 2^3 (tests/test_run.py:271-311)
             "fe_opt_conductivity": I + "Fe_W90-opt_conductivity_iter-0000.npz",
             # shift and injection current, GaAs (Wannier90), grid 6^3 / 3^3 (tests/test_run.py:584-611)
             "gaas_shift_current": I + "dynamic_GaAs_W90-shift_current_iter-0000.npz",
             "gaas_injection_current": I + "dynamic_GaAs_W90-injection_current_iter-0000.npz",
             "fe_opt_SHCryoo": I + "Fe_W90-opt_SHCryoo_iter-0000.npz",
             "fe_opt_SHCqiao": I + "Fe_W90-opt_SHCqiao_iter-0000.npz",
             # Fe_sym (Wannier90) with set_spin_interlaced, grid 4^3 / 2^3, no symmetry: Spin,
             # CumDOS and SHC 'simple' with and without external terms (tests/test_run.py:462-484)
             "fe_set_spin_spin": I + "Fe_set_spin-spin_iter-0000.npz",
             "fe_set_spin_cumdos": I + "Fe_set_spin-cumdos_iter-0000.npz",
             "fe_set_spin_SHCsimple": I + "Fe_set_spin-opt_SHCsimple_iter-0000.npz",
             "fe_set_spin_SHCsimple_internal": I + "Fe_set_spin-opt_SHCsimple_internal_iter-0000.npz",
             # tetrahedron method: Chiral model with C3z (tests/test_run.py:227-238,835-845)
             # and the Haldane C3z refinement with adpt_mesh 3 (tests/test_run.py:730-735)
             "chiral_tetra_ahc": I + "Chiral_tetra-ahc_iter-0000.npz",
             "chiral_tetra_berry_dipole": I + "Chiral_tetra-berry_dipole_iter-0000.npz",
             "chiral_tetra_berry_dipole_fsurf": I + "Chiral_tetra-berry_dipole_fsurf_iter-0000.npz",
             "chiral_tetra_dos": I + "Chiral_tetra-dos_iter-0000.npz",
             "chiral_tetra_cumdos": I + "Chiral_tetra-cumdos_iter-0000.npz",
             # the tetrahedral grid: GridTetra(length=8, NKFFT=[5, 5, 2]) (tests/test_run.py:958-968)
             "chiral_tetragrid_ahc": I + "Chiral_tetragrid-ahc_iter-0000.npz",
             "chiral_tetragrid_berry_dipole": I + "Chiral_tetragrid-berry_dipole_iter-0000.npz",
             "chiral_tetragrid_berry_dipole_fsurf": I + "Chiral_tetragrid-berry_dipole_fsurf_iter-0000.npz",
             "chiral_tetragrid_dos": I + "Chiral_tetragrid-dos_iter-0000.npz",
             "chiral_tetragrid_cumdos": I + "Chiral_tetragrid-cumdos_iter-0000.npz",
             "haldane_adpt3_tetra_ahc_iter0": I + "Haldane-ahc-PythTB-adpt3-tetra-sym_iter-0000.npz",
             "haldane_adpt3_tetra_ahc_iter1": I + "Haldane-ahc-PythTB-adpt3-tetra-sym_iter-0001.npz",
             "haldane_adpt3_tetra_dos_iter0": I + "Haldane-dos-PythTB-adpt3-tetra-sym_iter-0000.npz",
             "haldane_adpt3_tetra_dos_iter1": I + "Haldane-dos-PythTB-adpt3-tetra-sym_iter-0001.npz",
             # Ohmic conductivity, Fermi sea (InvMass) and surface (VelVel): Chiral (10x10x4 / 5x5x2,
             # also with the tetrahedron method and C3z), random (6^3 / 3^3), Fe (4^3 / 2^3)
             # (tests/test_run.py:119-128,186-238,271-303,653-669)
             "chiral_ohmic": I + "Chiral-conductivity_ohmic_iter-0000.npz",
             "chiral_ohmic_fsurf": I + "Chiral-conductivity_ohmic_fsurf_iter-0000.npz",
             "chiral_tetra_ohmic": I + "Chiral_tetra-conductivity_ohmic_iter-0000.npz",
             "chiral_tetra_ohmic_fsurf": I + "Chiral_tetra-conductivity_ohmic_fsurf_iter-0000.npz",
             "random_ohmic": I + "random-conductivity_ohmic_iter-0000.npz",
             "random_ohmic_fsurf": I + "random-conductivity_ohmic_fsurf_iter-0000.npz",
             "fe_ohmic": I + "Fe_W90-conductivity_ohmic_iter-0000.npz",
             "fe_ohmic_fsurf": I + "Fe_W90-conductivity_ohmic_fsurf_iter-0000.npz",
             # classic Hall conductivity, Fermi surface and sea, Chiral (also tetrahedra)
             "chiral_hall_classic": I + "Chiral-Hall_classic_iter-0000.npz",
             "chiral_hall_classic_fsurf": I + "Chiral-Hall_classic_fsurf_iter-0000.npz",
             "chiral_tetra_hall_classic": I + "Chiral_tetra-Hall_classic_iter-0000.npz",
             "chiral_tetra_hall_classic_fsurf": I + "Chiral_tetra-Hall_classic_fsurf_iter-0000.npz",
             # nonlinear Drude, Fermi sea (Der3E), Chiral (also tetrahedra)
             "chiral_der3e": I + "Chiral-Der3E_iter-0000.npz",
             "chiral_tetra_der3e": I + "Chiral_tetra-Der3E_iter-0000.npz",
             }, "results_run.npz")
    # FermiDiracSmoother(Efermi, T_Kelvin=1200, maxdE=8) on the Chiral AHC and Fermi-sea BCD
    # (tests/test_run.py:181-193): the .dat files hold the data and the smoothed data
    np.savez_compressed(os.path.join(HERE, "results_smooth.npz"),
                        chiral_ahc_dat=dat_columns(I + "Chiral-ahc_iter-0000.dat"),
                        chiral_berry_dipole_dat=dat_columns(I + "Chiral-berry_dipole_iter-0000.dat"))
    # Te (QE): tetra=True, irreducible K-points and symmetrisation, NK [3, 3, 4], NKFFT [1, 1, 4],
    # Efermi = linspace(4, 8, 11) (tests/test_run.py:154-170,1102-1119; common_systems.py:35)
    results({"te_" + k: I + f"Te_QE-{k}_iter-0000.npz"
             for k in ["GME_orb_FermiSea", "GME_orb_FermiSurf", "BerryDipole_FermiSea", "berry_dipole",
                       "dos", "cumdos", "NLDrude_FermiSurf", "NLDrude_Fermider2", "AHC_Zeeman_spin"]},
            "results_te_qe.npz")
    # k-resolved tabulation, Fe (Wannier90), grid 4^3 / 2^3, bands 5 and 6
    # (tests/test_run.py:271-288,313-345)
    with np.load(os.path.join(REF, I + "Fe_W90-tabulate.npz"), allow_pickle=False) as f:
        np.savez_compressed(os.path.join(HERE, "results_tabulate.npz"),
                            **{k: f[k] for k in ["Energy", "V", "spin", "berry", "Der_berry", "morb", "Der_morb",
                                                  "spin_berry"]})
    # per-formula gauge-invariant traces (tests/reference/formula/*.npz)
    forms = {}
    for name in ["Dcov", "DerDcov", "DEinv_ln", "InvMass", "Omega", "DerOmega", "Morb_H", "morb",
                 "DerMorb_upper", "Dermorb_lower", "VelOmega", "Hamiltonian", "Velocity", "Spin", "VelSpin",
                 "DerWln", "Der3E", "SpinOmega", "VelVel", "VelVelVel", "VelMassVel", "OmegaS", "OmegaOmega"]:
        with np.load(os.path.join(REF, "reference/formula", name + ".npz"), allow_pickle=False) as f:
            for k in f.files:
                forms[name + "/" + k] = f[k]
    np.savez_compressed(os.path.join(HERE, "results_formula.npz"), **forms)
    # k-path refined by a factor 4 (tests/test_path.py:76-95)
    with np.load(os.path.join(REF, "reference/path_refined.npz"), allow_pickle=False) as f:
        np.savez_compressed(os.path.join(HERE, "results_path.npz"), **{"refined/" + k: f[k] for k in f.files})
    with np.load(os.path.join(REF, "reference/evaluate_k.npz"), allow_pickle=False) as f:
        np.savez_compressed(os.path.join(HERE, "results_evaluate_k.npz"),
                            **{k: f[k] for k in ["kpoint", "energy", "berry_curvature", "ahc", "Morb"]})


if __name__ == "__main__":
    main()
