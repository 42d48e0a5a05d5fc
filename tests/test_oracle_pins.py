"""Pins of the FP64 oracle (oracle/wb_oracle.py) against the reference's own fixtures.

Every oracle function on the path is checked by at least one of:

* P1 `Haldane-ahc` (AHC end to end, internal terms, config-1 Hamiltonian);
* P2 `random-ahc`, `random-Morb` (external terms A, rotA, B, C; Morb non-additive cuts);
* P4 `Chiral-{ahc,berry_dipole,berry_dipole_fsurf}` (DerOmega sea, VelOmega surface, fder=1);
* P5 per-formula gauge-invariant traces on Fe (Dcov, DerDcov, DEinv, InvMass, Omega,
  DerOmega, Morb_H, morb, DerMorb, Dermorb, VelOmega, Hamiltonian, Velocity);
* P6 `calculator-Fe-ident-fder=0..3` (band grouping, Fermi scan, all fder stencils);
* P7 `evaluate_k` (eigenvalues, per-band Berry curvature, AHC, Morb at one k);
* P8 dynamic calculators: `Chiral-opt_conductivity`, `Chiral-jdos`, `random-opt_conductivity`,
  `random-opt_SHC{ryoo,qiao}`, `Fe_W90-opt_{conductivity,SHCryoo,SHCqiao}`,
  `dynamic_GaAs_W90-{shift,injection}_current` (at that test's absolute tolerances);
* P9 k-resolved tabulation: `Fe_W90-tabulate` (Energy, V, spin, berry, Der_berry, morb, Der_morb);
* the tetrahedron method is pinned in tests/test_oracle_tetra.py;
* closed forms and invariants: brute-force DFT == FFT (r2k_grid_slow), U^+HU = diag(E),
  sum_n Omega_n = 0 over all bands (internal terms), Tr dOmega = 0 (internal).

The tolerance of each pin is the reference test's own (tests/test_run.py:56,
tests/test_formula.py:207-213, tests/test_calc.py:50-58).
"""
import os

import numpy as np
import pytest

from oracle import wb_oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def gsys(name):
    return O.load_golden_system(os.path.join(GOLDEN, f"sys_{name}.npz"))


@pytest.fixture(scope="module")
def results_run():
    with np.load(os.path.join(GOLDEN, "results_run.npz"), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_haldane_ahc(results_run):
    s = gsys("haldane")
    Ef = np.linspace(-3, 3, 11)
    r = O.run(s, {"ahc": O.AHC(Ef)}, NK=[10, 10, 1], NKFFT=[5, 5, 1])
    assert rel_err(r["ahc"].data, results_run["haldane_ahc"]) < 1e-8


E2_H_ANGSTROM = 1.602176634e-19 ** 2 / 6.62607015e-34 / 1e-10   # e^2/(h * 1 A) in S/m


def test_haldane_chern_plateau():
    """Closed form: in the gap the Haldane layer (Chern number 1, interlayer 1 A)
    carries sigma_xy = e^2/(h * 1 A); the grid sum converges exponentially
    (10x10: 1.6e-3 off, the reference's own 3.86796e5; 40x40: ~1e-13)."""
    s = gsys("haldane")
    Ef = np.linspace(-3, 3, 1001)
    r = O.run(s, {"ahc": O.AHC(Ef)}, NK=[40, 40, 1], NKFFT=[20, 20, 1])
    gap = np.abs(Ef) < 0.5            # the gap is (-0.579, 0.579) eV
    z = r["ahc"].data[gap, 2]
    assert np.abs(z / E2_H_ANGSTROM - 1).max() < 1e-10
    # a 2-D layer has no in-plane Berry curvature
    assert np.abs(r["ahc"].data[:, :2]).max() < 1e-10 * E2_H_ANGSTROM


def test_run_adaptive_invariants():
    """Refinement bookkeeping (run_grid.py:352-389, Kpoint.py:145-175).

    Every K-block is a full shifted NKFFT grid, so any weights that sum to one
    give a valid estimate: at NKFFT 40x40 each block carries the Chern plateau
    to ~1e-13, and a lost or doubled block (factor sum != 1, a missing
    correction of a divided block) moves it.  The running sum must equal the
    direct sum over the final factors, and the divided blocks must be the
    ones with the largest factor-weighted metrics.
    """
    s = gsys("haldane")
    Ef = np.linspace(-3, 3, 61)
    calcs = {"ahc": O.AHC(Ef), "dos": O.CumDOS(Ef)}
    tot, blocks = O.run_adaptive(s, calcs, NK=[80, 80, 1], NKFFT=[40, 40, 1], adpt_num_iter=2)
    fac = np.array([b[2] for b in blocks])
    ndiv = int((fac == 0).sum())
    assert 2 <= ndiv <= 2 * 6                     # 1..6 metrics select per iteration
    assert len(blocks) == 4 + 4 * ndiv            # 2x2 sub-blocks (z is not periodic)
    assert abs(fac.sum() - 1) < 1e-14
    gap = np.abs(Ef) < 0.5
    assert np.abs(tot["ahc"].data[gap, 2] / E2_H_ANGSTROM - 1).max() < 1e-10
    direct = None
    for K, dK, f in blocks:
        if f == 0:
            continue
        d = O.OracleDataK(s, dK=K / np.array([40, 40, 1]), NKFFT=np.array([40, 40, 1]))
        r = calcs["ahc"](d) * f
        direct = r if direct is None else direct + r
    assert np.abs(direct.data - tot["ahc"].data).max() < 1e-12 * np.abs(tot["ahc"].data).max()


def test_chiral_ahc_bcd(results_run):
    s = gsys("chiral_left")
    Ef = np.linspace(-5, 8, 27)
    calcs = {"ahc": O.AHC(Ef),
             "bd": O.BerryDipole_FermiSea(Ef, use_factor=False),
             "bds": O.BerryDipole_FermiSurf(Ef, use_factor=False)}
    r = O.run(s, calcs, NK=[10, 10, 4], NKFFT=[5, 5, 2])
    assert rel_err(r["ahc"].data, results_run["chiral_ahc"]) < 1e-8
    assert rel_err(r["bd"].data, results_run["chiral_berry_dipole"]) < 1e-8
    assert rel_err(r["bds"].data, results_run["chiral_berry_dipole_fsurf"]) < 1e-8


def test_random_ahc_morb(results_run):
    s = gsys("random")
    Ef = np.linspace(-2, 2, 5)
    r = O.run(s, {"ahc": O.AHC(Ef), "Morb": O.Morb(Ef)}, NK=[6, 6, 6], NKFFT=[3, 3, 3])
    assert rel_err(r["ahc"].data, results_run["random_ahc"]) < 1e-8
    assert rel_err(r["Morb"].data, results_run["random_Morb"]) < 1e-8


def test_ohmic_chiral_random_fe(results_run):
    """Ohmic conductivity, Fermi sea (InvMass, fder 0) and surface (VelVel, fder 1), against the
    reference's Chiral (10x10x4 / 5x5x2), random (6^3 / 3^3) and Fe (4^3 / 2^3) fixtures
    (tests/test_run.py:119-128,186-190,271-303,653-669)."""
    for name, Ef, NK, NKFFT in (("chiral_left", np.linspace(-5, 8, 27), [10, 10, 4], [5, 5, 2]),
                                ("random", np.linspace(-2, 2, 5), [6, 6, 6], [3, 3, 3]),
                                ("fe_w90", np.linspace(17, 18, 11), [4, 4, 4], [2, 2, 2])):
        r = O.run(gsys(name), {"sea": O.Ohmic_FermiSea(Ef), "surf": O.Ohmic_FermiSurf(Ef)}, NK=NK, NKFFT=NKFFT)
        key = name.split("_")[0]
        assert rel_err(r["sea"].data, results_run[key + "_ohmic"]) < 1e-8, name
        assert rel_err(r["surf"].data, results_run[key + "_ohmic_fsurf"]) < 1e-8, name


def test_velocity_products_finite_difference():
    """VelVelVel and MassVel (formula/covariant.py:823-833) traced over an isolated band are
    v_a v_b v_c and (d_b v_a) v_c of that band, v = dE/dk: central differences of the energies
    on Fe (W90) at a general k, every band whose neighbours are at least 0.05 eV away.  A
    factor in the wrong slot of the product or a wrong InvMass index fails this."""
    o = gsys("fe_sym_osd")
    k0 = np.array([0.1, 0.2, -0.3])
    d0 = O.OracleDataK(o, k_list=k0[None])
    E0 = d0.E_K[0]
    gap = np.minimum(np.diff(E0, prepend=-np.inf), np.diff(E0, append=np.inf))
    bands = np.where(gap > 0.05)[0]
    assert len(bands) >= 6
    h = 1e-4
    binv = np.linalg.inv(o.recip_lattice)
    step = [h * binv[a] for a in range(3)]     # Cartesian step h along axis a, reduced coordinates

    def energies(shifts):
        return O.OracleDataK(o, k_list=np.array([k0 + s for s in shifts])).E_K

    v = np.stack([(lambda e: (e[0] - e[1]) / (2 * h))(energies([step[a], -step[a]])) for a in range(3)], axis=-1)
    m = np.zeros((o.num_wann, 3, 3))
    for a in range(3):
        for b in range(3):
            e = energies([step[a] + step[b], step[a] - step[b], -step[a] + step[b], -step[a] - step[b]])
            m[:, a, b] = (e[0] - e[1] - e[2] + e[3]) / (4 * h * h)
    vvv, mv = O.VelVelVel(d0), O.MassVel(d0)
    for n in bands:
        out = [x for x in range(o.num_wann) if x != n]
        t3 = vvv.trace(0, [n], out)
        tm = mv.trace(0, [n], out)
        ref3 = np.einsum("a,b,c->abc", v[n], v[n], v[n])
        refm = np.einsum("ab,c->abc", m[n], v[n])
        assert np.abs(t3 - ref3).max() < 1e-5 * max(1, np.abs(ref3).max()), n
        assert np.abs(tm - refm).max() < 1e-4 * max(1, np.abs(refm).max()), n


@pytest.mark.parametrize("kind", ["qiao", "ryoo"])
def test_shc_static_equals_dynamic_at_zero_frequency(kind):
    """The dc spin Hall conductivity (static SHC: Fermi-sea integral of the spin Berry
    curvature, static.py:647-660) equals the Kubo SHC at omega = 0 with a vanishing Gaussian
    width (1e-10 eV), on Fe (Wannier90), grid 4^3 / 2^3, Efermi 17 and 18 eV, at the reference's
    own tolerance (tests/test_run.py:1120-1151).  The Kubo SHC is pinned by the Fe fixtures."""
    Ef = np.array([17.0, 18.0])
    r = O.run(gsys("fe_w90"), {"dyn": O.SHC(Ef, np.arange(0.0, 7.1, 1.0), SHC_type=kind, smr_fixed_width=1e-10,
                                           smr_type="Gaussian"),
                               "st": O.SHC_static(Ef, kwargs_formula={"spin_current_type": kind})},
              NK=[4, 4, 4], NKFFT=[2, 2, 2])
    st, dyn = r["st"].data, r["dyn"].data[:, 0].real
    assert np.abs(st).max() > 0
    assert np.abs(st - dyn).max() <= max(np.average(np.abs(st)) / 1e10, 1e-8)


def test_chiral_hall_classic(results_run):
    """Classic Hall conductivity on Chiral (10x10x4 / 5x5x2): Fermi surface (VelMassVel) and
    sea (MassMass), the rank-4 products contracted as static.py:407-446
    (tests/test_run.py:186-200)."""
    Ef = np.linspace(-5, 8, 27)
    r = O.run(gsys("chiral_left"), {"surf": O.Hall_classic_FermiSurf(Ef), "sea": O.Hall_classic_FermiSea(Ef)},
              NK=[10, 10, 4], NKFFT=[5, 5, 2])
    assert rel_err(r["surf"].data, results_run["chiral_hall_classic_fsurf"]) < 1e-8
    assert rel_err(r["sea"].data, results_run["chiral_hall_classic"]) < 1e-8


def test_chiral_nldrude_fermi_sea(results_run):
    """Nonlinear Drude conductivity, Fermi sea (Der3E, static.py:519-529), on Chiral
    (10x10x4 / 5x5x2) against the reference's Chiral-Der3E fixture."""
    Ef = np.linspace(-5, 8, 27)
    r = O.run(gsys("chiral_left"), {"d3": O.NLDrude_FermiSea(Ef)}, NK=[10, 10, 4], NKFFT=[5, 5, 2])
    assert rel_err(r["d3"].data, results_run["chiral_der3e"]) < 1e-8


@pytest.mark.parametrize("fder", [0, 1, 2, 3])
def test_fe_scan_stencils(results_run, fder):
    s = gsys("fe_w90")
    d = O.OracleDataK(s, dK=[0.1, 0.2, 0.3], NKFFT=[3, 3, 3])
    c = O.StaticCalculator(np.linspace(17, 18, 11), Formula=O.Identity, fder=fder)
    ref = results_run[f"fe_ident_fder{fder}"]
    assert np.abs(c(d).data - ref).max() < max(1e-8 * np.abs(ref).max(), 1e-11)


def test_evaluate_k():
    s = gsys("fe_w90")
    with np.load(os.path.join(GOLDEN, "results_evaluate_k.npz"), allow_pickle=False) as f:
        ev = {k: f[k] for k in f.files}
    d = O.OracleDataK(s, dK=ev["kpoint"], NKFFT=[1, 1, 1])
    assert np.abs(d.E_K[0][4:6] - ev["energy"]).max() < 1e-8 * 17
    om = O.Omega(d)
    nb = s.num_wann
    berry = np.array([om.trace(0, [n], [m for m in range(nb) if m != n]) for n in (4, 5)])
    assert rel_err(berry, ev["berry_curvature"]) < 1e-8
    Ef = np.linspace(17, 18, 11)
    assert rel_err(O.AHC(Ef)(d).data, ev["ahc"]) < 1e-8
    assert rel_err(O.Morb(Ef)(d).data, ev["Morb"]) < 1e-8


FORMULAS = {
    "Dcov": O.Dcov, "DerDcov": O.DerDcov, "DEinv_ln": O.DEinv_ln, "InvMass": O.InvMass,
    "Omega": O.Omega, "DerOmega": O.DerOmega, "Morb_H": O.Morb_H, "morb": O.morb,
    "DerMorb_upper": O.DerMorb, "Dermorb_lower": O.Dermorb, "VelOmega": O.VelOmega,
    "Hamiltonian": lambda d: O.Matrix_ln(d.Xbar("Ham", 0)),
    # the reference's Velocity copies only the attributes of the covariant V,
    # so its ln block is the full off-diagonal block (formula/covariant.py:321-328)
    "Velocity": lambda d: O.Matrix_ln(d.Xbar("Ham", 1)),
    "Spin": O.Spin, "VelSpin": O.VelSpin,
    "DerWln": O.DerWln, "Der3E": O.Der3E, "VelVel": O.VelVel, "VelVelVel": O.VelVelVel, "VelMassVel": O.VelMassVel,
    # tests/test_formula.py:175-176: the Ryoo spin current
    "SpinOmega": lambda d: O.SpinOmega(d, spin_current_type="ryoo"),
    "OmegaS": O.OmegaS, "OmegaOmega": O.OmegaOmega,
}
LN_ONLY = {"Dcov", "DerDcov", "DEinv_ln"}
NN_ONLY = {"Omega", "DerOmega", "Morb_H", "morb", "DerMorb_upper", "Dermorb_lower", "VelOmega", "VelSpin", "Der3E",
           "VelVel", "VelVelVel", "VelMassVel", "SpinOmega", "OmegaS", "OmegaOmega"}


@pytest.fixture(scope="module")
def datak_fe_sym():
    # tests/test_formula.py:19-41: dK = 0, NKFFT = [1, 2, 3]
    return O.OracleDataK(gsys("fe_sym_osd"), dK=[0, 0, 0], NKFFT=[1, 2, 3])


@pytest.mark.parametrize("name", sorted(FORMULAS))
def test_formula_traces(datak_fe_sym, name):
    d = datak_fe_sym
    nb = d.num_wann
    groups = d.get_bands_in_range_groups(-10, 30, degen_thresh=1)
    with np.load(os.path.join(GOLDEN, "results_formula.npz"), allow_pickle=False) as f:
        ref = {k.split("/", 1)[1]: f[k] for k in f.files if k.startswith(name + "/")}
    val = {}
    for ik in range(d.nk):
        for key in ["Xnn", "Xll", "XlnXnl", "XnlXln", "XnnXnn", "XllXll"]:
            val[f"{key}_ik={ik}"] = []
        for n in groups[ik]:
            inn = np.arange(n[0], n[1])
            out = np.concatenate((np.arange(0, n[0]), np.arange(n[1], nb)))
            form = FORMULAS[name](d)
            if name not in NN_ONLY:
                Xnl, Xln = form.ln(ik, inn, out), form.ln(ik, out, inn)
                val[f"XnlXln_ik={ik}"].append(np.einsum("nl...,ln...->...", Xnl, Xln))
                val[f"XlnXnl_ik={ik}"].append(np.einsum("ln...,nl...->...", Xln, Xnl))
            if name not in LN_ONLY:
                # the reference calls ll/nn with out=inn (tests/test_formula.py:198-200)
                Xll, Xnn = form.ll(ik, inn, inn), form.nn(ik, inn, inn)
                val[f"Xll_ik={ik}"].append(np.einsum("ll...->...", Xll))
                val[f"Xnn_ik={ik}"].append(np.einsum("nn...->...", Xnn))
                val[f"XllXll_ik={ik}"].append(np.einsum("ll...,ll...->...", Xll, Xll))
                val[f"XnnXnn_ik={ik}"].append(np.einsum("nn...,nn...->...", Xnn, Xnn))
    der = name.startswith("Der") or name in ("VelOmega", "SpinOmega")
    rtol, atol = (1e-5, 2e-6) if der else (1e-6, 1e-10)
    assert len(ref) == 36
    for k, r in ref.items():
        o = np.array(val[k])
        assert o.shape == r.shape, k
        if r.size == 0:
            continue
        mx = np.abs(r).max()
        if mx < atol:
            assert np.abs(o).max() < atol, k
        else:
            assert np.abs(r - o).max() / mx < rtol, k


def test_derspin_finite_difference():
    """DerSpin (the generalised derivative of Xbar('SS'), formula/covariant.py:338-342) traced
    over an isolated band is the k-derivative of that band's spin expectation <n|S^b|n>:
    central differences in each Cartesian direction on Fe (W90, spin matrices), all bands
    whose neighbours are at least 0.05 eV away.  A dropped or sign-flipped D term, or the
    derivative index in the wrong slot, fails this."""
    o = gsys("fe_sym_osd")
    k0 = np.array([0.1, 0.2, -0.3])
    d0 = O.OracleDataK(o, k_list=k0[None])
    E = d0.E_K[0]
    gap = np.minimum(np.diff(E, prepend=-np.inf), np.diff(E, append=np.inf))
    bands = np.where(gap > 0.05)[0]
    assert len(bands) >= 6
    ds = O.DerSpin(d0)
    h = 1e-5
    binv = np.linalg.inv(o.recip_lattice)
    for d in range(3):
        dk = h * binv[d]           # Cartesian step h along axis d, in reduced coordinates
        sp = O.OracleDataK(o, k_list=np.array([k0 + dk, k0 - dk]))
        S = np.einsum("knnb->knb", sp.Xbar("SS")).real
        fd = (S[0] - S[1]) / (2 * h)
        for n in bands:
            out = [m for m in range(o.num_wann) if m != n]
            tr = ds.trace(0, [n], out)     # [b, d]
            assert np.abs(tr[:, d] - fd[n]).max() < 1e-5 * max(1, np.abs(fd[n]).max()), (n, d)


# -- closed forms and invariants ------------------------------------------

def test_fft_equals_bruteforce_dft():
    s = gsys("random")
    d = O.OracleDataK(s, dK=[0.13, -0.27, 0.41], NKFFT=[3, 4, 5])
    X = d.get_R_mat("AA")
    fast = O.r2k_grid(X, s.iRvec, d.NKFFT, hermitian=True)
    # direct sum over R at k = j/N (the dK phase is already inside X)
    slow = O.r2k_grid_slow(X, s.iRvec, d.NKFFT, hermitian=True)
    assert np.abs(fast - slow).max() < 1e-12 * np.abs(slow).max()


def test_eigh_invariants():
    s = gsys("random")
    d = O.OracleDataK(s, dK=[0.1, 0.2, 0.3], NKFFT=[2, 2, 2])
    U, E, H = d.UU_K, d.E_K, d.HH_K
    UhHU = np.einsum("kba,kbc,kcd->kad", U.conj(), H, U)
    assert np.abs(UhHU - np.eye(s.num_wann)[None] * E[:, :, None]).max() < 1e-12
    assert np.all(np.diff(E, axis=1) >= 0)


def test_omega_sum_rule_and_bcd_trace():
    # internal terms: sum_n Omega_n = 0 over all bands; Tr_c d_c Omega_c = 0 (tests/test_calc.py:110-120)
    s = gsys("chiral_left")
    d = O.OracleDataK(s, dK=[0.05, 0.1, 0.2], NKFFT=[3, 3, 2])
    om = O.Omega(d)
    dom = O.DerOmega(d)
    nb = s.num_wann
    for ik in range(d.nk):
        tot = sum(om.trace(ik, [n], [m for m in range(nb) if m != n]) for n in range(nb))
        assert np.abs(tot).max() < 1e-9
        for n in range(nb):
            t = dom.trace(ik, [n], [m for m in range(nb) if m != n])
            assert abs(np.trace(t)) < 1e-9 * max(1, np.abs(t).max())


def test_chiral_optical_conductivity(results_run):
    """Kubo optical conductivity (dynamic.py:26-196) on the Chiral model, Gaussian
    smearing 0.2 eV, 27 Fermi levels x 11 frequencies (tests/test_run.py:183-199)."""
    s = gsys("chiral_left")
    calc = O.OpticalConductivity(np.linspace(-5, 8, 27), np.linspace(0, 1, 11), smr_fixed_width=0.2,
                                 smr_type="Gaussian")
    r = O.run(s, {"oc": calc}, NK=[10, 10, 4], NKFFT=[5, 5, 2])
    assert rel_err(r["oc"].data, results_run["chiral_opt_conductivity"]) < 1e-8


def test_chiral_jdos(results_run):
    """Joint density of states (dynamic.py:145-167) on the Chiral model, Gaussian smearing
    0.2 eV, 27 Fermi levels x 11 frequencies (tests/test_run.py:183-202)."""
    s = gsys("chiral_left")
    calc = O.JDOS(np.linspace(-5, 8, 27), np.linspace(0, 1, 11), smr_fixed_width=0.2, smr_type="Gaussian")
    r = O.run(s, {"jdos": calc}, NK=[10, 10, 4], NKFFT=[5, 5, 2])
    assert rel_err(r["jdos"].data, results_run["chiral_jdos"]) < 1e-8


def test_random_optical_and_spin_hall(results_run):
    """Kubo optical conductivity (external terms) and spin Hall conductivity with the Ryoo
    and Qiao spin currents on the random model, Gaussian 0.2 eV, Efermi 17 and 18 eV, omega 0..7 eV
    (tests/test_run.py:653-667, default grid 6^3 / 3^3)."""
    s = gsys("random")
    kw = dict(smr_fixed_width=0.2, smr_type="Gaussian")
    Ef, om = np.array([17.0, 18.0]), np.arange(0.0, 7.1, 1.0)
    r = O.run(s, {"oc": O.OpticalConductivity(Ef, om, **kw), "shc": O.SHC(Ef, om, SHC_type="ryoo", **kw)},
              NK=[6, 6, 6], NKFFT=[3, 3, 3])
    assert rel_err(r["oc"].data, results_run["random_opt_conductivity"]) < 1e-8
    assert rel_err(r["shc"].data, results_run["random_opt_SHCryoo"]) < 1e-8
    rq = O.run(s, {"shc": O.SHC(Ef, om, SHC_type="qiao", **kw)}, NK=[6, 6, 6], NKFFT=[3, 3, 3])
    assert rel_err(rq["shc"].data, results_run["random_opt_SHCqiao"]) < 1e-8


def _tabulate_grid(s, quantities, ibands, NK, NKFFT):
    """Oracle tabulation over all K-blocks, assembled on the full grid (k index
    (i0*N1 + i1)*N2 + i2, tabresult.py:134-170)."""
    NKdiv, NKFFT, blocks = O.grid_blocks(NK, NKFFT)
    NK = NKdiv * NKFFT
    out = {}
    for dK, _ in blocks:
        d = O.OracleDataK(s, dK=dK, NKFFT=NKFFT)
        ki = np.rint(d.kpoints_all * NK[None, :]).astype(int) % NK[None, :]
        idx = (ki[:, 0] * NK[1] + ki[:, 1]) * NK[2] + ki[:, 2]
        for key, (F, kw) in quantities.items():
            v = O.tabulate(d, F, ibands, **kw)
            if key not in out:
                out[key] = np.zeros((int(np.prod(NK)),) + v.shape[1:])
            out[key][idx] = v
    return {k: v.reshape(tuple(NK) + v.shape[1:]) for k, v in out.items()}


TAB_QUANTITIES = {"Energy": (O.Eavln, {}),
                  # Velocity = covariant('Ham', gender=1), Spin = covariant('SS') (covariant.py:321-336)
                  "V": (lambda d: d.covariant("Ham", gender=1), {"degen_thresh": 5e-2}),
                  "spin": (lambda d: O.Matrix_ln(d.Xbar("SS")), {"degen_thresh": 5e-2}),
                  "berry": (O.Omega, {"degen_thresh": 5e-2}),
                  "Der_berry": (O.DerOmega, {"degen_thresh": 5e-2}),
                  "morb": (O.morb, {"degen_thresh": 5e-2, "constant_factor": O.FACTOR_MORB_EVA2_TO_MUB}),
                  "Der_morb": (O.Dermorb, {"degen_thresh": 5e-2, "constant_factor": O.FACTOR_MORB_EVA2_TO_MUB}),
                  # SpinOmega with the Ryoo spin current (covariant.py:759-789)
                  "spin_berry": (O.SpinOmega, {"degen_thresh": 5e-2})}


def test_fe_tabulate():
    """k-resolved tabulation (tabulate.py:20-48, tabresult.py:134-170) on Fe (Wannier90),
    grid 4^3 / 2^3, bands 5 and 6 (tests/test_run.py:271-288); tolerances of the reference
    test (2e-8, Der_berry 5e-8, relative to the largest value)."""
    with np.load(os.path.join(GOLDEN, "results_tabulate.npz"), allow_pickle=False) as f:
        ref = {k: f[k] for k in f.files}
    r = _tabulate_grid(gsys("fe_w90"), TAB_QUANTITIES, [5, 6], [4, 4, 4], [2, 2, 2])
    for k in TAB_QUANTITIES:
        assert rel_err(r[k], ref[k]) < (5e-8 if k == "Der_berry" else 2e-8), k


def test_fe_optical_and_spin_hall(results_run):
    """Kubo optical and spin Hall (Ryoo, Qiao) conductivity on Fe (Wannier90), grid 4^3 / 2^3,
    Gaussian 0.2 eV, Efermi 17 and 18 eV, omega 0..7 eV (tests/test_run.py:271-311)."""
    s = gsys("fe_w90")
    kw = dict(smr_fixed_width=0.2, smr_type="Gaussian")
    Ef, om = np.array([17.0, 18.0]), np.arange(0.0, 7.1, 1.0)
    r = O.run(s, {"oc": O.OpticalConductivity(Ef, om, **kw), "ry": O.SHC(Ef, om, SHC_type="ryoo", **kw),
                  "qi": O.SHC(Ef, om, SHC_type="qiao", **kw)}, NK=[4, 4, 4], NKFFT=[2, 2, 2])
    for k, ref in (("oc", "fe_opt_conductivity"), ("ry", "fe_opt_SHCryoo"), ("qi", "fe_opt_SHCqiao")):
        assert rel_err(r[k].data, results_run[ref]) < 1e-8, k


def test_gaas_shift_injection_current(results_run):
    """Shift and injection current (dynamic.py:244-361) on GaAs (Wannier90, nw = 16), grid
    6^3 / 3^3, Gaussian 0.2 eV, kBT 0.01 eV, Efermi 7..9 eV, omega 1..5 eV
    (tests/test_run.py:584-611), at the reference test's own absolute tolerances
    (shift 1e-9, injection 1e-7 against maxima of 5.7e-5 and 1.0e-5): the fixture was made
    from the Wannier90 files, the stored system reproduces it to about 1e-10 / 4e-8."""
    s = gsys("gaas_w90")
    kw = dict(smr_fixed_width=0.2, smr_type="Gaussian", kBT=0.01)
    Ef, om = np.linspace(7, 9, 11), np.arange(1.0, 5.1, 0.5)
    r = O.run(s, {"sc": O.ShiftCurrent(Ef, om, sc_eta=0.1, **kw), "ic": O.InjectionCurrent(Ef, om, **kw)},
              NK=[6, 6, 6], NKFFT=[3, 3, 3])
    assert np.abs(r["sc"].data - results_run["gaas_shift_current"]).max() < 1e-9
    assert np.abs(r["ic"].data - results_run["gaas_injection_current"]).max() < 1e-7


def test_set_spin_interlaced():
    """System_R.set_spin_interlaced (system_R.py:500-545) on the stored Fe_W90_proj system
    reproduces the SS that the reference's set_spin_interlaced wrote there."""
    import warnings
    import wannierberri_amd as wb
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_fe_w90_proj_set_spin.npz"))
    ref = s.get_R_mat("SS").copy()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")   # the pairs' centres differ, as the reference warns too
        s.set_spin_interlaced()
    assert np.abs(s.get_R_mat("SS") - ref).max() < 1e-14


def test_fe_set_spin(results_run):
    """Spin, CumDOS and SHC 'simple' with and without external terms on Fe (Wannier90,
    projections) with interlaced spins, grid 4^3 / 2^3, no symmetry (tests/test_run.py:462-484):
    the 'simple' spin current (formula/covariant.py:721-727) with the Velocity's external term
    (321-328), at the reference test's 1e-8."""
    o = gsys("fe_w90_proj_set_spin")
    Ef = np.linspace(17, 18, 11)
    kw = dict(smr_fixed_width=0.2, smr_type="Gaussian")
    Ef2, om = np.array([Ef[0], Ef[-1]]), np.arange(0.0, 7.1, 1.0)
    r = O.run(o, {"spin": O.SpinCalc(Ef), "cumdos": O.CumDOS(Ef),
                  "shc": O.SHC(Ef2, om, SHC_type="simple", kwargs_formula={"external_terms": True}, **kw),
                  "shc_int": O.SHC(Ef2, om, SHC_type="simple", kwargs_formula={"external_terms": False}, **kw)},
              NK=[4, 4, 4], NKFFT=[2, 2, 2])
    for k, ref in (("spin", "fe_set_spin_spin"), ("cumdos", "fe_set_spin_cumdos"), ("shc", "fe_set_spin_SHCsimple"),
                   ("shc_int", "fe_set_spin_SHCsimple_internal")):
        assert rel_err(r[k].data, results_run[ref]) < 1e-8, k


def test_select_bands_partition():
    """select_bands (static.py:133,136; utility.py:398-403; tetrahedron.py:147-149) on the
    random model with pairs of bands as groups (degen_Kramers): all bands selected gives the
    unselected result; the even and the odd bands, selected apart, each give half of it (each
    group of two adds half its value), and sum to it; listing the even bands twice counts twice;
    the Fermi sea refuses a selection."""
    s = gsys("random")
    nw = s.matrices["Ham"].shape[-1]
    E = np.linspace(-3, 3, 31)
    kw = dict(degen_Kramers=True)   # groups of two bands (even borders only)

    def run(sel):
        c = {"dos": O.DOS(E, select_bands=sel, **kw), "bcd": O.BerryDipole_FermiSurf(E, select_bands=sel, **kw)}
        return O.run(s, c, NK=[4, 4, 4], NKFFT=[2, 2, 2])
    full, every = run(None), run(list(range(nw)))
    even, odd = run(list(range(0, nw, 2))), run(list(range(1, nw, 2)))
    twice = run(list(range(0, nw, 2)) * 2)
    for k in ("dos", "bcd"):
        assert np.abs(full[k].data).max() > 0
        assert np.allclose(every[k].data, full[k].data, rtol=0, atol=1e-12 * np.abs(full[k].data).max())
        assert np.allclose(even[k].data + odd[k].data, full[k].data, rtol=0, atol=1e-12 * np.abs(full[k].data).max())
        assert np.allclose(even[k].data, 0.5 * full[k].data, rtol=0, atol=1e-12 * np.abs(full[k].data).max())
        assert np.allclose(twice[k].data, full[k].data, rtol=0, atol=1e-12 * np.abs(full[k].data).max())
    with pytest.raises(NotImplementedError, match="Fermi sea"):
        O.run(s, {"ahc": O.AHC(E, select_bands=[0])}, NK=[2, 2, 2], NKFFT=[2, 2, 2])
