"""Restart files of run() (run_grid.py:242-348,416-440) on the host, with the oracle as the
K-block evaluator (a data_k_class of its own, as the reference's plug-in point allows):

* two refinement iterations in one run equal one iteration written with dump_results,
  then one more iteration read back with restart=True (the reference's test_Fe_sym_refine
  pattern, tests/test_run.py:556-582), for adaptive refinement without and with a point group;
* restart_iteration picks the factors of an earlier iteration; a record of a tetrahedral grid
  or of tabulated results is refused.
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


class _Block:
    osys = None

    def __init__(self, system, dK=None, grid=None, Kpoint=None, **kw):
        self.d = O.OracleDataK(self.osys, dK=dK, NKFFT=grid.FFT)


class _Calc:
    allow_grid = True
    allow_path = False
    smoother = None
    parity = (-1, 1)

    def __init__(self, wb, calc):
        self.wb, self.calc = wb, calc

    def __call__(self, data):
        r = self.calc(data.d)
        return self.wb.EnergyResult(r.Efermi, r.data, transformTR=-1, transformInv=1)


@pytest.mark.parametrize("sym", [False, True])
def test_restart_refinement(wb, tmp_path, sym):
    path = os.path.join(GOLDEN, "sys_haldane.npz")
    s = wb.System_R.from_arrays_npz(path)
    if sym:
        s.set_pointgroup(["C3z"])
    _Block.osys = O.load_golden_system(path)
    Ef = np.linspace(-3, 3, 11)
    calcs = {"ahc": _Calc(wb, O.AHC(Ef)), "dos": _Calc(wb, O.DOS(Ef))}
    grid = wb.Grid(system=s, NK=[8, 8, 1], NKFFT=[4, 4, 1])
    kw = dict(use_irred_kpt=sym, symmetrize=sym, data_k_class=_Block, save_files=False, verbose=False)
    full = wb.run(s, grid, calcs, adpt_num_iter=2, **kw)
    rec = str(tmp_path / "rec")
    part = wb.run(s, grid, calcs, adpt_num_iter=1, dump_results=True, file_Klist_path=rec, **kw)
    assert os.path.exists(os.path.join(rec, "K_list.npz"))
    assert os.path.exists(os.path.join(rec, "factors_iter-00000001.npy"))
    cont = wb.run(s, grid, calcs, adpt_num_iter=1, restart=True, file_Klist_path=rec, **kw)
    for k in calcs:
        a, b = full.results[k].data, cont.results[k].data
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max(), k
        # the record after one iteration differs from the total after two
        assert np.abs(part.results[k].data - a).max() > 0 or k == "dos"
    # restart_iteration = 1 restores the state after the first iteration (factors of it)
    again = wb.run(s, grid, calcs, adpt_num_iter=0, restart=True, restart_iteration=1, file_Klist_path=rec, **kw)
    for k in calcs:
        assert np.abs(again.results[k].data - part.results[k].data).max() <= 1e-12 * np.abs(part.results[k].data).max()


def test_restart_refused(wb, tmp_path):
    s = wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_haldane.npz"))
    g = wb.grid.GridTetra.__new__(wb.grid.GridTetra)   # refused before the grid is used
    with pytest.raises(NotImplementedError):
        wb.run(s, g, {}, dump_results=True, file_Klist_path=str(tmp_path / "r"), save_files=False, verbose=False)
    # other result types are not written
    from wannierberri_amd import restart
    kb = wb.result.KBandResult(np.zeros((1, 2)))
    with pytest.raises(NotImplementedError):
        restart.dump_result(str(tmp_path), 0, wb.ResultDict({"kb": kb}))


def test_restart_tabresult_round_trip(wb, tmp_path):
    """A K-point's tabulated result (TABresult of KBandResults with their parities, path or grid
    mode) written to its restart record and read back equals the original, without pickling."""
    from wannierberri_amd import restart
    rng = np.random.default_rng(3)
    k = rng.random((5, 3))
    res = {"Energy": wb.result.KBandResult(rng.random((5, 4)), transformTR=1, transformInv=1),
           "berry": wb.result.KBandResult(rng.random((5, 4, 3)), transformTR=-1, transformInv=1),
           "plain": wb.result.KBandResult(rng.random((5, 4)))}
    tab = wb.result.TABresult(k, 2 * np.eye(3), results=res, mode="path", save_mode="txt")
    e = wb.result.EnergyResult(np.linspace(0, 1, 3), rng.random(3))
    restart.dump_result(str(tmp_path), 7, wb.ResultDict({"tab": tab, "e": e}))
    back = restart.load_result(str(tmp_path), 7, {"tab": None, "e": None}).results
    t = back["tab"]
    assert isinstance(t, wb.result.TABresult) and t.mode == "path" and t.save_mode == "txt"
    assert np.array_equal(t.kpoints, tab.kpoints) and np.array_equal(t.recip_lattice, tab.recip_lattice)
    assert sorted(t.results) == sorted(res)
    for q, kb in res.items():
        assert np.array_equal(t.results[q].data, kb.data)
        assert (t.results[q].transformTR, t.results[q].transformInv) == (kb.transformTR, kb.transformInv)
    assert np.array_equal(back["e"].data, e.data)
