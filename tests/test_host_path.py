"""k-paths (grid/path.py:9-294) on the host: the reference's own expected values.

* from_nodes with nk (one number, per node labels), with dk through the Haldane lattice
  (tests/test_path.py:14-55), nodes as numpy arrays (96-105);
* a path with breaks refined by a factor 4, against the reference's fixture
  path_refined.npz (K_list, labels, breaks; tests/test_path.py:76-95), and its npz round trip;
* sphere and spheroid k-lists against the reference's listed points (tests/test_path.py:126-183);
* an explicit k-list, getKline (cumulative Cartesian distance, zero steps at breaks), the
  K-batches of get_K_list, and run() refusing calculators that cannot run on a path
  (tests/test_path.py:276-297).
"""
import os

import numpy as np
import pytest
from pytest import approx

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def wb():
    import wannierberri_amd
    return wannierberri_amd


@pytest.fixture(scope="module")
def haldane(wb):
    return wb.System_R.from_arrays_npz(os.path.join(GOLDEN, "sys_haldane.npz"))


def test_path_nk(wb, haldane):
    p = wb.Path.from_nodes(system=haldane, nodes=[[0.0, 0.0, 0.0], [0.5, 0.5, 0.5]], nk=3)
    assert p.labels == {0: '1', 2: '2'}
    assert p.K_list == approx(np.array([[0, 0., 0.], [0.25, 0.25, 0.25], [0.5, 0.5, 0.5]]))
    p = wb.Path.from_nodes(system=haldane, nodes=[[0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.5, 0.0, 0.0]],
                           labels=["A", "B", "C"], nk=4)
    assert p.labels == {0: 'A', 3: 'B', 6: 'C'}
    assert p.K_list == approx(np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2], [3, 3, 3], [3, 2, 2], [3, 1, 1],
                                        [3, 0, 0]]) / 6)


def test_path_dk(wb, haldane):
    nodes = [[0.0, 0.0, 0.5], [0.0, 0.0, 0.0], [0.5, 0.5, 0.5]]
    p = wb.Path.from_nodes(haldane, nodes=nodes, dk=1.0)
    assert p.labels == {0: '1', 3: '2', 8: '3'}
    assert p.K_list[:4] == approx(np.array([[0, 0, 3], [0, 0, 2], [0, 0, 1], [0, 0, 0]]) / 6)
    assert p.K_list[3:] == approx(np.array([[0, 0, 0], [1, 1, 1], [2, 2, 2], [3, 3, 3], [4, 4, 4], [5, 5, 5]]) / 10)
    p2 = wb.Path.from_nodes(haldane, nodes=[np.array(k) for k in nodes], dk=1.0)
    assert p2.labels == p.labels and p2.K_list == approx(p.K_list)
    # length = 2 pi / dk
    p3 = wb.Path.from_nodes(haldane, nodes=nodes, length=2 * np.pi)
    assert p3.K_list == approx(p.K_list)
    with pytest.raises(ValueError):
        wb.Path.from_nodes(haldane, nodes=nodes, dk=1.0, nk=3)
    with pytest.raises(ValueError):
        wb.Path.from_nodes(haldane, nodes=nodes, dk=1.0, length=3.0)


def test_path_refined_fixture(wb, tmp_path):
    rec_lattice = 4 * (np.ones((3, 3)) - np.eye(3)) / 2
    nodes = [[0.0, 0.0, 0.5], [0.0, 0.0, 0.0], None, [0.5, 0.5, 0.5], [0, 0, 0], [0.5, 0.0, 0.5], None,
             [0.75, 0.25, 0.75], [0.125, 0.375, 0.625]]
    labels = ["A", "Gamma", "L", "Gamma", "B", "F", "G"]
    path = wb.Path.from_nodes(recip_lattice=rec_lattice, nodes=nodes, labels=labels, dk=0.1)
    ref = path.get_refined(factor=4)
    with np.load(os.path.join(GOLDEN, "results_path.npz"), allow_pickle=False) as f:
        fx = {k.split("/", 1)[1]: f[k] for k in f.files}
    d = ref.as_dict()
    assert d["K_list"].shape == fx["K_list"].shape
    assert d["K_list"] == approx(fx["K_list"])
    assert list(d["label_indices"]) == list(fx["label_indices"])
    assert list(d["label_values"]) == list(fx["label_values"])
    assert list(d["breaks"]) == list(fx["breaks"])
    assert d["recip_lattice"] == approx(fx["recip_lattice"])
    # every original point is kept, in order, every refined step is a quarter of the original
    assert len(ref.K_list) == len(path.K_list) + 3 * (len(path.K_list) - 1 - len(path.breaks))
    # npz round trip
    fn = str(tmp_path / "path.npz")
    ref.to_npz(fn)
    back = wb.Path.from_npz(fn)
    assert back.K_list == approx(ref.K_list) and back.labels == ref.labels and back.breaks == ref.breaks


def test_path_sphere(wb, haldane):
    p = wb.Path.sphere(system=haldane, r1=0.5, ntheta=3, nphi=5)
    assert p.K_list.shape == (15, 3)
    assert np.allclose(np.linalg.norm(p.K_list @ haldane.recip_lattice, axis=1), 0.5)
    expected = np.array([[0., 0., 0.07957747], [0.07957747, 0.03978874, 0.], [0., 0., -0.07957747],
                         [0., 0., 0.07957747], [0., 0.06891611, 0.], [0., 0., -0.07957747],
                         [0., 0., 0.07957747], [-0.07957747, -0.03978874, 0.], [-0., -0., -0.07957747],
                         [0., 0., 0.07957747], [-0., -0.06891611, 0.], [-0., -0., -0.07957747],
                         [0., 0., 0.07957747], [0.07957747, 0.03978874, 0.], [0., 0., -0.07957747]])
    assert np.allclose(p.K_list, expected)


def test_path_spheroid(wb, haldane):
    p = wb.Path.spheroid(system=haldane, r1=0.5, r2=1.0, ntheta=3, nphi=4)
    assert p.K_list.shape == (12, 3)
    kc = p.K_list @ haldane.recip_lattice
    theta = np.arccos(kc[:, 2] / np.linalg.norm(kc, axis=1))
    assert np.allclose(np.linalg.norm(kc, axis=1), np.sqrt((0.5 * np.sin(theta)) ** 2 + (1.0 * np.cos(theta)) ** 2))
    expected = np.array([[0., 0., 0.15915494], [0.07957747, 0.03978874, 0.], [0., 0., -0.15915494],
                         [0., 0., 0.15915494], [-0.03978874, 0.03978874, 0.], [-0., 0., -0.15915494],
                         [0., 0., 0.15915494], [-0.03978874, -0.07957747, 0.], [-0., -0., -0.15915494],
                         [0., 0., 0.15915494], [0.07957747, 0.03978874, 0.], [0., 0., -0.15915494]])
    assert np.allclose(p.K_list, expected)


def test_path_list_kline_batches(wb, haldane):
    k = np.random.default_rng(0).random((10, 3))
    p = wb.Path(system=haldane, k_list=k)
    assert p.K_list.shape == (10, 3) and np.allclose(p.K_list, k)
    p = wb.Path.from_nodes(haldane, nodes=[[0, 0, 0], [0.5, 0, 0], None, [0, 0.5, 0], [0, 0, 0]], nk=3)
    kl = p.getKline()
    steps = np.linalg.norm(np.diff(p.get_kpoints_cart(), axis=0), axis=1)
    steps[p.breaks] = 0
    assert kl[0] == 0 and kl == approx(np.concatenate(([0], np.cumsum(steps))))
    assert p.breaks == [2] and kl[3] == approx(kl[2])
    assert len(p.K_list) == 6
    blocks = p.get_K_list(k_batch=4)
    assert [len(b.K) for b in blocks] == [4, 2]
    assert np.allclose(np.concatenate([b.K for b in blocks]), p.K_list)
    assert all(b.Kp_fullBZ is None and b.factor == 1 for b in blocks)


def test_path_refuses_grid_calculators(wb, haldane):
    path = wb.Path.from_nodes(haldane, nodes=[[0.0, 0.0, 0.5], [0.0, 0.0, 0.0], [0.5, 0.5, 0.5]], dk=1.0)
    tab = wb.calculators.tabulate
    fails = {"tabulate_grid": tab.TabulatorAll({"Energy": tab.Energy(), "berry": tab.BerryCurvature()}, ibands=[0],
                                               mode="grid"),
             "ahc": wb.calculators.static.AHC(Efermi=np.array([0, 1, 2]))}
    for key, val in fails.items():
        with pytest.raises(ValueError, match=f"Calculation along a Path is running, but calculator `{key}` is not "
                                             "compatible with a Path"):
            wb.run(haldane, path, {key: val}, save_files=False, verbose=False)
    # and a path-mode table on a grid
    grid = wb.Grid(system=haldane, NK=[4, 4, 1], NKFFT=[2, 2, 1])
    with pytest.raises(ValueError, match="not compatible with a Grid"):
        wb.run(haldane, grid, {"t": tab.TabulatorAll({"Energy": tab.Energy()}, mode="path")}, save_files=False,
               verbose=False)
