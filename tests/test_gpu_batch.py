"""Batched k-list evaluation: the per-nw tile autotuner and the graph-captured multi-system step.

* every workgroup tile of the k-list transform (wbmi_r2k_klist_tile) gives the same
  output as the default tile (the same sums in the same order per element: equal to
  1e-14), the autotuner records one of them per (nw, precision), and its table survives
  a save / load;
* KlistGraph: k-lists of three systems with different nw captured as one HIP graph; three
  replays, with eager library work between them, each equal to run_klists on the same
  lists (1e-12 relative), in both precision forms.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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


def _klists(rng, n, centre_count=4, max_size=300):
    centres = rng.random((centre_count, 3))
    sizes = np.exp(rng.uniform(np.log(5), np.log(max_size), n)).astype(int)
    return [centres[i % centre_count] + 0.02 * rng.standard_normal((m, 3)) for i, m in enumerate(sizes)]


@pytest.mark.parametrize("nw,prec", [(8, 0), (24, 1)])
def test_autotune_klist_tiles(wb, nw, prec, tmp_path):
    import torch
    from wannierberri_amd import _lib, autotune, models
    from wannierberri_amd.data_k import device_system
    say(f"test_autotune_klist_tiles nw={nw} prec={prec}")
    ctx = _lib.get_context(0)
    s = models.synthetic(num_wann=nw, Rmax=2, seed=40 + nw)
    chans = [("Ham", 0), ("Ham", 1), ("AA", 0), ("rotAA", 0)]
    shapes = autotune.tile_shapes()
    assert len(shapes) >= 4 and shapes[0] == (64, 128)
    dev = device_system(s, ctx)
    stack = dev.stack(tuple(chans))
    kl = torch.from_numpy(np.random.default_rng(1).random((777, 3))).cuda()
    cdt = torch.complex128 if prec == 0 else torch.complex64
    outs = []
    for tile in range(len(shapes)):
        out = torch.empty((777, stack.shape[1], nw, nw), dtype=cdt, device="cuda")
        ctx.r2k_klist(dev.iRvec_dev, stack, kl, out, prec, tile=tile)
        outs.append(out.cpu().numpy())
    for tile in range(1, len(shapes)):
        assert np.abs(outs[tile] - outs[0]).max() <= 1e-14 * np.abs(outs[0]).max(), tile
    best, times = autotune.tune_klist(s, chans, nk=2048, prec=prec)
    say(f"  tiles {shapes}: ms {times} -> {best}")
    assert best in times and ctx.tuning[("klist_tile", nw, prec)] == best
    path = str(tmp_path / "tuning.json")
    autotune.save_table(path)
    ctx.tuning.clear()
    assert autotune.load_table(path) >= 1 and ctx.tuning[("klist_tile", nw, prec)] == best


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_klist_graph_mixed_nw(wb, precision):
    import torch
    from wannierberri_amd import models
    say(f"test_klist_graph_mixed_nw {precision}")
    rng = np.random.default_rng(4)
    Ef = np.linspace(-6, 6, 201)
    calcs = {"ahc": wb.calculators.static.AHC(Efermi=Ef)}
    jobs = [(models.synthetic(num_wann=nw, Rmax=2, seed=50 + nw), _klists(rng, 7)) for nw in (8, 16, 24)]
    pk = {"precision": precision}
    eager = [wb.run_klists(s, kls, calcs, parameters_K=pk, batch_points=600).results["ahc"].data.copy()
             for s, kls in jobs]
    g = wb.KlistGraph(jobs, calcs, parameters_K=pk, batch_points=600)
    assert g.npoints == sum(len(k) for _, kls in jobs for k in kls)
    for rep in range(3):
        res = g.replay()
        torch.cuda.synchronize()
        for (s, kls), r, e in zip(jobs, res, eager):
            got = r.results["ahc"].data
            err = np.abs(got - e).max() / np.abs(e).max()
            assert err < 1e-12, (rep, s.num_wann, err)
        # eager library work between replays
        wb.run_klists(jobs[0][0], jobs[0][1][:2], calcs, parameters_K=pk)
        torch.cuda.synchronize()
