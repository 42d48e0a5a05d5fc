"""wbmi_eigh_noalloc (the eigensolver with the library's fixed scratch) against wbmi_eigh."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from wannierberri_amd import _lib
    return _lib.get_context(0)


def _hermitian(rng, nk, nw):
    A = rng.standard_normal((nk, nw, nw)) + 1j * rng.standard_normal((nk, nw, nw))
    return 0.5 * (A + A.conj().swapaxes(1, 2))


@pytest.mark.parametrize("nw,prec", [(80, 0), (80, 1), (48, 0)])
def test_eigh_noalloc_matches_eigh(ctx, nw, prec):
    """Two calls in a row (the second with fewer k-points) through the C ABI: energies equal to
    those of wbmi_eigh with an allocated workspace to 1e-12 |H|, eigenvectors equal up to the
    phase of each column (non-degenerate random spectra), and the residual of both small."""
    import torch
    from wannierberri_amd import _lib
    rng = np.random.default_rng(21 + nw + prec)
    cdt = torch.complex128 if prec == 0 else torch.complex64
    for nk in (37, 9):
        say(f"test_eigh_noalloc nw={nw} prec={prec} nk={nk}")
        H = _hermitian(rng, nk, nw)
        Hd = torch.from_numpy(H).cuda()
        out = []
        for noalloc in (False, True):
            E = torch.empty((nk, nw), dtype=torch.float64, device="cuda")
            U = torch.empty((nk, nw, nw), dtype=cdt, device="cuda")
            f = ctx.eigh_noalloc if noalloc else ctx.eigh
            if prec == 0:
                f(Hd, nw * nw, E, U, 0, nw, refine=16)
            else:
                f(None, nw * nw, E, U, 1, nw, refine=16, H64=Hd)
            torch.cuda.synchronize()
            out.append((E.cpu().numpy(), U.cpu().numpy().astype(np.complex128)))
        (E0, U0), (E1, U1) = out
        scale = np.abs(E0).max()
        assert np.abs(E1 - E0).max() < 1e-12 * scale
        ov = np.abs(np.einsum("kji,kji->ki", U0.conj(), U1))
        assert np.abs(ov - 1).max() < (1e-10 if prec == 0 else 1e-5)
        res = np.abs(np.einsum("kab,kbc->kac", H, U1) - U1 * E1[:, None, :]).max() / scale
        assert res < (1e-12 if prec == 0 else 1e-6)


def test_eigh_noalloc_refuses_oversize(ctx):
    """More k-points than the scratch holds: a clear error, nothing launched."""
    import torch
    from wannierberri_amd import _lib
    nw = 128
    per_k = 4 * nw * nw * 16
    nk = _lib.lib().wbmi_eigh_noalloc_max_bytes() // per_k + 1
    E = torch.empty((nk, nw), dtype=torch.float64, device="cuda")
    U = torch.empty((nk, nw, nw), dtype=torch.complex128, device="cuda")
    H = torch.zeros((1, nw, nw), dtype=torch.complex128, device="cuda")
    with pytest.raises(ValueError, match="scratch"):
        ctx.eigh_noalloc(H, 0, E, U, 0, nw, refine=4)
