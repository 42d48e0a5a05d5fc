"""Model systems for the Berry-path workloads.

* `haldane()` -- the 2-orbital Haldane model (wannierberri/models.py:51-82,
  built as System_R with the PythTB conventions of
  system/system_tb_py.py:12-157: orbitals at (1/3,1/3) and (2/3,2/3),
  2-D lattice padded with z = 1 A, periodic [T, T, F], internal terms only).
* `synthetic(...)` -- seeded random Wannier models on a cube of R-vectors with
  Wigner-Seitz degeneracy weights, Hermitian H(-R) = H(R)^+ and position
  matrices A(-R) = A(R)^+ (the bcc nw=32 / nw=128 workloads).
* `spinor(...)` -- 2 x nw spin-block model with an SOC-like coupling block.

All generators use numpy.random.Generator(PCG64(seed)).
"""
import numpy as np

from .system import System_R

BCC_A = 3.0


def bcc_lattice(a=BCC_A):
    return 0.5 * a * np.array([[-1., 1., 1.], [1., -1., 1.], [1., 1., -1.]])


def _cube(Rmax):
    r = np.arange(-Rmax, Rmax + 1)
    return np.array([[x, y, z] for x in r for y in r for z in r], dtype=int)


def ws_weights(iRvec, lattice, ndiv, tol=1e-6):
    """1 / (number of supercell images of R at the same distance), supercell = ndiv x lattice.

    The degeneracy count of the Wigner-Seitz construction (fourier/rvectors.py:570-609).
    """
    shifts = np.array([[x, y, z] for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)]) * np.array(ndiv)
    d0 = np.linalg.norm(iRvec.dot(lattice), axis=1)
    dall = np.linalg.norm((iRvec[:, None, :] + shifts[None, :, :]).dot(lattice), axis=2)
    ndeg = np.sum(np.abs(dall - d0[:, None]) < tol * max(1.0, d0.max()), axis=1)
    return 1.0 / ndeg


def _conj_R(iRvec):
    index = {tuple(R): i for i, R in enumerate(iRvec)}
    minus = np.array([index[tuple(-R)] for R in iRvec])

    def conj(X):
        return X[minus].swapaxes(1, 2).conj()
    return conj


def haldane(delta=0.2, hop1=-1.0, hop2=0.15, phi=np.pi / 2):
    """Haldane model as System_R (models.py:51-82; PythTB adapter system_tb_py.py:118-145)."""
    lat2 = np.array([[1.0, 0.0], [0.5, np.sqrt(3.0) / 2.0]])
    orb = np.array([[1. / 3., 1. / 3.], [2. / 3., 2. / 3.]])
    t2 = hop2 * np.exp(1.j * phi)
    t2c = t2.conjugate()
    # (amplitude, from orbital i, to orbital j, R): H[R][i, j] += amp, H[-R][j, i] += conj(amp)
    hops = [(hop1, 0, 1, (0, 0)), (hop1, 1, 0, (1, 0)), (hop1, 1, 0, (0, 1)),
            (t2, 0, 0, (1, 0)), (t2, 1, 1, (1, -1)), (t2, 1, 1, (0, 1)),
            (t2c, 1, 1, (1, 0)), (t2c, 0, 0, (1, -1)), (t2c, 0, 0, (0, 1))]
    Rs = {(0, 0)}
    for _, _, _, R in hops:
        Rs.add(R)
        Rs.add((-R[0], -R[1]))
    iRvec2 = np.unique(np.array(sorted(Rs)), axis=0)
    iRvec = np.hstack([iRvec2, np.zeros((len(iRvec2), 1), dtype=int)])
    index = {tuple(R): i for i, R in enumerate(iRvec2)}
    H = np.zeros((len(iRvec), 2, 2), dtype=complex)
    for amp, i, j, R in hops:
        H[index[R], i, j] += amp
        H[index[(-R[0], -R[1])], j, i] += np.conjugate(amp)
    H[index[(0, 0)], 0, 0] = -delta
    H[index[(0, 0)], 1, 1] = delta
    real = np.eye(3)
    real[:2, :2] = lat2
    wcc_red = np.hstack([orb % 1.0, np.zeros((2, 1))])
    return System_R(real, iRvec, wcc_red.dot(real), {"Ham": H}, periodic=(True, True, False),
                    force_internal_terms_only=True, name="Haldane")


def synthetic(num_wann=32, Rmax=4, seed=2, lattice=None, decay=3.0, h_std=1.0, onsite=4.0, a_std=0.1,
              with_AA=True, with_BBCC=False):
    """Random Wannier model on R in {-Rmax..Rmax}^3 (the bcc nw=32 / nw=128 workloads).

    H_R: complex Gaussian, std h_std * exp(-|R|/decay) eV, times the WS weight,
    Hermitian in (R, m, n); onsite diagonal uniform in [-onsite, onsite] eV.
    AA_R: Wannier centres uniform in the cell (reduced coordinates), zero
    diagonal at R = 0, off-diagonal std a_std * exp(-|R|/decay) A.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    lattice = bcc_lattice() if lattice is None else np.array(lattice, dtype=float)
    iRvec = _cube(Rmax)
    nR = len(iRvec)
    dist = np.linalg.norm(iRvec.dot(lattice), axis=1)
    w = ws_weights(iRvec, lattice, [2 * Rmax + 1] * 3)
    conj = _conj_R(iRvec)
    iR0 = int(np.where((iRvec == 0).all(axis=1))[0][0])
    nw = num_wann

    def gauss(shape, std):
        z = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2)
        return z * std.reshape((-1,) + (1,) * (len(shape) - 1))

    H = gauss((nR, nw, nw), h_std * np.exp(-dist / decay) * w)
    H = 0.5 * (H + conj(H))
    H[iR0] += np.diag(rng.uniform(-onsite, onsite, nw))
    wcc_red = rng.uniform(0.0, 1.0, (nw, 3))
    mats = {"Ham": H}
    if with_AA:
        A = gauss((nR, nw, nw, 3), a_std * np.exp(-dist / decay) * w)
        A = 0.5 * (A + conj(A))
        A[iR0, np.arange(nw), np.arange(nw), :] = 0.0
        mats["AA"] = A
    if with_BBCC:
        for key in ("BB", "CC"):
            X = gauss((nR, nw, nw, 3), a_std * np.exp(-dist / decay) * w)
            if key == "CC":
                X = 0.5 * (X + conj(X))
            mats[key] = X
    return System_R(lattice, iRvec, wcc_red.dot(lattice), mats, name=f"synthetic_nw{nw}_seed{seed}")


def klist_batch(n_lists=512, nmin=8, nmax=4096, n_centres=16, sigma=0.02, seed=4):
    """Refinement-like k-lists: sizes log-uniform in [nmin, nmax], points
    ~ N(centre, sigma) (reduced coordinates, wrapped into [0, 1)) around
    `n_centres` centres uniform in the cell; list i uses centre i % n_centres."""
    rng = np.random.Generator(np.random.PCG64(seed))
    centres = rng.uniform(0.0, 1.0, (n_centres, 3))
    sizes = np.exp(rng.uniform(np.log(nmin), np.log(nmax + 1), n_lists)).astype(int).clip(nmin, nmax)
    return [(centres[i % n_centres] + sigma * rng.standard_normal((n, 3))) % 1.0 for i, n in enumerate(sizes)]


def spinor(num_wann_half=32, Rmax=5, seed=3, soc_std=0.05, decay=3.0, a_std=0.1):
    """2 x num_wann_half spin-block model: both spin blocks carry the same random
    Hamiltonian (synthetic recipe) and a random SOC-like block couples them
    (std soc_std * exp(-|R|/decay) eV); bands are ordered (orbital, spin)."""
    base = synthetic(num_wann=num_wann_half, Rmax=Rmax, seed=seed, decay=decay, a_std=a_std)
    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    iRvec = base.rvec.iRvec
    nR = len(iRvec)
    lattice = base.real_lattice
    dist = np.linalg.norm(iRvec.dot(lattice), axis=1)
    w = ws_weights(iRvec, lattice, [2 * Rmax + 1] * 3)
    conj = _conj_R(iRvec)
    nh = num_wann_half
    nw = 2 * nh
    S = (rng.standard_normal((nR, nh, nh)) + 1j * rng.standard_normal((nR, nh, nh))) / np.sqrt(2)
    S *= (soc_std * np.exp(-dist / decay) * w)[:, None, None]
    H = np.zeros((nR, nw, nw), dtype=complex)
    H0 = base.get_R_mat("Ham")
    H[:, 0::2, 0::2] = H0
    H[:, 1::2, 1::2] = H0
    H[:, 0::2, 1::2] = S
    H = 0.5 * (H + conj(H))
    A = np.zeros((nR, nw, nw, 3), dtype=complex)
    A0 = base.get_R_mat("AA")
    A[:, 0::2, 0::2] = A0
    A[:, 1::2, 1::2] = A0
    wcc = np.repeat(base.wannier_centers_cart, 2, axis=0)
    return System_R(lattice, iRvec, wcc, {"Ham": H, "AA": A}, name=f"spinor_nw{nw}_seed{seed}")
