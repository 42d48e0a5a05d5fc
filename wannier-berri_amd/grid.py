"""K-grid: NK = NKdiv x NKFFT mixed decomposition and K-block objects.

Follows wannierberri/grid/grid.py:86-266 (Grid, determineNK, autoNK) and
grid/Kpoint.py:9-216 (KpointBZparallel: star, absorb, equiv, divide;
exclude_equiv_points): with a point group the K-list holds one K-block per
star, weighted by the star's size, and refinement keeps only inequivalent
sub-blocks.
"""
from collections.abc import Iterable
import warnings

import numpy as np

from .symmetry import SYMMETRY_PRECISION, PointGroup


def one2three(nk):
    if nk is None:
        return None
    if isinstance(nk, Iterable):
        nk = list(nk)
        assert len(nk) == 3
    else:
        nk = (nk,) * 3
    assert all(isinstance(n, (int, np.integer)) and n > 0 for n in nk)
    return np.array(nk, dtype=int)


def _iterate_vector(v1, v2):
    return ((x, y, z) for x in range(v1[0], v2[0]) for y in range(v1[1], v2[1]) for z in range(v1[2], v2[2]))


def autoNK(NK, NKFFTrec, pointgroup=None):
    """Symmetric FFT grid between the smallest symmetric one (>= NKFFTrec) and twice it
    that best divides NK (grid.py:196-212)."""
    sym = (lambda fft: True) if pointgroup is None else pointgroup.symmetric_grid
    fft_sym = np.array([f for f in _iterate_vector(NKFFTrec, NKFFTrec * 3) if sym(f)])
    NKFFTmin = fft_sym[np.argmin(fft_sym.prod(axis=1))]
    cand = np.array([f for f in _iterate_vector(NKFFTmin, NKFFTmin * 2) if sym(f)])
    NKdiv = np.array(np.round(NK[None, :] / cand), dtype=int)
    NKdiv[NKdiv <= 0] = 1
    change = NKdiv * cand / NK[None, :]
    sel = change > 1
    change[sel] = 1. / change[sel]
    FFT = cand[np.argmax(change.min(axis=1))]
    NKdiv = np.array(np.round(NK / FFT), dtype=int)
    NKdiv[NKdiv <= 0] = 1
    return NKdiv, FFT


def determineNK(periodic, NKdiv, NKFFT, NK, NKFFT_recommended, recip_lattice, length=None, length_FFT=None,
                pointgroup=None):
    """grid.py:215-266."""
    NKdiv, NKFFT, NK = one2three(NKdiv), one2three(NKFFT), one2three(NK)
    if length is not None and NK is None:
        NK = np.array(np.round(length / (2 * np.pi) * np.linalg.norm(recip_lattice, axis=1)), dtype=int)
    if length_FFT is not None and NKFFT is None:
        NKFFT = np.array(np.round(length_FFT / (2 * np.pi) * np.linalg.norm(recip_lattice, axis=1)), dtype=int)
    if pointgroup is not None:
        for nkname, nk in (("NKdiv", NKdiv), ("NK", NK), ("NKFFT", NKFFT)):
            if nk is not None:
                assert pointgroup.symmetric_grid(nk), f" {nkname}={nk} is not consistent with the given symmetry "
    if (NKdiv is not None) and (NKFFT is not None):
        pass
    elif NK is not None:
        if NKFFT is not None:
            NKdiv = np.array(np.round(NK / NKFFT), dtype=int)
            NKdiv[NKdiv <= 0] = 1
        else:
            NKdiv, NKFFT = autoNK(NK, np.array(NKFFT_recommended), pointgroup)
    else:
        raise ValueError("you need to specify either NK or a pair (NKdiv,NKFFT) or (NK,NKFFT). "
                         f"found NK={NK}, NKdiv={NKdiv}, NKFFT={NKFFT} ")
    if NK is not None and not np.all(NK == NKFFT * NKdiv):
        warnings.warn(f" the requested k-grid {NK} was adjusted to {NKFFT * NKdiv}. ")
    notper = np.logical_not(periodic)
    NKdiv = np.array(NKdiv)
    NKFFT = np.array(NKFFT)
    NKdiv[notper] = 1
    NKFFT[notper] = 1
    return NKdiv, NKFFT


class KpointBZparallel:
    """A K-block: the NKFFT sub-grid shifted by K/NKFFT (Kpoint.py:9-181)."""

    def __init__(self, K=np.zeros(3), dK=np.ones(3), NKFFT=np.ones(3), factor=1., pointgroup=None,
                 refinement_level=0):
        self.K = np.array(K, dtype=float)
        self.dK = np.array(dK, dtype=float)
        self.NKFFT = np.array(NKFFT, dtype=int)
        self.factor = factor
        self.pointgroup = pointgroup
        self.refinement_level = refinement_level
        self.result = None
        self.result_loader = None   # restart: reads the result from its file (restart.py)
        self.was_evaluated_flag = False
        self._max = None
        self._star = None

    @property
    def Kp_fullBZ(self):
        return self.K / self.NKFFT   # Kpoint.py:75-77

    @property
    def dK_fullBZ(self):
        return self.dK / self.NKFFT

    @property
    def star(self):
        """Images of K under the point group (Kpoint.py:117-122)."""
        if self._star is None:
            self._star = np.array([self.K]) if self.pointgroup is None else self.pointgroup.star(self.K)
        return self._star

    @property
    def distGamma(self):
        """Distance of K from the nearest reciprocal lattice vector (Kpoint.py:177-181)."""
        sh = np.arange(-3, 4)
        corners = np.array([[x, y, z] for x in sh for y in sh for z in sh])
        return np.linalg.norm(((self.K % 1)[None, :] - corners).dot(self.pointgroup.recip_lattice), axis=1).min()

    def set_result(self, res):
        self.result = res
        self._max = res.max if hasattr(res, "max") else None
        self.was_evaluated_flag = True

    def get_result(self):
        if self.result is None:
            if self.result_loader is not None:
                return self.result_loader()
            raise RuntimeError("result for a K-point is called, which was never evaluated")
        return self.result

    def get_result_factor(self):
        return self.get_result() * self.factor

    def set_factor(self, factor):
        self.factor = factor

    def add_factor(self, factor):
        self.factor += factor

    @property
    def max(self):
        assert self.was_evaluated_flag, "max is called for a K-point which is not evaluated"
        return self._max * self.factor

    def absorb(self, other):
        """Take over the weight (and an evaluated result) of an equivalent K-point (Kpoint.py:124-133)."""
        if other is None:
            return
        if other.was_evaluated_flag:
            if self.was_evaluated_flag:
                raise RuntimeError(f"combining two evaluated K-points {self} and {other} should not happen")
            self.set_result(other.get_result())
        self.add_factor(other.factor)

    def equiv(self, other):
        """Same refinement level and K in each other's star (Kpoint.py:135-142)."""
        if self.refinement_level != other.refinement_level:
            return False
        dif = self.star[:, None, :] - other.star[None, :, :]
        return bool(np.linalg.norm(dif - np.round(dif), axis=2).min() < SYMMETRY_PRECISION)

    def divide(self, ndiv, periodic, use_symmetry=True):
        """Split into ndiv^3 sub-blocks (Kpoint.py:144-175); with symmetry, equivalent
        sub-blocks are merged."""
        ndiv = np.array(ndiv, dtype=int)
        ndiv[np.logical_not(periodic)] = 1
        dK_adpt = self.dK / ndiv
        shift = (-self.dK + dK_adpt) / 2.
        newfac = self.factor / np.prod(ndiv)
        out = []
        for x in range(ndiv[0]):
            for y in range(ndiv[1]):
                for z in range(ndiv[2]):
                    out.append(KpointBZparallel(K=self.K + shift + dK_adpt * np.array([x, y, z]), dK=dK_adpt,
                                                NKFFT=self.NKFFT, factor=newfac, pointgroup=self.pointgroup,
                                                refinement_level=self.refinement_level + 1))
        self.set_factor(0)
        if use_symmetry and self.pointgroup is not None:
            exclude_equiv_points(out)
        return out

    def __str__(self):
        return ("coord in rec.lattice = [ " + " , ".join(f"{x:10.6f}" for x in self.K) +
                f" ], refinement level:{self.refinement_level}, factor = {self.factor}")


def exclude_equiv_points(K_list, new_points=None):
    """Merge equivalent K-points in place (Kpoint.py:184-216).

    Points are bucketed by their distance from Gamma; within a bucket a later
    point equivalent to an earlier one is absorbed by it, except that two of
    the first n - new_points (old) points are never merged.
    """
    n = len(K_list)
    if new_points is None:
        new_points = n
    dist = np.array([K.distGamma for K in K_list])
    order = np.argsort(dist)
    dist = dist[order]
    walls = [0] + list(np.where(dist[1:] - dist[:-1] > 1e-4)[0] + 1) + [n]
    exclude = []
    for start, end in zip(walls[:-1], walls[1:]):
        for a in range(start, end):
            i = order[a]
            if i in exclude:
                continue
            for b in range(start, end):
                j = order[b]
                if i >= j:
                    continue
                if i < n - new_points and j < n - new_points:
                    continue
                if j not in exclude and K_list[i].equiv(K_list[j]):
                    exclude.append(j)
                    K_list[i].absorb(K_list[j])
    for i in sorted(exclude)[::-1]:
        del K_list[i]


class Grid:
    """k-grid NK = NKdiv * NKFFT (grid/grid.py:13-189).

    Exactly one of `system`, `real_lattice`, `recip_lattice`, `pointgroup`
    gives the lattice.  With use_symmetry (default) the point group is the
    system's (or the given one), otherwise the trivial group of the lattice
    (grid.py:44-60).
    """

    def __init__(self, system=None, real_lattice=None, recip_lattice=None, pointgroup=None, length=None,
                 NKdiv=None, NKFFT=None, NK=None, length_FFT=None, use_symmetry=True):
        found = [p is not None for p in (system, real_lattice, recip_lattice, pointgroup)]
        assert sum(found) == 1, "Exactly one of system, real_lattice, recip_lattice, pointgroup should be provided"
        periodic = np.ones(3, dtype=bool)
        NKFFT_rec = np.ones(3, dtype=int)
        if system is not None:
            real_lattice = system.real_lattice
            periodic = system.periodic
            NKFFT_rec = np.array(system.NKFFT_recommended)
        elif pointgroup is not None:
            real_lattice = pointgroup.real_lattice
        elif real_lattice is None:
            real_lattice = 2 * np.pi * np.linalg.inv(np.array(recip_lattice)).T
        self.real_lattice = np.array(real_lattice)
        self.recip_lattice = 2 * np.pi * np.linalg.inv(self.real_lattice).T
        if use_symmetry and system is not None:
            self.pointgroup = system.pointgroup
        elif use_symmetry and pointgroup is not None:
            self.pointgroup = pointgroup
        else:
            self.pointgroup = PointGroup([], real_lattice=self.real_lattice)
        self.div, self.FFT = determineNK(periodic, NKdiv, NKFFT, NK, NKFFT_rec, self.recip_lattice,
                                         length=length, length_FFT=length_FFT, pointgroup=self.pointgroup)

    @property
    def dense(self):
        return self.div * self.FFT

    @property
    def str_short(self):
        return f"Grid() with NKdiv={self.div}, NKFFT={self.FFT}, NKtot={self.dense}"

    @property
    def points_FFT(self):
        N = self.FFT
        return np.array([[i / N[0], j / N[1], l / N[2]] for i in range(N[0]) for j in range(N[1])
                         for l in range(N[2])])

    def get_K_list(self, use_symmetry=True, k_batch=None):
        """K-blocks of the grid; with use_symmetry one per star of the point group,
        carrying the star's weight (grid.py:148-189)."""
        dK = 1. / self.div
        factor = 1. / np.prod(self.div)
        K = [[[KpointBZparallel(K=np.array([x, y, z]) * dK, dK=dK, NKFFT=self.FFT, factor=factor,
                                pointgroup=self.pointgroup, refinement_level=0)
               for z in range(self.div[2])] for y in range(self.div[1])] for x in range(self.div[0])]
        if use_symmetry:
            # x fastest, as grid.py:172-181: the first point of each star met in this order
            # represents it, which matters where the block's own result is not symmetric
            # (tetrahedron corners)
            for z in range(self.div[2]):
                for y in range(self.div[1]):
                    for x in range(self.div[0]):
                        KP = K[x][y][z]
                        if KP is None:
                            continue
                        star = [tuple(k) for k in np.array(np.round(KP.star * self.div), dtype=int) % self.div]
                        for k in star:
                            if k != (x, y, z):
                                KP.absorb(K[k[0]][k[1]][k[2]])
                                K[k[0]][k[1]][k[2]] = None
        return [Kp for Kyz in K for Kz in Kyz for Kp in Kz if Kp is not None]


# ---------------------------------------------------------------------------
# tetrahedral K-blocks (grid/grid_tetra.py:8-152, grid/Kpoint_tetra.py:9-86)
# ---------------------------------------------------------------------------

# edges of a tetrahedron by their vertices, and the two vertices off each edge
TETRA_EDGES = [[0, 1], [0, 2], [0, 3], [1, 2], [1, 3], [2, 3]]
TETRA_EDGES_COMPLEMENT = [sorted({0, 1, 2, 3} - set(e)) for e in TETRA_EDGES]


def tetra_volume(vertices):
    return abs(np.linalg.det(vertices[1:] - vertices[0][None, :])) / 6.


class KpointBZtetra(KpointBZparallel):
    """A K-block of the tetrahedral grid: the NKFFT sub-grid shifted to the centre of a
    tetrahedron, whose vertices (relative to the centre, in units of the reciprocal lattice
    divided by NKFFT) are the corners of the tetrahedron method (Kpoint_tetra.py:13-86)."""

    def __init__(self, vertices, basis=np.eye(3), K=0, NKFFT=np.ones(3), factor=1., refinement_level=0,
                 split_level=0):
        vertices = np.array(vertices, dtype=float)
        assert vertices.shape == (4, 3)
        cntr = vertices.mean(axis=0)
        super().__init__(K=np.array(K, dtype=float) + cntr, dK=np.ones(3), NKFFT=NKFFT, factor=factor,
                         pointgroup=None, refinement_level=refinement_level)
        self.basis = basis
        self.vertices = vertices - cntr
        self.split_level = split_level

    def copy(self):
        return KpointBZtetra(vertices=self.vertices, K=self.K, NKFFT=self.NKFFT, factor=self.factor, basis=self.basis,
                             refinement_level=self.refinement_level, split_level=self.split_level)

    @property
    def vertices_fullBZ(self):
        return self.vertices / self.NKFFT

    @property
    def _edge_lengths(self):
        edges = np.array([self.vertices[j] - self.vertices[i] for i, j in TETRA_EDGES]).dot(self.basis)
        return np.linalg.norm(edges, axis=1)

    @property
    def size(self):
        return max(self._edge_lengths)

    def _i_max_edge(self):
        """The longest edge; among edges equal to 1e-6, the first in TETRA_EDGES (Kpoint_tetra.py:38-48)."""
        lengths = self._edge_lengths
        srt = np.argsort(lengths)
        return min(srt[np.where(lengths[srt[-1]] - lengths[srt] < 1e-6)])

    def divide(self, ndiv=2, periodic=(True, True, True), use_symmetry=True, refine=True):
        """Split along the longest edge into ndiv tetrahedra (Kpoint_tetra.py:54-86)."""
        if not np.all(periodic):
            raise ValueError("tetrahedron grid can be used only for 3D-periodic systems")
        if isinstance(ndiv, Iterable):
            ndiv = ndiv[0]
        ie = self._i_max_edge()
        edge, comp = TETRA_EDGES[ie], TETRA_EDGES_COMPLEMENT[ie]
        v0 = self.vertices[edge[0]]
        dv = (self.vertices[edge[1]] - v0) / ndiv
        out = [KpointBZtetra(K=self.K, vertices=np.array([self.vertices[comp[0]], self.vertices[comp[1]],
                                                          v0 + i * dv, v0 + (i + 1) * dv]),
                             factor=self.factor / ndiv, basis=self.basis,
                             refinement_level=self.refinement_level + int(refine),
                             split_level=self.split_level + int(not refine), NKFFT=self.NKFFT)
               for i in range(ndiv)]
        self.set_factor(0)
        return out


class GridTetra:
    """K-grid of tetrahedra (grid/grid_tetra.py:8-152): the reciprocal cell (or `IBZ_tetra`)
    split into tetrahedra by volume and by edge length, each the centre of an NKFFT
    sub-grid; use with the tetrahedron method."""

    def __init__(self, system, length, NKFFT=None, length_FFT=None, IBZ_tetra=None, weights=None,
                 refine_by_volume=True, refine_by_size=True, length_size=None):
        _, self.FFT = determineNK(system.periodic, 1, NKFFT, None, system.NKFFT_recommended, system.recip_lattice,
                                  length=None, length_FFT=length_FFT, pointgroup=system.pointgroup)
        self.div = np.ones(3, dtype=int)
        self.recip_lattice_reduced = system.recip_lattice / self.FFT[:, None]
        if IBZ_tetra is None:   # the reciprocal cell as 5 tetrahedra (grid_tetra.py:44-50)
            tetrahedra = np.array([[[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]],
                                   [[1, 0, 1], [0, 0, 1], [1, 0, 0], [1, 1, 1]],
                                   [[1, 1, 0], [1, 0, 0], [0, 1, 0], [1, 1, 1]],
                                   [[0, 1, 1], [0, 0, 1], [0, 1, 0], [1, 1, 1]],
                                   [[0, 0, 1], [0, 1, 0], [1, 0, 0], [1, 1, 1]]], dtype=float) - 0.5
        else:
            tetrahedra = np.array(IBZ_tetra, dtype=float)
        vol = np.array([tetra_volume(t) for t in tetrahedra])
        weights = vol / vol.sum() if weights is None else np.array(weights) * vol
        self.K_list = [KpointBZtetra(vertices=t, K=0, NKFFT=self.FFT, factor=w, basis=self.recip_lattice_reduced)
                       for t, w in zip(tetrahedra, weights)]
        if refine_by_volume:
            dkmax = 2 * np.pi / length
            self._split_volume(dkmax ** 3 / np.linalg.det(self.recip_lattice_reduced))
        if refine_by_size:
            if length_size is None:
                length_size = 0.5 * length
            self._split_size((2 * np.pi / length_size) * np.sqrt(2))

    def _split_size(self, dkmax):
        while max(K.size for K in self.K_list) >= dkmax:
            kl = []
            for K in self.K_list:
                kl += K.divide(ndiv=2, refine=False) if K.size > dkmax else [K]
            self.K_list = kl

    def _split_volume(self, vmax):
        while True:
            vols = [tetra_volume(K.vertices) for K in self.K_list]
            if max(vols) < vmax:
                break
            kl = []
            for K, v in zip(self.K_list, vols):
                kl += K.divide(ndiv=2, refine=False) if v > vmax else [K]
            self.K_list = kl

    @property
    def dense(self):
        return self.FFT

    @property
    def points_FFT(self):
        N = self.FFT
        return np.array([[i / N[0], j / N[1], l / N[2]] for i in range(N[0]) for j in range(N[1])
                         for l in range(N[2])])

    def get_K_list(self, use_symmetry=True, k_batch=None):
        return [K.copy() for K in self.K_list]
