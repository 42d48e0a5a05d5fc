"""k-paths for tabulation along lines (wannierberri/grid/path.py:9-294).

A Path holds an explicit list of k-points in reduced coordinates, the labels
of some of them ({index: label}) and the indices after which the line is broken.
run() splits it into KpointBZpath blocks of `k_batch` points, each evaluated as
a Data_K on that k-list (the R -> k phase GEMM), and puts the tabulated
results back in path order (tabresult.py:171-185).
"""
import warnings
from collections.abc import Iterable

import numpy as np

from .symmetry import PointGroup


class KpointBZpath:
    """A batch of path k-points (grid/Kpoint.py:95-101): K is [n, 3] reduced coordinates;
    no FFT grid and no symmetry (factor 1)."""

    def __init__(self, K=np.zeros(3), pointgroup=None):
        self.K = np.array(K, dtype=float).reshape(-1, 3)
        self.pointgroup = pointgroup
        self.factor = 1.
        self.was_evaluated_flag = False

    @property
    def Kp_fullBZ(self):
        return None   # a k-list, not a shifted FFT grid

    def __str__(self):
        return "coord in rec.lattice = [ " + " , ".join(f"{x:10.6f}" for x in self.K.reshape(-1)) + " ]"


class Path:
    """k-path (path.py:9-64): `k_list` [n, 3] reduced coordinates, `labels` {index: name},
    `breaks` (indices after which the line is not continued).  Exactly one of system,
    real_lattice, recip_lattice, pointgroup gives the lattice; the point group is always the
    trivial group of that lattice (no symmetry along a path)."""

    def __init__(self, system=None, pointgroup=None, real_lattice=None, recip_lattice=None, k_list=(), labels=None,
                 breaks=None):
        found = [p is not None for p in (system, real_lattice, recip_lattice, pointgroup)]
        assert sum(found) == 1, "Exactly one of system, real_lattice, recip_lattice, pointgroup should be provided"
        if system is not None:
            real_lattice = system.real_lattice
        elif pointgroup is not None:
            real_lattice = pointgroup.real_lattice
        if real_lattice is not None:
            self.pointgroup = PointGroup([], real_lattice=real_lattice)
        else:
            self.pointgroup = PointGroup([], recip_lattice=recip_lattice)
        self.breaks = [] if breaks is None else list(breaks)
        self.labels = {} if labels is None else labels
        self.K_list = np.array(k_list, dtype=float).reshape(-1, 3)
        self.div = self.K_list.shape[0]

    @property
    def recip_lattice(self):
        return self.pointgroup.recip_lattice

    @property
    def real_lattice(self):
        return self.pointgroup.real_lattice

    # ------------------------------------------------------------------ constructors
    @classmethod
    def from_nodes(cls, system=None, real_lattice=None, recip_lattice=None, pointgroup=None, nodes=(), labels=None,
                   length=None, dk=None, nk=None):
        """Straight segments between consecutive nodes (path.py:133-196).  A None node breaks
        the path.  Each segment from a to b holds n - 1 points a + j (b - a) / (n - 1),
        j < n - 1, with n from `nk` (one number, or one per segment) or
        n = max(2, round(|b - a|_cart / dk) + 1), dk = 2 pi / length; the last node closes the
        path.  Labels: one per non-None node (default '1', '2', ...), put at the index where
        the node lands."""
        self = cls(system=system, real_lattice=real_lattice, recip_lattice=recip_lattice, pointgroup=pointgroup)
        nodes = [None if k is None else np.array(k, dtype=float).reshape(3) for k in nodes]
        real = [k for k in nodes if k is not None]
        if labels is None:
            labels = [str(i + 1) for i in range(len(real))]
        it = iter(labels)
        node_label = [None if k is None else next(it) for k in nodes]
        if length is not None:
            assert length > 0
            if dk is not None:
                raise ValueError("'length' and  'dk' cannot be set together")
            dk = 2 * np.pi / length
        if dk is not None and nk is not None:
            raise ValueError("'nk' cannot be set together with 'length' or 'dk' ")
        if dk is None and nk is None:
            raise ValueError("one of 'nk', 'dk' or 'length' is needed")
        nk_iter = iter(nk) if isinstance(nk, Iterable) else None
        points, new_labels, breaks = [], {}, []
        for i in range(len(nodes) - 1):
            a, b = nodes[i], nodes[i + 1]
            if a is None:
                continue
            new_labels[len(points)] = node_label[i]
            if b is None:
                points.append(a)
                breaks.append(len(points) - 1)
                continue
            if nk is not None:
                n = next(nk_iter) if nk_iter is not None else nk
            else:
                n = max(2, int(round(np.linalg.norm((b - a) @ self.recip_lattice) / dk)) + 1)
            for t in np.linspace(0, 1., n - 1, endpoint=False):
                points.append(a + t * (b - a))
        points.append(nodes[-1])
        new_labels[len(points) - 1] = node_label[-1]
        self.K_list = np.array(points, dtype=float).reshape(-1, 3)
        self.div = self.K_list.shape[0]
        self.labels = new_labels
        self.breaks = breaks
        return self

    @classmethod
    def spheroid(cls, system=None, real_lattice=None, recip_lattice=None, pointgroup=None, r1=0.1, r2=0.1, ntheta=10,
                 nphi=10, origin=None):
        """k-points on the spheroid (x^2 + y^2) / r1^2 + z^2 / r2^2 = 1 (Cartesian, path.py:68-91):
        theta in [0, pi] (ntheta values), phi in [0, 2 pi] (nphi values), phi the slow index;
        the origin is subtracted in reduced coordinates."""
        self = cls(system=system, real_lattice=real_lattice, recip_lattice=recip_lattice, pointgroup=pointgroup,
                   labels={0: "sphere"})
        origin = np.zeros(3) if origin is None else np.array(origin, dtype=float)
        theta, phi = np.meshgrid(np.linspace(0, np.pi, ntheta), np.linspace(0, 2 * np.pi, nphi))
        cart = np.stack([r1 * np.cos(phi) * np.sin(theta), r1 * np.sin(phi) * np.sin(theta), r2 * np.cos(theta)],
                        axis=-1).reshape(-1, 3)
        self.K_list = cart @ np.linalg.inv(self.recip_lattice) - origin[None, :]
        self.div = self.K_list.shape[0]
        return self

    @classmethod
    def sphere(cls, system=None, real_lattice=None, recip_lattice=None, pointgroup=None, r1=0.1, ntheta=10, nphi=10,
               origin=None):
        return cls.spheroid(system=system, real_lattice=real_lattice, recip_lattice=recip_lattice,
                            pointgroup=pointgroup, r1=r1, r2=r1, ntheta=ntheta, nphi=nphi, origin=origin)

    @classmethod
    def seekpath(cls, cell=None, dk=0.05, with_time_reversal=True, lattice=None, positions=None, numbers=None,
                 twoD_direction=None):
        raise NotImplementedError("Path.seekpath needs the seekpath package, which is not available here; "
                                  "use Path.from_nodes")

    # ------------------------------------------------------------------ files
    def as_dict(self):
        idx = sorted(self.labels.keys())
        return {"recip_lattice": self.recip_lattice, "K_list": self.K_list, "label_indices": idx,
                "label_values": [self.labels[i] for i in idx], "breaks": self.breaks}

    @classmethod
    def from_dict(cls, dct):
        return cls(recip_lattice=np.asarray(dct["recip_lattice"]), k_list=np.asarray(dct["K_list"]),
                   labels={int(i): str(v) for i, v in zip(dct["label_indices"], dct["label_values"])},
                   breaks=[int(b) for b in np.asarray(dct["breaks"]).reshape(-1)])

    def to_npz(self, filename):
        np.savez(filename, **self.as_dict())

    @classmethod
    def from_npz(cls, filename):
        with np.load(filename, allow_pickle=False) as f:
            return cls.from_dict({k: f[k] for k in f.files})

    # ------------------------------------------------------------------ use
    def get_refined(self, factor=2):
        """Each unbroken segment i -> i+1 split into `factor` equal steps (path.py:223-249);
        labels and breaks follow their points."""
        K = self.K_list
        last = len(K) - 1
        pts, labels, breaks = [], {}, []
        for i in range(last + 1):
            pts.append(K[i])
            if i in self.labels:
                labels[len(pts) - 1] = self.labels[i]
            if i in self.breaks:
                breaks.append(len(pts) - 1)
            if i < last and i not in self.breaks:
                step = (K[i + 1] - K[i]) / factor
                pts.extend(K[i] + j * step for j in range(1, factor))
        return Path(pointgroup=self.pointgroup, k_list=pts, labels=labels, breaks=breaks)

    def get_kpoints(self):
        return np.array(self.K_list)

    def get_kpoints_cart(self):
        return self.K_list @ self.recip_lattice

    @property
    def str_short(self):
        return f"Path() with {len(self.K_list)} points and labels {self.labels}"

    def __str__(self):
        rows = []
        for i, k in enumerate(self.K_list):
            s = "  ".join(f"{x:10.6f}" for x in k)
            if i in self.labels:
                s += " <--- " + str(self.labels[i])
            if i in self.breaks:
                s += "\n" + "-" * 20
            rows.append(s)
        return "\n" + "\n".join(rows)

    def get_K_list(self, use_symmetry=False, k_batch=None):
        """Blocks of `k_batch` consecutive points (path.py:269-284); symmetry is not used."""
        if use_symmetry:
            warnings.warn("symmetry is not used for a tabulation along path")
        n = len(self.K_list)
        k_batch = n if not k_batch else int(k_batch)
        return [KpointBZpath(K=self.K_list[i:i + k_batch], pointgroup=self.pointgroup) for i in range(0, n, k_batch)]

    def getKline(self, break_thresh=np.inf):
        """Distance along the path (Cartesian, 1/Angstrom): steps longer than break_thresh and
        steps after a break count zero (path.py:286-294)."""
        kc = self.get_kpoints_cart()
        step = np.linalg.norm(kc[1:] - kc[:-1], axis=1)
        step[step > break_thresh] = 0.0
        if self.breaks:
            step[self.breaks] = 0.0
        return np.concatenate(([0.0], np.cumsum(step)))
