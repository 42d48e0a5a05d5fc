"""Point-group symmetry of the reciprocal-space results.

Restates wannierberri/symmetry/point_symmetry.py:49-423 for what the path
needs: point operations (proper rotation R, inversion flag, time-reversal
flag) acting on reduced k-vectors and on result tensors, the group generated
by a list of operations, the star of a K-point and the symmetrisation of a
result (average of its images under the group).

Result tensors transform by rotating each of their last `rank` Cartesian
axes with R, times -1 for an operation with time reversal when the quantity
is TR-odd and times -1 for one with inversion when it is inversion-odd (the
factor-only transforms of formula/covariant.py and data_K.py:16-46).
"""
from collections.abc import Iterable

import numpy as np

SYMMETRY_PRECISION = 1e-6


def _rotation_matrix(n, axis):
    """Rotation by 2 pi / n about `axis` (Cartesian), Rodrigues' formula."""
    axis = np.array(axis, dtype=float)
    norm = np.linalg.norm(axis)
    if norm < 1e-10:
        raise ValueError(f"the axis vector is too small : {norm}")
    u = axis / norm
    th = 2 * np.pi / n
    K = np.array([[0, -u[2], u[1]], [u[2], 0, -u[0]], [-u[1], u[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


class PointSymmetry:
    """A point operation in Cartesian coordinates (point_symmetry.py:49-119).

    `R` may be improper (det -1): it is stored as the proper rotation with
    `Inv` set.  A reduced k-vector maps to iTR * iInv * (R k).
    """

    def __init__(self, R, TR=False):
        R = np.array(R, dtype=float)
        self.TR = bool(TR)
        self.Inv = bool(np.linalg.det(R) < 0)
        self.R = -R if self.Inv else R

    @property
    def iTR(self):
        return -1 if self.TR else 1

    @property
    def iInv(self):
        return -1 if self.Inv else 1

    def __mul__(self, other):
        return PointSymmetry((self.R @ other.R) * (self.iInv * other.iInv), self.TR != other.TR)

    def __eq__(self, other):
        return np.linalg.norm(self.R - other.R) < 1e-12 and self.TR == other.TR and self.Inv == other.Inv

    def __str__(self):
        return f"rotation:\n{np.round(self.R, decimals=4)} , TR: {self.TR} , I: {self.Inv}"

    def transform_reduced_vector(self, vec, basis):
        """Rows of `vec` in the basis `basis` (rows) mapped by the operation."""
        return vec @ (basis @ self.R.T @ np.linalg.inv(basis)) * (self.iTR * self.iInv)

    def transform_tensor(self, data, rank, factor_TR=1, factor_Inv=1):
        """Rotate the last `rank` axes of `data` (NumPy or torch) and apply the TR / inversion factors."""
        nd = len(data.shape)
        if rank > 0 and tuple(data.shape[nd - rank:]) != (3,) * rank:
            raise RuntimeError(f"all dimensions of rank-{rank} tensor should be 3, found: {data.shape[nd - rank:]}")
        factor = (factor_TR if self.TR else 1) * (factor_Inv if self.Inv else 1)
        if hasattr(data, "device") and not isinstance(data, np.ndarray):
            import torch
            R = torch.as_tensor(self.R, dtype=data.dtype, device=data.device)
            res = data
            for i in range(nd - rank, nd):
                res = torch.movedim(torch.movedim(res, i, -1) @ R.T, -1, i)
            return res * factor
        res = np.array(data)
        for i in range(nd - rank, nd):
            res = np.moveaxis(np.moveaxis(res, i, -1) @ self.R.T, -1, i)
        return res * factor


class Rotation(PointSymmetry):
    """n-fold rotation about `axis` (point_symmetry.py:122-142)."""

    def __init__(self, n, axis=(0, 0, 1)):
        if not isinstance(n, int) or n == 0:
            raise ValueError("Only non-zero integer rotations are supported")
        super().__init__(_rotation_matrix(n, axis))


class Mirror(PointSymmetry):
    """Mirror plane normal to `axis` (point_symmetry.py:145-156)."""

    def __init__(self, axis=(0, 0, 1)):
        super().__init__(-_rotation_matrix(2, axis))


Identity = PointSymmetry(np.eye(3))
Inversion = PointSymmetry(-np.eye(3))
TimeReversal = PointSymmetry(np.eye(3), True)
Mx = Mirror([1, 0, 0])
My = Mirror([0, 1, 0])
Mz = Mirror([0, 0, 1])
C2x = Rotation(2, [1, 0, 0])
C2y = Rotation(2, [0, 1, 0])
C2z = Rotation(2, [0, 0, 1])
C3z = Rotation(3, [0, 0, 1])
C4x = Rotation(4, [1, 0, 0])
C4y = Rotation(4, [0, 1, 0])
C4z = Rotation(4, [0, 0, 1])
C6z = Rotation(6, [0, 0, 1])

dict_sym = {"Identity": Identity, "Inversion": Inversion, "TimeReversal": TimeReversal, "Mx": Mx, "My": My,
            "Mz": Mz, "C2x": C2x, "C2y": C2y, "C2z": C2z, "C3z": C3z, "C4x": C4x, "C4y": C4y, "C4z": C4z,
            "C6z": C6z}


def product(lst):
    """lst[0] * lst[1] * ... (point_symmetry.py:195-201)."""
    assert isinstance(lst, Iterable) and len(lst) > 0
    res = Identity
    for op in lst[::-1]:
        res = op * res
    return res


def from_string(string):
    try:
        return dict_sym[string]
    except KeyError:
        raise ValueError(f"The symmetry {string} is not defined. Use Rotation(n, axis) or Mirror(axis)")


def from_string_prod(string):
    """'TimeReversal*C2x' -> the product (point_symmetry.py:214-218)."""
    return product([from_string(s) for s in string.split("*")])


class PointGroup:
    """The group generated by `generator_list` (point_symmetry.py:221-423).

    Generators are PointSymmetry objects or names ('C3z', 'TimeReversal*C2x').
    """

    def __init__(self, generator_list=(), real_lattice=None, recip_lattice=None):
        if real_lattice is None and recip_lattice is None:
            real_lattice = np.eye(3)
        if real_lattice is not None:
            self.real_lattice = np.array(real_lattice, dtype=float)
            self.recip_lattice = 2 * np.pi * np.linalg.inv(self.real_lattice).T
        else:
            self.recip_lattice = np.array(recip_lattice, dtype=float)
            self.real_lattice = 2 * np.pi * np.linalg.inv(self.recip_lattice).T
        ops = [op if isinstance(op, PointSymmetry) else from_string_prod(op) for op in generator_list]
        if not ops:
            ops = [Identity]
        # closure under products
        while True:
            n_old = len(ops)
            for s1 in list(ops):
                for s2 in list(ops):
                    s3 = s1 * s2
                    if s3 not in ops:
                        ops.append(s3)
                        if len(ops) > 1000:
                            raise RuntimeError("Cannot define a finite group")
            if len(ops) == n_old:
                break
        self.symmetries = ops
        msg = " : the symmetries are not consistent with the lattice vectors"
        assert self.check_basis_symmetry(self.real_lattice), "real_lattice is not symmetric" + msg
        assert self.check_basis_symmetry(self.recip_lattice), "recip_lattice is not symmetric" + msg

    @property
    def size(self):
        return len(self.symmetries)

    def as_dict(self):
        """Keys real_lattice, nsym, symm{i}_R (improper matrices kept improper), symm{i}_TR
        (point_symmetry.py:294-302): the `pointgroup.npz` of a saved system."""
        d = dict(real_lattice=self.real_lattice, nsym=self.size)
        for i, sym in enumerate(self.symmetries):
            d[f"symm{i}_R"] = sym.R * sym.iInv
            d[f"symm{i}_TR"] = sym.TR
        return d

    @classmethod
    def from_dict(cls, d):
        ops = [PointSymmetry(d[f"symm{i}_R"], bool(d[f"symm{i}_TR"])) for i in range(int(d["nsym"]))]
        return cls(ops, real_lattice=d["real_lattice"])

    def check_basis_symmetry(self, basis, tol=1e-6, rel_tol=None):
        """True if every operation maps the basis vectors onto integer combinations of them."""
        if rel_tol is not None:
            tol = rel_tol * tol
        for sym in self.symmetries:
            rot = sym.transform_reduced_vector(np.eye(3), basis)
            if np.abs(np.round(rot) - rot).max() > tol:
                return False
        return True

    def symmetric_grid(self, nk):
        return self.check_basis_symmetry(self.recip_lattice / np.array(nk)[:, None], rel_tol=10)

    def symmetrize(self, result):
        """Average of the result's images under the group (point_symmetry.py:337-338)."""
        if self.size == 1:
            return result
        return sum(result.transform(s) for s in self.symmetries) / self.size

    def star(self, k):
        """Distinct images of the reduced vector k, modulo reciprocal lattice vectors (point_symmetry.py:413-419)."""
        st = [S.transform_reduced_vector(np.asarray(k, dtype=float), self.recip_lattice) for S in self.symmetries]
        for i in range(len(st) - 1, 0, -1):
            diff = np.array(st[:i]) - np.array(st[i])[None, :]
            if np.linalg.norm(diff - diff.round(), axis=-1).min() < SYMMETRY_PRECISION:
                del st[i]
        return np.array(st)


__all__ = ["PointSymmetry", "Rotation", "Mirror", "PointGroup", "Identity", "Inversion", "TimeReversal", "Mx",
           "My", "Mz", "C2x", "C2y", "C2z", "C3z", "C4x", "C4y", "C4z", "C6z", "dict_sym", "product",
           "from_string", "from_string_prod", "SYMMETRY_PRECISION"]
