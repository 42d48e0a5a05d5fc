"""Data_K: one K-block of the Berry path on the MI355X (the `data_k_class` of run()).

Constructed as the reference constructs its data_K per K-block
(run_grid.py:258-265, data_K/data_K_R.py:8-22):

    Data_K(system, dK=Kp.Kp_fullBZ, grid=grid, Kpoint=Kp, **parameters_K)

and consumed by the calculators of `wannierberri_amd.calculators`.  All
numerics run in libwbmi (HIP, gfx950):

  R-space channel stack (once per system)  -> wbmi_build_channels
  R -> k on the NKFFT sub-grid              -> wbmi_r2k_exec  (k-list: wbmi_r2k_klist)
  eigh of 0.5 (H + H^+)                     -> wbmi_eigh
  U^+ X U of every other channel            -> wbmi_rotate
  formula traces + Fermi scan               -> wbmi_static_scan (called by calculators)

A "channel" is one complex nw x nw matrix per k: a component of
Xbar(name, der) for name in Ham, AA, rotAA, BB, CC and der in 0, 1 (2 for Ham),
or of the spin matrices Xbar('SS'), Xbar('SA'), Xbar('SHA'), Xbar('SH'),
Xbar('SR'), Xbar('SHR') (der 0).
"""
import collections

import numpy as np

from . import _lib
from ._lib import WBMI_F32, WBMI_F32_SPLIT, WBMI_F64
from .path import KpointBZpath

# components of each Xbar(name, der)
CHANNEL_NCOMP = {("Ham", 0): 1, ("Ham", 1): 3, ("Ham", 2): 9, ("Ham", 3): 27,
                 ("AA", 0): 3, ("AA", 1): 9, ("rotAA", 0): 3, ("rotAA", 1): 9,
                 ("BB", 0): 3, ("BB", 1): 9, ("CC", 0): 3, ("CC", 1): 9,
                 ("SS", 0): 3, ("SS", 1): 9, ("SA", 0): 9, ("SHA", 0): 9, ("SH", 0): 3, ("SR", 0): 9, ("SHR", 0): 9}
CHANNEL_ORDER = list(CHANNEL_NCOMP)
# matrices forced Hermitian after R->k (data_K/data_K_R.py:87)
HERMITIAN = {"AA", "rotAA", "SS"}
BASES = ["Ham", "AA", "BB", "CC", "SS", "SA", "SHA", "SH", "SR", "SHR"]
# At most this many FP64 Ogita-Aishima steps after the FP32 Jacobi (the
# kernel stops at ||E||_F < 1e-8 (complex64 vectors) or 1e-10 (complex128):
# two or three steps for separated spectra, up to about ten for a pair
# closer than ~1e-5 |H|, whose first-order correction is large)
OA_STEPS = 16
_ALPHA = [1, 2, 0]
_BETA = [2, 0, 1]


def channel_terms(name, der, base_offset):
    """Per component: two terms (sign, base component, d1, d2) of X_R * prod(i r_d).

    Ham/AA/BB/CC derivatives: rvectors.py:487-506 (one Cartesian axis per order);
    rotAA_c = i r_alpha A_beta - i r_beta A_alpha: data_K_R.py:48-51.
    """
    out = []
    if name == "rotAA":
        a0 = base_offset["AA"]
        for c in range(3):
            al, be = _ALPHA[c], _BETA[c]
            for d in ([-1] if der == 0 else [0, 1, 2]):
                out.append([[1, a0 + be, al, d], [-1, a0 + al, be, d]])
        return out
    b0 = base_offset[name]
    if name == "Ham":
        if der == 0:
            return [[[1, b0, -1, -1], [0, 0, -1, -1]]]
        if der == 1:
            return [[[1, b0, a, -1], [0, 0, -1, -1]] for a in range(3)]
        if der == 3:
            # the first factor i r_a is in the base (Ham_d = i r_a Ham, DeviceSystem.stack)
            h0 = base_offset["Ham_d"]
            return [[[1, h0 + a, b, c], [0, 0, -1, -1]] for a in range(3) for b in range(3) for c in range(3)]
        return [[[1, b0, a, b], [0, 0, -1, -1]] for a in range(3) for b in range(3)]
    if der == 0:
        # SA and SHA carry two Cartesian indices (a, s), row-major
        return [[[1, b0 + a, -1, -1], [0, 0, -1, -1]] for a in range(CHANNEL_NCOMP[(name, 0)])]
    return [[[1, b0 + a, d, -1], [0, 0, -1, -1]] for a in range(3) for d in range(3)]


class DeviceSystem:
    """Device copies of a System_R's R-space data (cached per system and device)."""

    # bytes of z-pass outputs kept resident for the K-blocks of grids (zpass_cached); 0 = none
    zcache_limit = 16 << 30

    def __init__(self, system, ctx):
        import torch
        self.torch = torch
        self.ctx = ctx
        dev = ctx.device
        self.system = system
        self.nw = system.num_wann
        self.nR = system.nRvec
        present = [b for b in BASES if system.has_R_mat(b)]
        self.base_offset = {}
        comps = []
        for b in present:
            X = system.get_R_mat(b)
            X = X.reshape(X.shape[:3] + (-1,))
            self.base_offset[b] = sum(c.shape[3] for c in comps)
            comps.append(X)
        base = np.concatenate(comps, axis=3)
        self.base = torch.from_numpy(np.ascontiguousarray(base)).to(dev)
        self.cR = torch.from_numpy(np.ascontiguousarray(system.cRvec_shifted)).to(dev)
        self.iRvec = np.ascontiguousarray(system.rvec.iRvec, dtype=np.int32)
        self.iRvec_dev = torch.from_numpy(self.iRvec).to(dev)
        self._stacks = {}
        self._plans = {}
        self._zcache = collections.OrderedDict()   # z-pass outputs, least recently used first
        self._zcache_used = 0
        self.eig_hint = None   # the last grid K-block's eigenvectors (starting basis of the next)

    def has(self, name):
        return name in self.base_offset or (name == "rotAA" and "AA" in self.base_offset)

    def stack(self, keys):
        """R-space channel stack [nR, nch, nw, nw] (complex128) for the channel keys."""
        keys = tuple(keys)
        if ("Ham", 3) in keys and "Ham_d" not in self.base_offset:
            # third derivative: one factor i r_a applied here, two by the channel builder
            h = self.base[..., self.base_offset["Ham"]:self.base_offset["Ham"] + 1]
            hamd = h * (1j * self.cR)
            self.base_offset["Ham_d"] = self.base.shape[3]
            self.base = self.torch.cat([self.base, hamd], dim=3).contiguous()
        if keys not in self._stacks:
            spec = []
            for name, der in keys:
                if not self.has(name):
                    raise ValueError(f"The real-space matrix elements '{name}' are not set in the system, "
                                     "but are required for the current calculation")
                spec += channel_terms(name, der, self.base_offset)
            spec = np.array(spec, dtype=np.int32)
            out = self.torch.empty((self.nR, len(spec), self.nw, self.nw), dtype=self.torch.complex128,
                                   device=self.ctx.device)
            self.ctx.build_channels(self.cR, self.base, spec, out)
            self._stacks[keys] = out
        return self._stacks[keys]

    def plan(self, keys, NKFFT):
        keys = tuple(keys)
        key = (keys, tuple(int(x) for x in NKFFT))
        if key not in self._plans:
            st = self.stack(keys)
            nelem = st[0].numel()
            plan = self.ctx.r2k_plan(self.iRvec, NKFFT, nelem)
            aliased = _is_aliased(self.iRvec, NKFFT)
            box = None   # aliased: rebuilt per call into a box of the caller's
            if not aliased:
                box = self.torch.empty(plan.box_bytes // 16, dtype=self.torch.complex128, device=self.ctx.device)
                plan.box(st, box)
            self._plans[key] = (plan, box, aliased)
        return self._plans[key]


    def zcache_clear(self):
        self._zcache.clear()
        self._zcache_used = 0

    def zpass_cached(self, keys, NKFFT, plan, box, dK, prec, work):
        """The z-pass output of the grid transform of `keys` for shifts with this dK[2]
        (wbmi_r2k_zpass): the K-blocks of a grid share a few distinct dK[2], so it is kept
        resident (up to `zcache_limit` bytes, least recently used dropped first) and each
        K-block runs only the y and x passes.  None when it does not fit (or under graph
        capture): the caller then runs the whole transform.

        Each use is ordered after the stream that wrote the buffer (an event) and recorded on
        the using stream, so a dropped buffer is reused only after every stream that read it."""
        torch = self.torch
        if self.zcache_limit <= 0 or torch.cuda.is_current_stream_capturing():
            return None
        key = (tuple(keys), tuple(int(x) for x in NKFFT), int(prec), float(dK[2]))
        cur = self.ctx.stream
        ent = self._zcache.get(key)
        if ent is not None:
            self._zcache.move_to_end(key)
            w1, ev = ent
            cur.wait_event(ev)
            w1.record_stream(cur)
            return w1
        nb = plan.w1_bytes(prec)
        if nb > self.zcache_limit:
            return None
        while self._zcache_used + nb > self.zcache_limit:
            _, (old, _) = self._zcache.popitem(last=False)
            self._zcache_used -= old.numel()
        w1 = torch.empty(nb, dtype=torch.uint8, device=self.ctx.device)
        w1.record_stream(cur)
        plan.zpass(box, dK, work, w1, prec)
        ev = torch.cuda.Event()
        ev.record(cur)
        self._zcache[key] = (w1, ev)
        self._zcache_used += nb
        return w1



def _is_aliased(iRvec, NKFFT):
    ext = iRvec.max(axis=0) - iRvec.min(axis=0) + 1
    return bool(np.any(ext > np.array(NKFFT)))


def device_system(system, ctx):
    # one copy per device: the context launches on whatever stream it holds at the time
    # (a graph capture switches it to the capture stream)
    key = ctx.device.index
    cache = system._device_cache
    if key not in cache:
        cache[key] = DeviceSystem(system, ctx)
    return cache[key]


class Data_K:
    """K-block data on the GPU (data_K/data_K.py:49-334, data_K/data_K_R.py:8-97).

    Parameters (beyond the reference's): `precision` = 'fp64' (default) or
    'fp32' (complex64 k-space channels; energies, denominators and sums stay
    FP64); `channels` = iterable of (name, der) computed together on first use;
    `split_rotation` (mixed form only, default True): rotate every channel but
    the velocity dH/dk with split-fp16 matrix-core products (FP32-accumulation
    accuracy) instead of FP64 arithmetic.
    """

    accepts_channels = True

    def __init__(self, system, dK=None, grid=None, Kpoint=None, k_list=None, Emin=-np.inf, Emax=np.inf,
                 fftlib=None, precision="fp64", channels=None, device=0, random_gauge=False, split_rotation=True,
                 k_list_dev=None, **unused):
        import torch
        self.torch = torch
        if not torch.cuda.is_available():
            raise _lib.WbmiLibraryError("the MI355X path needs a GPU (torch.cuda is not available)")
        if np.isfinite(Emin) or np.isfinite(Emax):
            raise NotImplementedError("band selection by Emin/Emax is not supported")
        if random_gauge:
            raise NotImplementedError("random_gauge is a test-only option of the reference and is not supported")
        if k_list is None and isinstance(Kpoint, KpointBZpath):
            k_list, dK = Kpoint.K, None   # a batch of path points is a k-list (data_K.py:84-87)
        if k_list is None and dK is None:
            raise AssertionError("either dK or k_list must be specified, but not both")
        self.system = system
        self.ctx = _lib.get_context(device)
        self.dev = device_system(system, self.ctx)
        self.num_wann = system.num_wann
        self.nbands = self.num_wann
        self.cell_volume = system.cell_volume
        self.force_internal_terms_only = system.force_internal_terms_only
        self.grid = grid
        self.Kpoint = Kpoint
        self.k_list = None if k_list is None else np.array(k_list, dtype=float).reshape(-1, 3)
        # the k-list already on the device (float64 [nk, 3]): no host-to-device copy per evaluation
        self.k_list_dev = k_list_dev
        if self.k_list is None:
            self.dK = np.array(dK, dtype=float)
            self.NKFFT = np.array(grid.FFT, dtype=int)
            self.nk = int(np.prod(self.NKFFT))
        else:
            self.dK = None
            self.NKFFT = None
            self.nk = len(self.k_list)
        if precision not in ("fp64", "fp32"):
            raise ValueError(f"precision must be 'fp64' or 'fp32', not {precision}")
        self.precision = precision
        self.prec = WBMI_F64 if precision == "fp64" else WBMI_F32
        self.cdtype = torch.complex128 if precision == "fp64" else torch.complex64
        self.split_rotation = bool(split_rotation) and precision == "fp32"
        self._requested = set(channels) if channels else set()
        self._slots = None
        self.Xk = None
        self.E_dev = None
        self.U_dev = None
        self._Ec = None
        self._pref = None

    # -- evaluation -------------------------------------------------------------
    def request(self, keys):
        self._requested |= set(keys)

    def ensure(self, keys):
        """Make sure the channels `keys` are available (rotated to the band basis)."""
        keys = set(keys) | {("Ham", 0)}
        if self._slots is not None and (keys - {("Ham", 0)}) <= set(self._slots):
            return
        self._requested |= keys
        if self._slots is not None:
            self._requested |= set(self._slots)
        self._evaluate(sorted(self._requested, key=CHANNEL_ORDER.index))

    def _transform(self, keys, dtype, prec, shift=None, keep=None):
        """R -> k of the channel keys into a new [nk, nch, nw, nw] tensor (grid blocks:
        at the points j/NKFFT + dK, or + shift in place of dK).  `keep`: a list that takes the
        work buffer instead of freeing it (work launched on another stream)."""
        torch = self.torch
        nw, nk = self.num_wann, self.nk
        stack = self.dev.stack(keys)
        X = torch.empty((nk, stack.shape[1], nw, nw), dtype=dtype, device=self.ctx.device)
        if self.k_list is None:
            plan, box, aliased = self.dev.plan(keys, self.NKFFT)
            dK = self.dK if shift is None else shift
            work = torch.empty(plan.work_bytes, dtype=torch.uint8, device=self.ctx.device)
            w1 = None if aliased else self.dev.zpass_cached(keys, self.NKFFT, plan, box, dK, prec, work)
            if w1 is not None:
                plan.yx(w1, dK, work, X, prec)
            else:
                if aliased:
                    box = torch.empty(plan.box_bytes // 16, dtype=torch.complex128, device=self.ctx.device)
                plan.execute(stack if aliased else None, box, dK, work, X, prec)
            if keep is not None:
                keep.extend([work, box])
            del work
        else:
            kl = self.k_list_dev if self.k_list_dev is not None else torch.from_numpy(self.k_list).to(self.ctx.device)
            self.ctx.r2k_klist(self.dev.iRvec_dev, stack, kl, X, prec)
        return X

    def _channel_keys(self):
        return sorted(self._requested | {("Ham", 0)}, key=CHANNEL_ORDER.index)

    def _transforms(self, keys, keep=None):
        """R -> k of `keys`: (the stack in the form's type, the FP64 Hamiltonian of the mixed form)."""
        torch = self.torch
        mixed = self.prec == WBMI_F32
        main = [k for k in keys if not (mixed and k == ("Ham", 0))]
        X = self._transform(main, self.cdtype, self.prec, keep=keep) if main else None
        H64 = self._transform([("Ham", 0)], torch.complex128, WBMI_F64, keep=keep) if mixed else None
        return X, H64

    def prefetch(self, stream):
        """Start this grid K-block's R -> k transforms of the requested channels on `stream`
        (a side stream), so that they run while the previous K-block is still being evaluated
        on the context's stream; the evaluation waits for them with an event.

        Every buffer comes from the context stream's memory pool: `stream` first waits for the
        work queued on the context's stream so far (so memory freed there is no longer in use),
        and the work buffers stay referenced until the context's stream has waited for the
        event, after which they are freed in its order."""
        torch = self.torch
        if self.k_list is not None or self._slots is not None or self._pref is not None:
            return
        keys = self._channel_keys()
        main = self.ctx.stream
        stream.wait_stream(main)
        keep = []
        prev = self.ctx.set_stream(stream)
        try:
            X, H64 = self._transforms(keys, keep=keep)
            ev = torch.cuda.Event()
            ev.record(stream)
        finally:
            self.ctx.set_stream(prev)
        self._pref = (keys, X, H64, ev, keep)

    def __del__(self):
        # a prefetched block dropped unevaluated: its buffers return to the context stream's
        # pool only after the side-stream work that writes them
        pref = getattr(self, "_pref", None)
        if pref is not None:
            try:
                self.ctx.stream.wait_event(pref[3])
            except Exception:
                pass

    def _take_prefetch(self, keys):
        """The prefetched transforms when they hold exactly `keys`, ordered before the context
        stream's next work; else (None, None)."""
        pref, self._pref = self._pref, None
        if pref is None:
            return None, None
        self.ctx.stream.wait_event(pref[3])
        if pref[0] != list(keys):
            return None, None
        return pref[1], pref[2]

    def _evaluate(self, keys):
        """Channels `keys` (always including the Hamiltonian) -> E, U and the rotated stack.

        FP64 form: one complex128 stack; FP32 Jacobi + FP64 Ogita-Aishima
        refinement to FP64-accurate energies and complex128 vectors.  Mixed
        form: the
        Hamiltonian is transformed in FP64 on its own, the other channels in
        complex64; FP32 Jacobi + FP64 Ogita-Aishima steps (at most OA_STEPS) give FP64
        energies and refined eigenvectors (stored as complex64).
        """
        torch = self.torch
        nw, nk = self.num_wann, self.nk
        mixed = self.prec == WBMI_F32
        main = [k for k in keys if not (mixed and k == ("Ham", 0))]
        slots, s = {}, 0
        for key in main:
            slots[key] = s
            s += CHANNEL_NCOMP[key]
        nch = s
        E = torch.empty((nk, nw), dtype=torch.float64, device=self.ctx.device)
        U = torch.empty((nk, nw, nw), dtype=self.cdtype, device=self.ctx.device)
        X, H64 = self._take_prefetch(keys)
        if X is None and H64 is None:
            X, H64 = self._transforms(keys)
        # starting eigenvectors: the previous K-block's at the same k index (one grid step
        # away on the NK grid), when it was a grid block of the same shape and form
        U0 = None
        hint = getattr(self.dev, "eig_hint", None)
        if self.k_list is None and hint is not None and hint.shape == U.shape and hint.dtype == U.dtype:
            U0 = hint
        if mixed:
            self.ctx.eigh(None, nw * nw, E, U, self.prec, nw, refine=OA_STEPS, H64=H64, U0=U0)
            del H64
        else:
            # FP32 Jacobi + FP64 refinement to complex128 vectors
            h0 = slots[("Ham", 0)]
            self.ctx.eigh(X[:, h0], nch * nw * nw, E, U, self.prec, nw, refine=OA_STEPS, U0=U0)
        if self.k_list is None:
            self.dev.eig_hint = U
        if X is not None:
            herm = np.zeros(nch, dtype=np.int32)
            skip = np.zeros(nch, dtype=np.int32)
            for key, sl in slots.items():
                if key[0] in HERMITIAN:
                    herm[sl:sl + CHANNEL_NCOMP[key]] = 1
            if ("Ham", 0) in slots:
                skip[slots[("Ham", 0)]] = 1
            if self.split_rotation:
                # the velocity in FP64 arithmetic (its elements between close bands enter the
                # Berry curvature through 1 / (E_n - E_m)^2), the rest on split-fp16 products
                vel = np.zeros(nch, dtype=np.int32)
                if ("Ham", 1) in slots:
                    vel[slots[("Ham", 1)]:slots[("Ham", 1)] + 3] = 1
                if (skip | (1 - vel)).sum() < nch:
                    self.ctx.rotate(U, X, herm, skip | (1 - vel), self.prec)
                if (skip | vel).sum() < nch:
                    self.ctx.rotate(U, X, herm, skip | vel, WBMI_F32_SPLIT)
            elif skip.sum() < nch:
                self.ctx.rotate(U, X, herm, skip, self.prec)
        self.Xk, self.E_dev, self.U_dev, self._slots, self.nch = X, E, U, slots, nch

    def corner_energies(self):
        """Band energies [8, nk, nw] (FP64, device) at the corners
        k + ((2i-1), (2j-1), (2l-1)) * dK_fullBZ / 2 of the parallelepiped around each k,
        corner i*4 + j*2 + l (data_K_R.py:126-141), or [4, nk, nw] at the vertices of a
        tetrahedral K-block (99-119): the Hamiltonian transformed to the shifted grid and
        diagonalised (FP32 Jacobi + FP64 refinement, FP64-accurate energies) once per corner."""
        if self._Ec is None:
            if self.k_list is not None or self.Kpoint is None:
                raise ValueError("the tetrahedron method needs a grid K-block (Kpoint with dK)")
            self.ensure(set())
            torch = self.torch
            nw, nk = self.num_wann, self.nk
            if hasattr(self.Kpoint, "vertices_fullBZ"):
                # tetrahedral K-block: the 4 vertices (data_K_R.py:99-119)
                shifts = [self.dK + v for v in np.asarray(self.Kpoint.vertices_fullBZ, dtype=float)]
            else:
                half = np.asarray(self.Kpoint.dK_fullBZ, dtype=float) / 2
                shifts = [self.dK + (np.array([(c >> 2) & 1, (c >> 1) & 1, c & 1]) * 2 - 1) * half for c in range(8)]
            Ec = torch.empty((len(shifts), nk, nw), dtype=torch.float64, device=self.ctx.device)
            U = torch.empty((nk, nw, nw), dtype=torch.complex128, device=self.ctx.device)
            for c, shift in enumerate(shifts):
                H = self._transform([("Ham", 0)], torch.complex128, WBMI_F64, shift=shift)
                self.ctx.eigh(H[:, 0], nw * nw, Ec[c], U, WBMI_F64, nw, refine=OA_STEPS)
                del H
            self._Ec = Ec
        return self._Ec

    def tabulate(self, params, out):
        """Per-band table [nk, nw, ncomp] (see wbmi_tabulate)."""
        for i, v in enumerate(self.role_table()):
            params.ch[i] = v
        self.ctx.tabulate(self.E_dev, self.Xk, self.prec, params, out)

    def _scan_input(self, params, bands):
        """(X, prec, nch) of a scan: the channel stack, or per-band values [nk, nw, nvals]
        (float64) for WBMI_FORMULA_BANDVALUES."""
        for i, v in enumerate(self.role_table()):
            params.ch[i] = v
        if bands is None:
            return self.Xk, self.prec, None
        return bands, WBMI_F64, int(bands.shape[2])

    def static_scan_tetra_k(self, params, Efermi, hist, weight=1.0, bands=None):
        """k-resolved tetrahedron-weighted scan into hist [nk, 2 nEF, ncomp]."""
        Ec = self.corner_energies()
        X, prec, nch = self._scan_input(params, bands)
        self.ctx.static_scan_tetra_k(self.E_dev, Ec, X, prec, params, Efermi, weight, hist, nch=nch)

    def static_scan_tetra(self, params, Efermi, hist, weight=1.0, bands=None):
        """Tetrahedron-weighted scan into hist [2 nEF, ncomp] (see wbmi_static_scan_tetra)."""
        Ec = self.corner_energies()
        X, prec, nch = self._scan_input(params, bands)
        self.ctx.static_scan_tetra(self.E_dev, Ec, X, prec, params, Efermi, weight, hist, nch=nch)

    def role_table(self):
        """Channel slot per formula role (include/wbmi.h WBMI_ROLE_*), -1 if absent."""
        ch = [-1] * _lib.NROLES
        s = self._slots or {}

        def put(role, key, n):
            if key in s:
                for i in range(n):
                    ch[role + i] = s[key] + i
        put(_lib.ROLE_V, ("Ham", 1), 3)
        put(_lib.ROLE_W, ("Ham", 2), 9)
        put(_lib.ROLE_A, ("AA", 0), 3)
        put(_lib.ROLE_AD, ("AA", 1), 9)
        put(_lib.ROLE_O, ("rotAA", 0), 3)
        put(_lib.ROLE_OD, ("rotAA", 1), 9)
        put(_lib.ROLE_B, ("BB", 0), 3)
        put(_lib.ROLE_BD, ("BB", 1), 9)
        put(_lib.ROLE_C, ("CC", 0), 3)
        put(_lib.ROLE_CD, ("CC", 1), 9)
        put(_lib.ROLE_S, ("SS", 0), 3)
        put(_lib.ROLE_SD, ("SS", 1), 9)
        return ch

    def static_scan(self, params, hist, weight=1.0, bands=None):
        """Accumulate weight * (group traces) into hist [nEF_extra, ncomp] (FP64, device)."""
        X, prec, nch = self._scan_input(params, bands)
        self.ctx.static_scan(self.E_dev, X, prec, params, weight, hist, nch=nch)

    def static_scan_k(self, params, hist, weight=1.0, bands=None):
        """k-resolved scan into hist [nk, nEF_extra, ncomp] (FP64, device)."""
        X, prec, nch = self._scan_input(params, bands)
        self.ctx.static_scan_k(self.E_dev, X, prec, params, weight, hist, nch=nch)

    def subset(self, k0, k1):
        """The k-points [k0, k1) of an evaluated k-list block as a Data_K of their own.

        Shares (slices) the device arrays of this block, so calculators on the
        subset cost one scan launch each; channels that this block did not
        evaluate are computed for the subset's points alone.
        """
        if self.k_list is None:
            raise ValueError("subset() needs a k-list block")
        if not 0 <= k0 < k1 <= self.nk:
            raise ValueError(f"subset [{k0}, {k1}) out of range for {self.nk} k-points")
        self.ensure(set())
        sub = object.__new__(type(self))
        sub.__dict__.update(self.__dict__)
        sub.k_list = self.k_list[k0:k1]
        sub.nk = k1 - k0
        sub._requested = set(self._requested)
        sub.E_dev = self.E_dev[k0:k1]
        sub.U_dev = self.U_dev[k0:k1]
        sub.Xk = None if self.Xk is None else self.Xk[k0:k1]
        return sub

    # -- reference-style accessors (synchronising; for tests and users) ---------
    @property
    def E_K(self):
        self.ensure(set())
        return self.E_dev.cpu().numpy()

    @property
    def UU_K(self):
        self.ensure(set())
        return self.U_dev.cpu().numpy()

    def Xbar(self, name, der=0):
        """Rotated matrix [nk, nw, nw, 3^der...] as NumPy (data_K_R.py:69-89)."""
        key = (name, der)
        self.ensure({key})
        if key == ("Ham", 0):
            # U^+ H U = diag(E): the Hamiltonian channel itself is consumed by eigh
            E = self.E_K
            return E[:, :, None] * np.eye(self.num_wann)[None]
        sl = self._slots[key]
        n = CHANNEL_NCOMP[key]
        x = self.Xk[:, sl:sl + n].cpu().numpy()
        x = np.moveaxis(x, 1, -1)
        if n == 9:
            x = x.reshape(x.shape[:3] + (3, 3))
        elif n == 1:
            x = x[..., 0]
        return x

    @property
    def kpoints_all(self):
        if self.k_list is not None:
            return self.k_list
        return (self.grid.points_FFT + self.dK[None]) % 1
