"""ctypes binding of libwbmi.so (the C ABI declared in include/wbmi.h).

The library is required: importing the GPU path without it raises
`WbmiLibraryError`; there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# WBMI_LIB: another build of the library (e.g. one compiled with different tuning macros)
LIB_PATH = os.environ.get("WBMI_LIB") or os.path.join(HERE, "libwbmi.so")

WBMI_OK = 0
WBMI_ERR_ARG = 1
WBMI_ERR_HIP = 2
WBMI_ERR_UNSUPPORTED = 3
WBMI_F64 = 0
WBMI_F32 = 1
WBMI_F32_SPLIT = 2

FORMULA_IDENTITY = 0
FORMULA_OMEGA = 1
FORMULA_DEROMEGA = 2
FORMULA_VELOMEGA = 3
FORMULA_MORB_HPM = 4
FORMULA_DERMORB = 5
FORMULA_VELHPLUS = 6
FORMULA_DERSPIN = 7
FORMULA_VELSPIN = 8
FORMULA_SPIN = 9
FORMULA_INVMASS = 10
FORMULA_VELVEL = 11
FORMULA_VELVELVEL = 12
FORMULA_MASSVEL = 13
FORMULA_HALL_SURF = 14
FORMULA_HALL_SEA = 15
FORMULA_BANDVALUES = 16
FORMULA_OMEGASPIN = 17
FORMULA_OMEGAOMEGA = 18
FORMULA_OMEGAHPLUS = 19
WBMI_SHC_RYOO, WBMI_SHC_QIAO, WBMI_SHC_SIMPLE = 0, 1, 2

ROLE_V, ROLE_W, ROLE_A, ROLE_AD, ROLE_O, ROLE_OD = 0, 3, 12, 15, 24, 27
ROLE_B, ROLE_BD, ROLE_C, ROLE_CD = 36, 39, 48, 51
ROLE_S, ROLE_SD = 60, 63
NROLES = 72

# exported symbols (must match include/wbmi.h)
SYMBOLS = [
    "wbmi_version", "wbmi_ctx_create", "wbmi_ctx_destroy", "wbmi_ctx_set_stream", "wbmi_sync",
    "wbmi_last_error", "wbmi_build_channels", "wbmi_r2k_plan_create", "wbmi_r2k_plan_destroy",
    "wbmi_r2k_plan_sizes", "wbmi_r2k_box", "wbmi_r2k_exec", "wbmi_r2k_klist", "wbmi_eigh",
    "wbmi_eigh_work_bytes", "wbmi_eigh_guess", "wbmi_optical_conductivity", "wbmi_rotate", "wbmi_rotate_work_bytes", "wbmi_static_scan", "wbmi_finalize",
    "wbmi_static_scan_tetra", "wbmi_finalize_tetra", "wbmi_jdos", "wbmi_spin_hall", "wbmi_spin_hall_work_bytes",
    "wbmi_tabulate", "wbmi_tabulate_diag", "wbmi_shift_current", "wbmi_injection_current",
    "wbmi_static_scan_k", "wbmi_finalize_k", "wbmi_static_scan_tetra_k", "wbmi_finalize_tetra_k",
    "wbmi_tabulate_invmass", "wbmi_eigh_noalloc", "wbmi_eigh_noalloc_max_bytes", "wbmi_r2k_klist_tile",
    "wbmi_r2k_klist_tile_shape", "wbmi_tabulate_spin_berry", "wbmi_spin_berry_bands", "wbmi_der3e_bands",
    "wbmi_optical_conductivity_work_bytes", "wbmi_optical_conductivity_ws", "wbmi_r2k_w1_bytes", "wbmi_r2k_zpass",
    "wbmi_r2k_yx",
]


class WbmiLibraryError(RuntimeError):
    pass


class WbmiError(RuntimeError):
    pass


class ScanParams(ctypes.Structure):
    _fields_ = [
        ("formula", ctypes.c_int), ("sign", ctypes.c_int), ("internal_terms", ctypes.c_int),
        ("external_terms", ctypes.c_int), ("fder", ctypes.c_int), ("nEF_extra", ctypes.c_int),
        ("EFmin", ctypes.c_double), ("EFmax", ctypes.c_double), ("dEF", ctypes.c_double),
        ("degen_thresh", ctypes.c_double), ("degen_Kramers", ctypes.c_int),
        ("ch", ctypes.c_int * NROLES), ("comp_slice", ctypes.c_int), ("band_weight", ctypes.c_void_p),
        ("erange", ctypes.c_int), ("Emin", ctypes.c_double), ("Emax", ctypes.c_double),
    ]


class OptcondParams(ctypes.Structure):
    _fields_ = [
        ("nEF", ctypes.c_int), ("nomega", ctypes.c_int), ("Efermi", ctypes.c_void_p), ("omega", ctypes.c_void_p),
        ("EFmin", ctypes.c_double), ("EFmax", ctypes.c_double), ("omega_min", ctypes.c_double),
        ("omega_max", ctypes.c_double), ("kBT", ctypes.c_double), ("eta", ctypes.c_double),
        ("smr_gaussian", ctypes.c_int), ("degen_thresh", ctypes.c_double), ("degen_Kramers", ctypes.c_int),
        ("external_terms", ctypes.c_int), ("ch", ctypes.c_int * NROLES),
    ]


class ShcParams(ctypes.Structure):
    _fields_ = [("base", OptcondParams), ("kind", ctypes.c_int), ("ch_S", ctypes.c_int * 3),
                ("ch_SA", ctypes.c_int * 9), ("ch_SHA", ctypes.c_int * 9), ("ch_SH", ctypes.c_int * 3),
                ("ch_SR", ctypes.c_int * 9), ("ch_SHR", ctypes.c_int * 9)]


_lib = None
_lock = threading.Lock()

_vp, _i, _i64, _d = ctypes.c_void_p, ctypes.c_int, ctypes.c_int64, ctypes.c_double
_PROTOS = {
    "wbmi_version": (_i, []),
    "wbmi_ctx_create": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "wbmi_ctx_destroy": (_i, [_vp]),
    "wbmi_ctx_set_stream": (_i, [_vp, _vp]),
    "wbmi_sync": (_i, [_vp]),
    "wbmi_last_error": (ctypes.c_char_p, [_vp]),
    "wbmi_build_channels": (_i, [_vp, _i, _i, _vp, _vp, _i, _vp, _i, _vp]),
    "wbmi_r2k_plan_create": (_i, [_vp, _i, _vp, _vp, _i64, ctypes.POINTER(_vp)]),
    "wbmi_r2k_plan_destroy": (_i, [_vp]),
    "wbmi_r2k_plan_sizes": (_i, [_vp, ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]),
    "wbmi_r2k_box": (_i, [_vp, _vp, _vp, _vp]),
    "wbmi_r2k_exec": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "wbmi_r2k_w1_bytes": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_size_t)]),
    "wbmi_r2k_zpass": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "wbmi_r2k_yx": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _i]),
    "wbmi_r2k_klist": (_i, [_vp, _i, _vp, _i64, _vp, _i, _vp, _vp, _i]),
    "wbmi_r2k_klist_tile": (_i, [_vp, _i, _vp, _i64, _vp, _i, _vp, _vp, _i, _i]),
    "wbmi_r2k_klist_tile_shape": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "wbmi_eigh": (_i, [_vp, _i, _i, _vp, _i64, _vp, _vp, _i, _i, _vp, _vp, ctypes.c_size_t]),
    "wbmi_eigh_noalloc": (_i, [_vp, _i, _i, _vp, _i64, _vp, _vp, _i, _i, _vp]),
    "wbmi_eigh_noalloc_max_bytes": (ctypes.c_size_t, []),
    "wbmi_eigh_work_bytes": (ctypes.c_size_t, [_i, _i, _i, _i]),
    "wbmi_eigh_guess": (_i, [_vp, _i, _i, _vp, _i64, _vp, _vp, _i, _i, _vp, _vp, _vp, ctypes.c_size_t]),
    "wbmi_optical_conductivity": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(OptcondParams), _vp]),
    "wbmi_rotate": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, ctypes.c_size_t]),
    "wbmi_rotate_work_bytes": (ctypes.c_size_t, [_i, _i, _i]),
    "wbmi_static_scan": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(ScanParams), _d, _vp]),
    "wbmi_finalize": (_i, [_vp, _vp, _i, _i, _i, _d, _d, _vp]),
    "wbmi_static_scan_k": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(ScanParams), _d, _vp]),
    "wbmi_finalize_k": (_i, [_vp, _vp, _i, _i, _i, _i, _d, _d, _vp]),
    "wbmi_static_scan_tetra": (_i, [_vp, _i, _i, _i, _vp, _i, _vp, _vp, _i, ctypes.POINTER(ScanParams), _vp, _i,
                                    _d, _vp]),
    "wbmi_finalize_tetra": (_i, [_vp, _vp, _i, _i, _d, _vp]),
    "wbmi_static_scan_tetra_k": (_i, [_vp, _i, _i, _i, _vp, _i, _vp, _vp, _i, ctypes.POINTER(ScanParams), _vp, _i,
                                      _d, _vp]),
    "wbmi_finalize_tetra_k": (_i, [_vp, _vp, _i, _i, _i, _d, _vp]),
    "wbmi_jdos": (_i, [_vp, _i, _i, _vp, ctypes.POINTER(OptcondParams), _vp]),
    "wbmi_spin_hall_work_bytes": (ctypes.c_size_t, [_i, _i]),
    "wbmi_tabulate": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(ScanParams), _vp]),
    "wbmi_tabulate_diag": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _vp, _d, _i, _vp]),
    "wbmi_tabulate_invmass": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _d, _i, _vp]),
    "wbmi_shift_current": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(OptcondParams), _d, _vp,
                                ctypes.c_size_t, _vp]),
    "wbmi_injection_current": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(OptcondParams), _vp,
                                    ctypes.c_size_t, _vp]),
    "wbmi_spin_hall": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(ShcParams), _vp, ctypes.c_size_t, _vp]),
    "wbmi_tabulate_spin_berry": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(ShcParams), _vp, ctypes.c_size_t,
                                      _vp]),
    "wbmi_spin_berry_bands": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(ShcParams), _vp, ctypes.c_size_t,
                                   _vp]),
    "wbmi_der3e_bands": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _d, _i, _i, _vp]),
    "wbmi_optical_conductivity_work_bytes": (ctypes.c_size_t, [_i, _i]),
    "wbmi_optical_conductivity_ws": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, ctypes.POINTER(OptcondParams), _vp,
                                          ctypes.c_size_t, _vp]),
}


def lib():
    """Load libwbmi.so (once).  Raises WbmiLibraryError if it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise WbmiLibraryError(
                    f"{LIB_PATH} not found: build it with `python -m wannierberri_amd.build_lib` "
                    "(the MI355X path has no CPU fallback)")
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _PROTOS.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
        return _lib


def _ptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


def _host_i32(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(ctypes.c_void_p)


class Context:
    """One library context bound to a device and a HIP stream (torch's current stream by default)."""

    def __init__(self, device=0, stream=None):
        import torch
        self.torch = torch
        self.device = torch.device("cuda", device)
        if stream is None:
            stream = torch.cuda.current_stream(self.device)
        self.stream = stream
        h = ctypes.c_void_p()
        rc = lib().wbmi_ctx_create(device, ctypes.c_void_p(stream.cuda_stream), ctypes.byref(h))
        if rc != WBMI_OK:
            raise WbmiError(f"wbmi_ctx_create failed with code {rc} (device {device}; gfx950 required)")
        self.h = h
        self.timer = None   # optional {stage: [(start_event, end_event, units)]}
        self.tuning = {}    # (knob, nw, prec) -> choice, filled by wannierberri_amd.autotune
        self._work = None   # device workspace shared by the calls on this context's stream

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _lib is not None:
                _lib.wbmi_ctx_destroy(self.h)
        except Exception:
            pass

    def _timed(self, stage, units, fn, *args):
        """Run a library call; with a timer set, bracket it with events on this stream."""
        if self.timer is None:
            return fn(*args)
        torch = self.torch
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(self.stream)
        rc = fn(*args)
        e1.record(self.stream)
        self.timer.setdefault(stage, []).append((e0, e1, units))
        return rc

    def check(self, rc, what):
        if rc != WBMI_OK:
            msg = lib().wbmi_last_error(self.h).decode(errors="replace")
            if rc == WBMI_ERR_ARG:
                raise ValueError(f"{what}: {msg}")
            if rc == WBMI_ERR_UNSUPPORTED:
                raise NotImplementedError(f"{what}: {msg}")
            raise WbmiError(f"{what}: {msg}")

    def sync(self):
        self.check(lib().wbmi_sync(self.h), "wbmi_sync")

    def set_stream(self, stream):
        """Launch on `stream` (a torch stream) from now on; returns the previous one."""
        prev = self.stream
        self.check(lib().wbmi_ctx_set_stream(self.h, ctypes.c_void_p(stream.cuda_stream)), "wbmi_ctx_set_stream")
        self.stream = stream
        return prev

    # -- stages --------------------------------------------------------------
    def build_channels(self, cRshift, base, spec, out):
        nR, nw = cRshift.shape[0], cRshift.shape[1]
        spec_a, spec_p = _host_i32(spec)
        self.check(lib().wbmi_build_channels(self.h, nR, nw, _ptr(cRshift), _ptr(base), base.shape[3], spec_p,
                                             spec_a.shape[0], _ptr(out)), "wbmi_build_channels")

    def r2k_plan(self, iRvec, NKFFT, nelem):
        return R2KPlan(self, iRvec, NKFFT, nelem)

    def r2k_klist(self, iRvec_dev, XR, klist_dev, out, prec, tile=None):
        """Phase-GEMM R -> k; `tile` defaults to the tuned tile for out's nw and `prec`
        (self.tuning[("klist_tile", nw, prec)], set by wannierberri_amd.autotune), else tile 0."""
        nR = iRvec_dev.shape[0]
        nelem = XR[0].numel()
        if tile is None:
            tile = self.tuning.get(("klist_tile", int(out.shape[-1]), int(prec)), 0)
        self.check(self._timed("r2k_klist", (klist_dev.shape[0], nelem, nR), lib().wbmi_r2k_klist_tile, self.h, nR,
                               _ptr(iRvec_dev), nelem, _ptr(XR), klist_dev.shape[0], _ptr(klist_dev), _ptr(out),
                               prec, int(tile)), "wbmi_r2k_klist")

    def _workspace(self, nbytes):
        """Device workspace of at least nbytes (kept and grown on demand; None for 0)."""
        if nbytes <= 0:
            return None
        import torch
        if self._work is None or self._work.numel() < nbytes:
            self._work = None
            self._work = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self._work

    def eigh(self, H, Hstride, E, U, prec, nw, refine=0, H64=None, U0=None):
        """Batched eigh; refine=n >= 1: FP32 Jacobi + up to n FP64 Ogita-Aishima steps.
        U0: optional starting eigenvectors (same shape and dtype as U; may be U itself)."""
        nk = E.shape[0]
        wbytes = int(lib().wbmi_eigh_work_bytes(nk, nw, prec, int(refine)))
        work = self._workspace(wbytes)
        if U0 is None:
            self.check(self._timed("eigh", (nk, nw), lib().wbmi_eigh, self.h, nk, nw, _ptr(H), Hstride,
                                   _ptr(E), _ptr(U), prec, int(refine), _ptr(H64), _ptr(work), wbytes), "wbmi_eigh")
        else:
            if tuple(U0.shape) != tuple(U.shape) or U0.dtype != U.dtype:
                raise ValueError("eigh: U0 must have the shape and dtype of U")
            self.check(self._timed("eigh", (nk, nw), lib().wbmi_eigh_guess, self.h, nk, nw, _ptr(H), Hstride,
                                   _ptr(E), _ptr(U), prec, int(refine), _ptr(H64), _ptr(U0), _ptr(work), wbytes),
                       "wbmi_eigh_guess")

    def eigh_noalloc(self, H, Hstride, E, U, prec, nw, refine=0, H64=None):
        """wbmi_eigh_noalloc: eigh with the library's fixed scratch instead of a workspace."""
        nk = E.shape[0]
        self.check(lib().wbmi_eigh_noalloc(self.h, nk, nw, _ptr(H), Hstride, _ptr(E), _ptr(U), prec, int(refine),
                                           _ptr(H64)), "wbmi_eigh_noalloc")

    def rotate(self, U, X, herm, skip, prec):
        nk, nch, nw = X.shape[0], X.shape[1], X.shape[2]
        h_a, h_p = _host_i32(herm)
        s_a, s_p = _host_i32(skip)
        nrot = int(nch - (0 if s_a is None else np.count_nonzero(s_a)))
        wbytes = int(lib().wbmi_rotate_work_bytes(nk, nw, prec))
        work = self._workspace(wbytes)
        stage = "rotate_split" if prec == WBMI_F32_SPLIT else "rotate"
        self.check(self._timed(stage, (nk, nw, nrot), lib().wbmi_rotate, self.h, nk, nw, nch, _ptr(U), _ptr(X),
                               h_p, s_p, prec, _ptr(work), wbytes), "wbmi_rotate")

    def static_scan(self, E, X, prec, params, weight, hist, nch=None):
        nk, nw = E.shape
        if nch is None:
            nch = X.shape[1] if X is not None else 1
        self.check(self._timed(f"scan_f{params.formula}", (nk, nw), lib().wbmi_static_scan, self.h, nk, nw, nch,
                               _ptr(E), _ptr(X), prec, ctypes.byref(params), float(weight), _ptr(hist)),
                   "wbmi_static_scan")

    def static_scan_tetra(self, E, Ec, X, prec, params, Efermi, weight, hist, nch=None):
        """Tetrahedron-weighted scan: Ec [8, nk, nw], Efermi [nEF] (device, ascending),
        hist [2 nEF, ncomp]."""
        nk, nw = E.shape
        if nch is None:
            nch = X.shape[1] if X is not None else 1
        if tuple(Ec.shape[1:]) != (nk, nw) or Ec.shape[0] not in (4, 8) or \
                tuple(hist.shape) != (2 * Efermi.shape[0], hist.shape[1]):
            raise ValueError("static_scan_tetra: Ec must be [8 or 4, nk, nw] and hist [2 nEF, ncomp]")
        self.check(self._timed(f"scan_tetra_f{params.formula}", (nk, nw), lib().wbmi_static_scan_tetra, self.h, nk,
                               nw, nch, _ptr(E), int(Ec.shape[0]), _ptr(Ec), _ptr(X), prec, ctypes.byref(params),
                               _ptr(Efermi),
                               int(Efermi.shape[0]), float(weight), _ptr(hist)),
                   "wbmi_static_scan_tetra")

    def static_scan_tetra_k(self, E, Ec, X, prec, params, Efermi, weight, hist, nch=None):
        """k-resolved tetrahedron scan: hist [nk, 2 nEF, ncomp]."""
        nk, nw = E.shape
        if nch is None:
            nch = X.shape[1] if X is not None else 1
        if tuple(Ec.shape[1:]) != (nk, nw) or Ec.shape[0] not in (4, 8) or \
                tuple(hist.shape[:2]) != (nk, 2 * Efermi.shape[0]):
            raise ValueError("static_scan_tetra_k: Ec must be [8 or 4, nk, nw] and hist [nk, 2 nEF, ncomp]")
        self.check(self._timed(f"scan_tetra_k_f{params.formula}", (nk, nw), lib().wbmi_static_scan_tetra_k, self.h,
                               nk, nw, nch, _ptr(E), int(Ec.shape[0]), _ptr(Ec), _ptr(X), prec, ctypes.byref(params),
                               _ptr(Efermi),
                               int(Efermi.shape[0]), float(weight), _ptr(hist)), "wbmi_static_scan_tetra_k")

    def finalize_tetra_k(self, hist, scale, out):
        nk, nEF2, ncomp = hist.shape
        self.check(lib().wbmi_finalize_tetra_k(self.h, _ptr(hist), nk, nEF2 // 2, ncomp, float(scale), _ptr(out)),
                   "wbmi_finalize_tetra_k")

    def finalize_tetra(self, hist, scale, out):
        nEF2, ncomp = hist.shape
        self.check(lib().wbmi_finalize_tetra(self.h, _ptr(hist), nEF2 // 2, ncomp, float(scale), _ptr(out)),
                   "wbmi_finalize_tetra")

    def optical_conductivity(self, E, X, prec, params, out):
        nk, nw = E.shape
        wbytes = int(lib().wbmi_optical_conductivity_work_bytes(nk, nw))
        if wbytes == 0:
            self.check(self._timed("optcond", (nk, nw), lib().wbmi_optical_conductivity, self.h, nk, nw, X.shape[1],
                                   _ptr(E), _ptr(X), prec, ctypes.byref(params), _ptr(out)),
                       "wbmi_optical_conductivity")
            return
        work = self._workspace(wbytes)   # nw > 32: A of every k
        self.check(self._timed("optcond", (nk, nw), lib().wbmi_optical_conductivity_ws, self.h, nk, nw, X.shape[1],
                               _ptr(E), _ptr(X), prec, ctypes.byref(params), _ptr(work), wbytes, _ptr(out)),
                   "wbmi_optical_conductivity_ws")

    def jdos(self, E, params, out):
        nk, nw = E.shape
        self.check(self._timed("jdos", (nk, nw), lib().wbmi_jdos, self.h, nk, nw, _ptr(E), ctypes.byref(params),
                               _ptr(out)), "wbmi_jdos")

    def spin_hall(self, E, X, prec, params, out):
        nk, nw = E.shape
        wbytes = int(lib().wbmi_spin_hall_work_bytes(nk, nw))
        work = self._workspace(wbytes)
        self.check(self._timed("shc", (nk, nw), lib().wbmi_spin_hall, self.h, nk, nw, X.shape[1], _ptr(E), _ptr(X),
                               prec, ctypes.byref(params), _ptr(work), wbytes, _ptr(out)), "wbmi_spin_hall")

    def tabulate_spin_berry(self, E, X, prec, params, out, per_band=False):
        nk, nw = E.shape
        wbytes = int(lib().wbmi_spin_hall_work_bytes(nk, nw))
        work = self._workspace(wbytes)
        fn = lib().wbmi_spin_berry_bands if per_band else lib().wbmi_tabulate_spin_berry
        self.check(self._timed("spin_berry_bands" if per_band else "tab_spin_berry", (nk, nw), fn, self.h, nk, nw,
                               X.shape[1], _ptr(E), _ptr(X), prec, ctypes.byref(params), _ptr(work), wbytes, _ptr(out)),
                   "wbmi_spin_berry_bands" if per_band else "wbmi_tabulate_spin_berry")

    def tabulate(self, E, X, prec, params, out):
        nk, nw = E.shape
        nch = X.shape[1] if X is not None else 1
        self.check(self._timed(f"tab_f{params.formula}", (nk, nw), lib().wbmi_tabulate, self.h, nk, nw, nch, _ptr(E),
                               _ptr(X), prec, ctypes.byref(params), _ptr(out)), "wbmi_tabulate")

    def der3e_bands(self, E, X, prec, chV, chW, chW3, degen_thresh, degen_Kramers, mode, out):
        nk, nw = E.shape
        v_a, v_p = _host_i32(np.asarray(chV, dtype=np.int32))
        w_a, w_p = _host_i32(np.asarray(chW, dtype=np.int32))
        w3_a, w3_p = _host_i32(np.asarray(chW3, dtype=np.int32))
        self.check(self._timed("der3e", (nk, nw), lib().wbmi_der3e_bands, self.h, nk, nw, X.shape[1], _ptr(E), _ptr(X),
                               prec, v_p, w_p, w3_p, float(degen_thresh), int(bool(degen_Kramers)), int(mode),
                               _ptr(out)), "wbmi_der3e_bands")

    def tabulate_diag(self, E, X, prec, ch3, degen_thresh, degen_Kramers, out):
        nk, nw = E.shape
        c_a, c_p = _host_i32(np.asarray(ch3, dtype=np.int32))
        self.check(self._timed("tab_diag", (nk, nw), lib().wbmi_tabulate_diag, self.h, nk, nw, X.shape[1], _ptr(E),
                               _ptr(X), prec, c_p, float(degen_thresh), int(bool(degen_Kramers)), _ptr(out)),
                   "wbmi_tabulate_diag")

    def shift_current(self, E, X, prec, params, sc_eta, out):
        nk, nw = E.shape
        wbytes = int(lib().wbmi_spin_hall_work_bytes(nk, nw))
        work = self._workspace(wbytes)
        self.check(self._timed("shift", (nk, nw), lib().wbmi_shift_current, self.h, nk, nw, X.shape[1], _ptr(E),
                               _ptr(X), prec, ctypes.byref(params), float(sc_eta), _ptr(work), wbytes, _ptr(out)),
                   "wbmi_shift_current")

    def injection_current(self, E, X, prec, params, out):
        nk, nw = E.shape
        wbytes = int(lib().wbmi_spin_hall_work_bytes(nk, nw))
        work = self._workspace(wbytes)
        self.check(self._timed("injection", (nk, nw), lib().wbmi_injection_current, self.h, nk, nw, X.shape[1],
                               _ptr(E), _ptr(X), prec, ctypes.byref(params), _ptr(work), wbytes, _ptr(out)),
                   "wbmi_injection_current")

    def static_scan_k(self, E, X, prec, params, weight, hist, nch=None):
        nk, nw = E.shape
        if nch is None:
            nch = X.shape[1] if X is not None else 1
        self.check(self._timed(f"scan_k_f{params.formula}", (nk, nw), lib().wbmi_static_scan_k, self.h, nk, nw, nch,
                               _ptr(E), _ptr(X), prec, ctypes.byref(params), float(weight), _ptr(hist)),
                   "wbmi_static_scan_k")

    def finalize_k(self, hist, fder, dEF, scale, out):
        nk, nEF, ncomp = hist.shape
        self.check(lib().wbmi_finalize_k(self.h, _ptr(hist), nk, nEF, ncomp, fder, float(dEF), float(scale),
                                         _ptr(out)), "wbmi_finalize_k")

    def tabulate_invmass(self, E, X, prec, chV, chW, degen_thresh, degen_Kramers, out):
        nk, nw = E.shape
        v_a, v_p = _host_i32(np.asarray(chV, dtype=np.int32))
        w_a, w_p = _host_i32(np.asarray(chW, dtype=np.int32))
        self.check(self._timed("tab_invmass", (nk, nw), lib().wbmi_tabulate_invmass, self.h, nk, nw, X.shape[1],
                               _ptr(E), _ptr(X), prec, v_p, w_p, float(degen_thresh), int(bool(degen_Kramers)),
                               _ptr(out)), "wbmi_tabulate_invmass")

    def finalize(self, hist, fder, dEF, scale, out):
        nEF, ncomp = hist.shape
        self.check(lib().wbmi_finalize(self.h, _ptr(hist), nEF, ncomp, fder, float(dEF), float(scale), _ptr(out)),
                   "wbmi_finalize")


class R2KPlan:
    def __init__(self, ctx, iRvec, NKFFT, nelem):
        self.ctx = ctx
        iR_a, iR_p = _host_i32(np.asarray(iRvec).reshape(-1, 3))
        nk_a = (ctypes.c_int32 * 3)(*[int(x) for x in NKFFT])
        h = ctypes.c_void_p()
        ctx.check(lib().wbmi_r2k_plan_create(ctx.h, iR_a.shape[0], iR_p, nk_a, int(nelem), ctypes.byref(h)),
                  "wbmi_r2k_plan_create")
        self.h = h
        b = ctypes.c_size_t()
        w = ctypes.c_size_t()
        ctx.check(lib().wbmi_r2k_plan_sizes(h, ctypes.byref(b), ctypes.byref(w)), "wbmi_r2k_plan_sizes")
        self.box_bytes, self.work_bytes = b.value, w.value

    def __del__(self):
        try:
            if getattr(self, "h", None) is not None and _lib is not None:
                _lib.wbmi_r2k_plan_destroy(self.h)
        except Exception:
            pass

    def box(self, XR, box):
        self.ctx.check(lib().wbmi_r2k_box(self.ctx.h, self.h, _ptr(XR), _ptr(box)), "wbmi_r2k_box")

    def execute(self, XR, box, dK, work, out, prec):
        dk = (ctypes.c_double * 3)(*[float(x) for x in dK])
        self.ctx.check(self.ctx._timed("r2k", (out.shape[0], out[0].numel(), self.box_bytes, self.work_bytes),
                                       lib().wbmi_r2k_exec, self.ctx.h, self.h, _ptr(XR), _ptr(box), dk, _ptr(work),
                                       _ptr(out), prec), "wbmi_r2k_exec")

    def w1_bytes(self, prec):
        """Bytes of the z-pass output (wbmi_r2k_zpass) in the form `prec`."""
        b = ctypes.c_size_t()
        self.ctx.check(lib().wbmi_r2k_w1_bytes(self.h, prec, ctypes.byref(b)), "wbmi_r2k_w1_bytes")
        return b.value

    def zpass(self, box, dK, work, w1, prec):
        dk = (ctypes.c_double * 3)(*[float(x) for x in dK])
        self.ctx.check(self.ctx._timed("r2k_z", (w1.numel(), self.box_bytes), lib().wbmi_r2k_zpass, self.ctx.h, self.h,
                                       _ptr(box), dk, _ptr(work), _ptr(w1), prec), "wbmi_r2k_zpass")

    def yx(self, w1, dK, work, out, prec):
        dk = (ctypes.c_double * 3)(*[float(x) for x in dK])
        self.ctx.check(self.ctx._timed("r2k", (out.shape[0], out[0].numel(), self.box_bytes, self.work_bytes),
                                       lib().wbmi_r2k_yx, self.ctx.h, self.h, _ptr(w1), dk, _ptr(work), _ptr(out),
                                       prec), "wbmi_r2k_yx")


_contexts = {}


def get_context(device=0):
    """Per-device context on torch's current stream."""
    import torch
    key = (device, torch.cuda.current_stream(device).cuda_stream)
    if key not in _contexts:
        _contexts[key] = Context(device)
    return _contexts[key]
