"""ctypes binding of libstaqdic.so (the C ABI in include/staq_dic.h).

The library is the product: nothing here computes on the CPU.  A missing
library, or a call on a machine without a GPU, raises.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libstaqdic.so")

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class StaqParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("mu", C.c_double), ("max_iter", C.c_int),
                ("plain", C.c_int), ("block_size", C.c_int), ("persistent", C.c_int),
                ("guard", C.c_double)]


class BatchPtrs(C.Structure):
    _fields_ = [("coords", _vp), ("U0", _vp), ("UOld", _vp), ("FOld", _vp), ("vdual", _vp),
                ("U", _vp), ("F", _vp), ("Hpar", _vp), ("status", _vp), ("wx", _vp),
                ("wy", _vp), ("slot", _vp), ("N", C.c_long), ("wmax", C.c_int)]


# name -> (restype, argtypes); every symbol include/staq_dic.h declares
SIGNATURES = {
    "staq_params_default": (None, [C.POINTER(StaqParams)]),
    "staq_ctx_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "staq_ctx_destroy": (None, [_vp]),
    "staq_last_error": (C.c_char_p, [_vp]),
    "staq_ctx_stream": (_vp, [_vp]),
    "staq_set_frame": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, C.c_long, C.c_long]),
    "staq_set_frame_f32": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, C.c_long, C.c_long]),
    "staq_set_frame_dev": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, C.c_long, C.c_long, _vp]),
    "staq_get_gradient": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "staq_ba_interp2": (C.c_int, [_vp, _dp, C.c_long, C.c_long, C.c_long, _dp, _dp, C.c_long, _dp]),
    "staq_ba_interp2_f32": (C.c_int, [_vp, _fp, C.c_long, C.c_long, C.c_long, _fp, _fp, C.c_long,
                                      _fp]),
    "staq_ba_interp2_dev": (C.c_int, [_vp, _vp, C.c_long, C.c_long, C.c_long, _vp, _vp, C.c_long,
                                      _vp, C.c_int, _vp]),
    "staq_img_gradient": (C.c_int, [_vp, _dp, _dp, C.c_long, C.c_long, _dp, _dp]),
    "staq_integer_search": (C.c_int, [_vp, C.c_int, C.c_long, C.c_long, _lp, _lp, C.c_long,
                                      C.c_long, C.c_long, _lp, _lp, _dp, _dp, _dp, _dp, _dp]),
    "staq_xcorr_full_peak": (C.c_int, [_vp, C.c_int, _lp, _lp, _dp]),
    "staq_xcorr_boxes": (C.c_int, [_vp, C.c_int, C.c_long, _lp, _lp, C.c_long, _dp, _ip]),
    "staq_xcorr_qfactors": (C.c_int, [_vp, C.c_int, C.c_long, _lp, _lp, _dp, _ip]),
    "staq_natural_interp": (C.c_int, [C.c_long, _dp, C.c_long, _ip, _ip, _lp, _ip, C.c_long, _dp,
                                      C.c_long, _dp, _ip, _dp]),
    "staq_natural_weights": (C.c_int, [C.c_long, _dp, C.c_long, _ip, _ip, _lp, _ip, C.c_long, _dp,
                                       _ip, C.c_long, _ip, _dp, _ip]),
    "staq_local_icgn": (C.c_int, [_vp, C.POINTER(StaqParams), C.c_long, _dp, _dp, _lp, _ip, _dp,
                                  _dp, _dp, _ip]),
    "staq_subpb1": (C.c_int, [_vp, C.POINTER(StaqParams), C.c_long, _dp, _dp, _dp, _dp, _dp, _lp,
                              _lp, _ip, _dp, _ip]),
    "staq_local_icgn_noalloc": (C.c_int, [_vp, C.POINTER(StaqParams), C.c_int, C.c_long, C.c_int,
                                          _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "staq_znssd_fwd": (C.c_int, [_vp, C.c_long, _dp, _dp, _lp, _ip, _dp, _ip]),
    "staq_znssd_bwd": (C.c_int, [_vp, C.c_long, _dp, _dp, _lp, _ip, _dp, _dp, _dp]),
    "staq_znssd_fwd_dev": (C.c_int, [_vp, C.c_long, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "staq_znssd_bwd_dev": (C.c_int, [_vp, C.c_long, C.c_int, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "staq_batch_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_long, _dp, _lp, _lp, _ip]),
    "staq_batch_destroy": (None, [_vp]),
    "staq_batch_set_input": (C.c_int, [_vp, C.c_int, _dp]),
    "staq_batch_icgn6": (C.c_int, [_vp, C.POINTER(StaqParams)]),
    "staq_batch_icgn2": (C.c_int, [_vp, C.POINTER(StaqParams)]),
    "staq_batch_icgn6_per_subset": (C.c_int, [_vp, C.POINTER(StaqParams), _dp]),
    "staq_batch_step_timed": (C.c_int, [_vp, C.POINTER(StaqParams), C.c_int, _fp]),
    "staq_batch_get_output": (C.c_int, [_vp, _dp, _dp, _dp, _ip]),
    "staq_batch_chain_outputs": (C.c_int, [_vp]),
    "staq_batch_reorder": (C.c_int, [_vp]),
    "staq_batch_autotune": (C.c_int, [_vp, C.POINTER(StaqParams), C.c_int]),
    "staq_batch_graph_capture": (C.c_int, [_vp, C.POINTER(StaqParams), C.c_int]),
    "staq_batch_graph_launch": (C.c_int, [_vp]),
    "staq_batch_exact_count": (C.c_int, [_vp, _lp]),
    "staq_batch_get_ptrs": (C.c_int, [_vp, C.POINTER(BatchPtrs)]),
    "staq_batch_bins": (C.c_int, [_vp, _ip, _ip, _lp, _ip, _ip, _ip, C.c_int]),
    "staq_batch_bin_of": (C.c_int, [_vp, _ip]),
    "staq_batch_set_form": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
}

IN_U0, IN_UOLD, IN_FOLD, IN_VDUAL = 0, 1, 2, 3

_lib = None


def load():
    """Load libstaqdic.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m staq_dic_amd.build` "
                "(or __graft_entry__.build())")
        # PyTorch-ROCm ships its own libamdhip64 (soname libamdhip64.so.7) and
        # loads it under the name "libamdhip64.so"; loaded after this library
        # it would bring a second HIP runtime into the process, which then
        # finds no GPU.  Loading torch first makes this library bind to the
        # runtime torch already holds, so tensors and streams are shared.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def d(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def f32(a: np.ndarray):
    return a.ctypes.data_as(_fp)


def lp(a: np.ndarray):
    return a.ctypes.data_as(_lp)


def ip(a: np.ndarray):
    return a.ctypes.data_as(_ip)


class StaqError(RuntimeError):
    pass


def params(tol=1e-2, mu=1e-3, max_iter=100, plain=False, block_size=0, persistent=False,
           guard=None):
    """staq_params with the library's defaults; guard None keeps the default
    guard band (0 turns the FP64 re-solve off)."""
    p = StaqParams()
    load().staq_params_default(C.byref(p))
    p.tol, p.mu, p.max_iter = tol, mu, max_iter
    p.plain, p.block_size, p.persistent = int(plain), block_size, int(persistent)
    if guard is not None:
        p.guard = guard
    return p


class Context:
    """One device context: a HIP stream plus the frame pairs resident in HBM."""

    def __init__(self, device: int = 0):
        self.L = load()
        h = _vp()
        rc = self.L.staq_ctx_create(C.byref(h), device)
        if rc != 0:
            raise StaqError(f"staq_ctx_create failed ({rc}): no usable HIP device {device}")
        self.h = h
        self.frames = {}
        self._batches = weakref.WeakSet()

    def close(self):
        # batches hold a pointer to the context: destroy them first
        for b in list(getattr(self, "_batches", ())):
            b.close()
        if getattr(self, "h", None):
            self.L.staq_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc, what="call"):
        if rc != 0:
            msg = self.L.staq_last_error(self.h)
            raise StaqError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    @property
    def stream(self) -> int:
        return self.L.staq_ctx_stream(self.h)

    def set_frame(self, f, g, mask=None, slot=0):
        """Upload f, g (and mask) given as (H, W) arrays indexed [x-1, y-1]."""
        f = np.asfortranarray(f, dtype=np.float64)
        g = np.asfortranarray(g, dtype=np.float64)
        H, W = f.shape
        m = None if mask is None else np.asfortranarray(mask, dtype=np.float64)
        self.check(self.L.staq_set_frame(self.h, slot, d(f), d(g), None if m is None else d(m),
                                         H, W), "staq_set_frame")
        self.frames[slot] = (H, W)

    def gradient(self, slot=0):
        H, W = self.frames[slot]
        dx = np.zeros((H - 6, W - 6), order="F")
        dy = np.zeros((H - 6, W - 6), order="F")
        self.check(self.L.staq_get_gradient(self.h, slot, d(dx), d(dy)), "staq_get_gradient")
        return dx, dy


class Batch:
    """Device-resident subset list with its inputs and outputs."""

    def __init__(self, ctx: Context, coords, wx, wy=None, slot=None):
        self.ctx = ctx
        self.L = ctx.L
        coords = np.asfortranarray(coords, dtype=np.float64)
        self.N = coords.shape[0]
        self.coords = coords
        self.wx = np.ascontiguousarray(np.broadcast_to(wx, (self.N,)), dtype=np.int64)
        self.wy = None if wy is None else np.ascontiguousarray(np.broadcast_to(wy, (self.N,)),
                                                               dtype=np.int64)
        self.slot = None if slot is None else np.ascontiguousarray(slot, dtype=np.int32)
        h = _vp()
        ctx.check(self.L.staq_batch_create(ctx.h, C.byref(h), self.N, d(coords), lp(self.wx),
                                           None if self.wy is None else lp(self.wy),
                                           None if self.slot is None else ip(self.slot)),
                  "staq_batch_create")
        self.h = h
        ctx._batches.add(self)

    def close(self):
        if getattr(self, "h", None) and getattr(self.ctx, "h", None):
            self.L.staq_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_input(self, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        n = (4 if which == IN_FOLD else 2) * self.N
        if arr.size != n:
            raise ValueError(f"input {which} needs {n} values, got {arr.size}")
        self.ctx.check(self.L.staq_batch_set_input(self.h, which, d(arr)), "staq_batch_set_input")

    def icgn6(self, p):
        self.ctx.check(self.L.staq_batch_icgn6(self.h, C.byref(p)), "staq_batch_icgn6")

    def icgn2(self, p):
        self.ctx.check(self.L.staq_batch_icgn2(self.h, C.byref(p)), "staq_batch_icgn2")

    def icgn6_per_subset(self, p, timed=False):
        ms = C.c_double(0.0)
        self.ctx.check(self.L.staq_batch_icgn6_per_subset(self.h, C.byref(p),
                                                          C.byref(ms) if timed else None),
                       "staq_batch_icgn6_per_subset")
        return ms.value

    def step_timed(self, p, mode):
        """One eager step; returns the device milliseconds of each bin's launch
        followed by the fallback launch's."""
        nb = len(self.bins())
        ms = np.zeros(nb + 1, dtype=np.float32)
        self.ctx.check(self.L.staq_batch_step_timed(self.h, C.byref(p), mode,
                                                    ms.ctypes.data_as(C.POINTER(C.c_float))),
                       "staq_batch_step_timed")
        return ms.astype(np.float64)

    def reorder(self):
        self.ctx.check(self.L.staq_batch_reorder(self.h), "staq_batch_reorder")

    def chain(self):
        self.ctx.check(self.L.staq_batch_chain_outputs(self.h), "staq_batch_chain_outputs")

    def autotune(self, p, mode):
        self.ctx.check(self.L.staq_batch_autotune(self.h, C.byref(p), mode), "staq_batch_autotune")

    def graph_capture(self, p, mode):
        self.ctx.check(self.L.staq_batch_graph_capture(self.h, C.byref(p), mode),
                       "staq_batch_graph_capture")

    def graph_launch(self):
        self.ctx.check(self.L.staq_batch_graph_launch(self.h), "staq_batch_graph_launch")

    def outputs(self):
        N = self.N
        U = np.zeros(2 * N)
        F = np.zeros(4 * N)
        Hp = np.zeros((N, 21), order="F")
        st = np.zeros(N, dtype=np.int32)
        self.ctx.check(self.L.staq_batch_get_output(self.h, d(U), d(F), d(Hp), ip(st)),
                       "staq_batch_get_output")
        return U, F, Hp, st

    def exact_count(self) -> int:
        """Subsets the last step solved again in FP64 (guard band)."""
        n = C.c_long(0)
        self.ctx.check(self.L.staq_batch_exact_count(self.h, C.byref(n)), "staq_batch_exact_count")
        return n.value

    def ptrs(self) -> BatchPtrs:
        p = BatchPtrs()
        self.ctx.check(self.L.staq_batch_get_ptrs(self.h, C.byref(p)), "staq_batch_get_ptrs")
        return p

    def set_forms(self, forms, mode):
        """Kernel form of every bin for mode 6 or 2 (a list as in bins())."""
        for k, f in enumerate(forms):
            self.ctx.check(self.L.staq_batch_set_form(self.h, k, mode, int(f)), "staq_batch_set_form")

    def bin_of(self):
        """Bin index of every subset, in the caller's order."""
        out = np.zeros(self.N, dtype=np.int32)
        self.ctx.check(self.L.staq_batch_bin_of(self.h, ip(out)), "staq_batch_bin_of")
        return out

    def bins(self):
        n = C.c_int()
        wc = np.zeros(64, dtype=np.int32)
        cnt = np.zeros(64, dtype=np.int64)
        b6 = np.zeros(64, dtype=np.int32)
        b2 = np.zeros(64, dtype=np.int32)
        pp = np.zeros(64, dtype=np.int32)
        self.ctx.check(self.L.staq_batch_bins(self.h, C.byref(n), ip(wc), lp(cnt), ip(b6), ip(b2),
                                              ip(pp), 64), "staq_batch_bins")
        k = n.value
        return [dict(wcap=int(wc[i]), count=int(cnt[i]), bs6=int(b6[i]), bs2=int(b2[i]),
                     ppt=int(pp[i])) for i in range(k)]
