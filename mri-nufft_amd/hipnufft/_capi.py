"""ctypes binding of ``libhipnufft.so`` (the C ABI declared in include/hipnufft.h).

The library links against ``libamdhip64.so.7``; it is loaded only after
``torch`` so that it binds to the HIP runtime torch already loaded (one HIP
runtime per process, shared device pointers and streams).  There is no CPU
fallback: if the library is missing or the device is absent every call raises.
"""

from __future__ import annotations

import ctypes
import os
import threading

import numpy as np
from pathlib import Path

import torch  # noqa: F401  (must be imported before the library is loaded)

LIB_PATH = Path(os.environ.get("HIPNUFFT_LIBRARY")
                or Path(__file__).resolve().parent / "_lib" / "libhipnufft.so")

ERR_ARG = -1
ERR_HIP = -2
ERR_NOPTS = -3
ERR_ALLOC = -4

STAGES = ("rows_fwd", "cols_fwd", "slabs_fwd", "interp", "spread", "slabs_adj", "cols_adj",
          "rows_adj", "grad_traj")

_lock = threading.Lock()
_lib = None


class HipNufftError(RuntimeError):
    """Failure reported by libhipnufft (HIP error, allocation failure...)."""


class Opts(ctypes.Structure):
    """Mirror of ``hipnufft_opts``."""

    _fields_ = [
        ("upsampfac", ctypes.c_double),
        ("spreadinterponly", ctypes.c_int),
        ("tile", ctypes.c_int * 3),
        ("spread_group", ctypes.c_int),
        ("interp_group", ctypes.c_int),
        ("max_subproblem", ctypes.c_int),
        ("deterministic", ctypes.c_int),
        ("spread_max", ctypes.c_int),
        ("spread_method", ctypes.c_int),
        ("interp_span", ctypes.c_int),
    ]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_F = ctypes.c_float

_SIGNATURES = {
    "hipnufft_default_opts": [ctypes.POINTER(Opts)],
    "hipnufft_plan_create": [_I, _I, ctypes.POINTER(_I64), _I, _I, ctypes.c_double,
                             ctypes.POINTER(Opts), _P, ctypes.POINTER(_P)],
    "hipnufft_setpts": [_P, _I64, _P, _P, _P],
    "hipnufft_setpts_sets": [_P, _I, ctypes.POINTER(_I64), _P, _P, _P],
    "hipnufft_execute": [_P, _P, _P],
    "hipnufft_execute_adjoint": [_P, _P, _P],
    "hipnufft_sense_op": [_P, _P, _P, _P, _I, _I, _F],
    "hipnufft_sense_adj": [_P, _P, _P, _P, _P, _I, _I, _F],
    "hipnufft_sense_dc": [_P, _P, _P, _P, _P, _P, _I, _I, _F],
    "hipnufft_orc_op": [_P, _P, _P, _P, _P, _I, _I64, _P, _I, _I, _F],
    "hipnufft_orc_adj": [_P, _P, _P, _P, _P, _I, _I64, _P, _P, _I, _I, _F],
    "hipnufft_sense_grad_traj": [_P, _P, _P, _P, _I, _I, _F, _I, _I, _P],
    "hipnufft_dot_c64": [_P, _P, _I64, _I, _P, _P, _P],
    "hipnufft_dot_c64_rows": [_P, _P, _I64, _I64, _I, _P, _P, _P],
    "hipnufft_dot_c64_rows_noalloc": [_P, _P, _I64, _I64, _I, _P, _P],
    "hipnufft_get_sort": [_P, _P, _P, _P, _P, ctypes.POINTER(_I64)],
    "hipnufft_plan_info": [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I),
                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I),
                           ctypes.POINTER(_I64)],
    "hipnufft_workspace_bytes": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "hipnufft_set_stream": [_P, _P],
    "hipnufft_set_timing": [_P, _I],
    "hipnufft_get_timing": [_P, _P, _P, _I],
    "hipnufft_plan_destroy": [_P],
    "hipnufft_gram_create": [_I, ctypes.POINTER(_I64), _I64, _P, _P, _P, _P, ctypes.c_double, _I,
                             _F, _I, _P, ctypes.POINTER(_P)],
    "hipnufft_gram_apply": [_P, _P, _P, _P, _I, _I],
    "hipnufft_gram_set_stream": [_P, _P],
    "hipnufft_gram_workspace_bytes": [_P, ctypes.POINTER(ctypes.c_size_t)],
    "hipnufft_gram_destroy": [_P],
    "hipnufft_stack_create": [_I64, _I64, _I64, ctypes.POINTER(ctypes.c_int32), _I, _P,
                              ctypes.POINTER(_P)],
    "hipnufft_stack_fwd": [_P, _P, _P, _I, _I, _I, _F, _P],
    "hipnufft_stack_adj": [_P, _P, _P, _I, _I, _I, _F, _P],
    "hipnufft_stack_set_stream": [_P, _P],
    "hipnufft_stack_destroy": [_P],
    "hipnufft_kspace_window": [_P, _I64, _I, ctypes.POINTER(_F), _I, _P, _P, _I, _P, _P],
    "hipnufft_rss_normalize": [_P, _I, _I64, _F, _P],
    "hipnufft_blur_magnitude": [_P, _I, _I, ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_double),
                                ctypes.c_double, _P, _P],
}

EXPORTED = sorted(list(_SIGNATURES) + ["hipnufft_last_error"])


def lib():
    """Load the library once (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not LIB_PATH.exists():
                raise HipNufftError(
                    f"{LIB_PATH} is missing: build it with `python -m hipnufft._build` "
                    "(or __graft_entry__.build())"
                )
            h = ctypes.CDLL(str(LIB_PATH))
            for name, args in _SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = _I
            h.hipnufft_last_error.argtypes = []
            h.hipnufft_last_error.restype = ctypes.c_char_p
            _lib = h
    return _lib


def check(rc: int, what: str = ""):
    """Raise ValueError for argument errors and HipNufftError otherwise."""
    if rc == 0:
        return
    msg = lib().hipnufft_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == ERR_ARG or rc == ERR_NOPTS:
        raise ValueError(text)
    raise HipNufftError(text)


def ptr(t) -> int | None:
    """Device pointer of a tensor (None for None).

    The library reads raw memory, so lazy conj / neg views and strided
    tensors are refused rather than silently read as their storage.
    """
    if t is None:
        return None
    if t.is_conj() or t.is_neg() or not t.is_contiguous():
        raise ValueError("libhipnufft needs contiguous tensors without lazy conj/neg views")
    return t.data_ptr()


def current_stream() -> int:
    return torch.cuda.current_stream().cuda_stream


SPREAD_METHODS = {"auto": 0, "cells": 1, "blocks": 2, "columns": 3, "matrix": 4}


def spread_method_id(m):
    """HIPNUFFT_SPREAD_* value of a spreading-kernel name or number."""
    if isinstance(m, str):
        if m not in SPREAD_METHODS:
            raise ValueError(f"spread_method must be one of {sorted(SPREAD_METHODS)}")
        return SPREAD_METHODS[m]
    return int(m)


class RawPlan:
    """Owns one ``hipnufft_plan`` (both transform directions).

    Mirrors ``RawCufinufftPlan`` (``operators/interfaces/cufinufft.py:51-137``)
    with a single plan object for type 1 and type 2, like finufft's
    ``Plan.execute`` / ``execute_adjoint`` (``interfaces/finufft.py:64-76``).
    """

    def __init__(self, shape, n_trans=1, eps=1e-6, isign=-1, upsampfac=2.0,
                 spreadinterponly=False, tile=None, spread_group=0,
                 interp_group=0, max_subproblem=0, spread_max=0, spread_method=0,
                 interp_span=0, stream=None):
        L = lib()
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.n_trans = int(n_trans)
        self.eps = float(eps)
        self.isign = int(isign)
        o = Opts()
        check(L.hipnufft_default_opts(ctypes.byref(o)))
        o.upsampfac = float(upsampfac)
        o.spreadinterponly = 1 if spreadinterponly else 0
        if tile is not None:
            for d, t in enumerate(tile):
                o.tile[d] = int(t)
        o.spread_group = int(spread_group)
        o.interp_group = int(interp_group)
        o.max_subproblem = int(max_subproblem)
        o.spread_max = int(spread_max)
        o.spread_method = spread_method_id(spread_method)
        o.interp_span = int(interp_span)
        self.opts = o
        nm = (_I64 * self.ndim)(*self.shape)
        h = _P()
        self.stream = current_stream() if stream is None else stream
        check(L.hipnufft_plan_create(2, self.ndim, nm, self.isign, self.n_trans, self.eps,
                                     ctypes.byref(o), self.stream, ctypes.byref(h)),
              "plan_create")
        self._h = h
        self.n_samples = 0

    @property
    def handle(self):
        return self._h

    def set_stream(self, stream):
        self.stream = stream
        check(lib().hipnufft_set_stream(self._h, stream))

    def _sync_stream(self):
        s = current_stream()
        if s != self.stream:
            self.set_stream(s)

    def setpts(self, coords):
        """``coords``: (M, ndim) float32 radians on the device."""
        self._sync_stream()
        c = coords.to(torch.float32)
        cols = [c[:, d].contiguous() for d in range(self.ndim)]
        args = [ptr(x) for x in cols] + [None] * (3 - self.ndim)
        check(lib().hipnufft_setpts(self._h, c.shape[0], *args), "setpts")
        self.n_samples = int(c.shape[0])

    def setpts_sets(self, coords_list):
        """Several point sets at once: a list of (M_k, ndim) float32 radians
        tensors on the device (one trajectory each)."""
        self._sync_stream()
        Ms = [int(c.shape[0]) for c in coords_list]
        c = torch.cat([x.to(torch.float32).reshape(-1, self.ndim) for x in coords_list], 0)
        cols = [c[:, d].contiguous() for d in range(self.ndim)]
        args = [ptr(x) for x in cols] + [None] * (3 - self.ndim)
        arr = (_I64 * len(Ms))(*Ms)
        check(lib().hipnufft_setpts_sets(self._h, len(Ms), arr, *args), "setpts_sets")
        self.n_samples = int(c.shape[0])
        self.set_sizes = Ms

    def info(self):
        nf = (_I64 * 3)()
        w = _I()
        beta = ctypes.c_double()
        tile = (_I * 3)()
        M = _I64()
        check(lib().hipnufft_plan_info(self._h, nf, ctypes.byref(w), ctypes.byref(beta), tile,
                                       ctypes.byref(M)))
        return dict(nf=tuple(nf), w=w.value, beta=beta.value, tile=tuple(tile), M=M.value)

    def set_timing(self, enable=True):
        check(lib().hipnufft_set_timing(self._h, 1 if enable else 0))

    def get_timing(self):
        """{stage: (total_ms, launches)} since the last read (synchronises)."""
        n = len(STAGES)
        ms = (ctypes.c_double * n)()
        cnt = (_I64 * n)()
        check(lib().hipnufft_get_timing(self._h, ms, cnt, n))
        return {s: (ms[i], cnt[i]) for i, s in enumerate(STAGES)}

    def workspace_bytes(self):
        b = ctypes.c_size_t()
        check(lib().hipnufft_workspace_bytes(self._h, ctypes.byref(b)))
        return b.value

    def get_sort(self):
        """Permutation, bin offsets and spread lists (int32 device tensors)."""
        sizes = (_I64 * 3)()
        check(lib().hipnufft_get_sort(self._h, None, None, None, None, sizes))
        M, nbins, nlist = (int(v) for v in sizes)
        dev = torch.device("cuda")
        perm = torch.empty(max(M, 1), dtype=torch.int32, device=dev)
        bs = torch.empty(nbins + 1, dtype=torch.int32, device=dev)
        lst = torch.empty(max(nlist, 1), dtype=torch.int32, device=dev)
        ls = torch.empty(nbins + 1, dtype=torch.int32, device=dev)
        self._sync_stream()
        check(lib().hipnufft_get_sort(self._h, ptr(perm), ptr(bs), ptr(lst), ptr(ls), sizes))
        torch.cuda.synchronize()
        return perm[:M], bs, lst[:nlist], ls

    # --- transforms on contiguous complex64 device tensors ------------------
    def sense_op(self, img, smaps, ksp, B, C, scale):
        self._sync_stream()
        check(lib().hipnufft_sense_op(self._h, ptr(img), ptr(smaps), ptr(ksp), B, C, scale), "op")

    def sense_adj(self, ksp, smaps, density, img, B, C, scale):
        self._sync_stream()
        check(lib().hipnufft_sense_adj(self._h, ptr(ksp), ptr(smaps), ptr(density), ptr(img), B, C,
                                       scale), "adj_op")

    def sense_dc(self, img, smaps, obs, density, grad, B, C, scale):
        self._sync_stream()
        check(lib().hipnufft_sense_dc(self._h, ptr(img), ptr(smaps), ptr(obs), ptr(density),
                                      ptr(grad), B, C, scale), "data_consistency")

    def orc_op(self, img, smaps, cmap, bt, ksp, B, C, scale):
        """Off-resonance forward: cmap (L, *shape), bt (NK, L) complex64 on the device."""
        self._sync_stream()
        L, NK = int(cmap.shape[0]), int(bt.shape[0])
        if bt.shape[1] != L:
            raise ValueError("the temporal interpolators must be (NK, L)")
        check(lib().hipnufft_orc_op(self._h, ptr(img), ptr(smaps), ptr(cmap), ptr(bt), L, NK,
                                    ptr(ksp), B, C, scale), "orc op")

    def orc_adj(self, ksp, smaps, cmap, bt, density, img, B, C, scale):
        self._sync_stream()
        L, NK = int(cmap.shape[0]), int(bt.shape[0])
        if bt.shape[1] != L:
            raise ValueError("the temporal interpolators must be (NK, L)")
        check(lib().hipnufft_orc_adj(self._h, ptr(ksp), ptr(smaps), ptr(cmap), ptr(bt), L, NK,
                                     ptr(density), ptr(img), B, C, scale), "orc adj_op")

    def sense_grad_traj(self, img, smaps, dy, B, C, scale, isign, conj_smaps, gtraj):
        self._sync_stream()
        check(lib().hipnufft_sense_grad_traj(self._h, ptr(img), ptr(smaps), ptr(dy), B, C, scale,
                                             isign, 1 if conj_smaps else 0, ptr(gtraj)),
              "grad_traj")

    def execute(self, inp, out):
        self._sync_stream()
        check(lib().hipnufft_execute(self._h, ptr(inp), ptr(out)), "execute")

    def execute_adjoint(self, inp, out):
        self._sync_stream()
        check(lib().hipnufft_execute_adjoint(self._h, ptr(inp), ptr(out)), "execute_adjoint")

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.hipnufft_plan_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


WINDOWS = {"ellipse": 0, "rect": 1, "hann": 2, "hanning": 2, "hamming": 3}


def kspace_window(traj, threshold, kind, ksp=None, weights=True):
    """Window of the k-space centre on the device: (weights (M,) or None,
    windowed copy of ksp (C, M) or None).  traj (M, d) float32 in cycles."""
    t = traj.to(torch.float32).contiguous()
    M, d = t.shape
    thr = (_F * d)(*[float(x) for x in threshold])
    w = torch.empty(M, dtype=torch.float32, device=t.device) if weights else None
    k = out = None
    C = 0
    if ksp is not None:
        k = ksp.reshape(-1, M).contiguous()
        C = k.shape[0]
        out = torch.empty_like(k)
    check(lib().hipnufft_kspace_window(ptr(t), M, d, thr, WINDOWS[kind], ptr(k), ptr(out), C,
                                       ptr(w), current_stream()), "window")
    return w, out


def rss_normalize(maps, eps=0.0):
    """maps (C, *shape) complex64 device tensor, divided in place by its
    root sum of squares over coils (+ eps)."""
    C = maps.shape[0]
    check(lib().hipnufft_rss_normalize(ptr(maps), C, maps[0].numel(), float(eps),
                                       current_stream()), "rss_normalize")
    return maps


def blur_magnitude(maps, sigma, truncate=4.0):
    """Gaussian blur of |maps| per image axis (sigma per axis), phase kept, in place."""
    C, shape = maps.shape[0], tuple(maps.shape[1:])
    work = torch.empty(2 * maps.numel(), dtype=torch.float32, device=maps.device)
    sh = (_I64 * len(shape))(*shape)
    sg = (ctypes.c_double * len(shape))(*[float(x) for x in sigma])
    check(lib().hipnufft_blur_magnitude(ptr(maps), C, len(shape), sh, sg, float(truncate),
                                        ptr(work), current_stream()), "blur_magnitude")
    return maps


class RawStack:
    """Owns one ``hipnufft_stack``: the z-axis passes of a stacked 2.5D
    transform between images [.][X][Y][Z] and planes [T][NZ][X][Y]."""

    def __init__(self, shape, z_index, stream=None):
        L = lib()
        X, Y, Z = (int(s) for s in shape)
        self.shape = (X, Y, Z)
        zi = np.ascontiguousarray(np.asarray(z_index, dtype=np.int32).reshape(-1))
        self.n_planes = int(zi.size)
        self.stream = current_stream() if stream is None else stream
        h = _P()
        check(L.hipnufft_stack_create(X, Y, Z, zi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                      self.n_planes, self.stream, ctypes.byref(h)),
              "stack_create")
        self._h = h

    def _sync_stream(self):
        s = current_stream()
        if s != self.stream:
            self.stream = s
            check(lib().hipnufft_stack_set_stream(self._h, s))

    def fwd(self, img, smaps, planes, C, t0, T, scale):
        self._sync_stream()
        check(lib().hipnufft_stack_fwd(self._h, ptr(img), ptr(smaps), int(C), int(t0), int(T),
                                       float(scale), ptr(planes)), "stack_fwd")

    def adj(self, planes, smaps, img, C, t0, T, scale):
        self._sync_stream()
        check(lib().hipnufft_stack_adj(self._h, ptr(planes), ptr(smaps), int(C), int(t0),
                                       int(T), float(scale), ptr(img)), "stack_adj")

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.hipnufft_stack_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RawGram:
    """Owns one ``hipnufft_gram`` (Toeplitz form of ``scale * A^H W A``)."""

    def __init__(self, shape, coords, weights=None, eps=1e-6, isign=-1, scale=1.0, n_trans=1,
                 stream=None):
        L = lib()
        self.shape = tuple(int(s) for s in shape)
        self.ndim = len(self.shape)
        self.stream = current_stream() if stream is None else stream
        c = coords.to(torch.float32)
        cols = [c[:, d].contiguous() for d in range(self.ndim)]
        args = [ptr(x) for x in cols] + [None] * (3 - self.ndim)
        w = None if weights is None else weights.to(torch.float32).contiguous()
        nm = (_I64 * self.ndim)(*self.shape)
        h = _P()
        check(L.hipnufft_gram_create(self.ndim, nm, c.shape[0], *args, ptr(w), float(eps),
                                     int(isign), float(scale), int(n_trans), self.stream,
                                     ctypes.byref(h)), "gram_create")
        self._h = h

    def _sync_stream(self):
        s = current_stream()
        if s != self.stream:
            self.stream = s
            check(lib().hipnufft_gram_set_stream(self._h, s))

    def apply(self, img, smaps, out, B, C):
        self._sync_stream()
        check(lib().hipnufft_gram_apply(self._h, ptr(img), ptr(smaps), ptr(out), B, C), "gram")

    def workspace_bytes(self):
        b = ctypes.c_size_t()
        check(lib().hipnufft_gram_workspace_bytes(self._h, ctypes.byref(b)))
        return b.value

    def close(self):
        h, self._h = getattr(self, "_h", None), None
        if h and _lib is not None:
            _lib.hipnufft_gram_destroy(h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


DOT_WS = 2048


def dot_c64(a, b, conj=True, out=None, ws=None):
    """FP64-accumulated ``sum(conj(a) * b)`` (or ``sum(a * b)``) of complex64
    device tensors.  Returns a (2,) float64 device tensor [re, im] without a
    host synchronisation; ``ws`` is an optional (DOT_WS,) float64 workspace.
    """
    a = a.resolve_conj().resolve_neg().contiguous()
    b = b.resolve_conj().resolve_neg().contiguous()
    if a.numel() != b.numel():
        raise ValueError("dot_c64: size mismatch")
    if out is None:
        out = torch.empty(2, dtype=torch.float64, device=a.device)
    check(lib().hipnufft_dot_c64(ptr(a), ptr(b), a.numel(), 1 if conj else 0, ptr(out), ptr(ws),
                                 current_stream()), "dot")
    return out


def dot_c64_rows(a, b, conj=True, out=None, ws=None):
    """Row-wise FP64 dots of (R, n) complex64 device tensors -> (R, 2) float64
    [re, im] per row (row r equals ``dot_c64(a[r], b[r])`` bit for bit);
    ``ws`` is an optional (R * DOT_WS,) float64 workspace."""
    a = a.resolve_conj().resolve_neg().contiguous()
    b = b.resolve_conj().resolve_neg().contiguous()
    if a.shape != b.shape or a.ndim != 2:
        raise ValueError("dot_c64_rows: need two (rows, n) tensors of the same shape")
    R, n = a.shape
    if out is None:
        out = torch.empty((R, 2), dtype=torch.float64, device=a.device)
    check(lib().hipnufft_dot_c64_rows(ptr(a), ptr(b), R, n, 1 if conj else 0, ptr(out), ptr(ws),
                                      current_stream()), "dot_rows")
    return out


NOALLOC_DOT_ROWS = 64


def dot_c64_rows_noalloc(a, b, conj=True, out=None):
    """``dot_c64_rows`` through the entry point that allocates nothing (its
    partial sums live in the library's static device array): at most
    NOALLOC_DOT_ROWS rows; concurrent calls on one GPU from different
    streams race.  ``out`` must be given by callers that cannot allocate."""
    a = a.resolve_conj().resolve_neg().contiguous()
    b = b.resolve_conj().resolve_neg().contiguous()
    if a.shape != b.shape or a.ndim != 2:
        raise ValueError("dot_c64_rows_noalloc: need two (rows, n) tensors of the same shape")
    R, n = a.shape
    if out is None:
        out = torch.empty((R, 2), dtype=torch.float64, device=a.device)
    check(lib().hipnufft_dot_c64_rows_noalloc(ptr(a), ptr(b), R, n, 1 if conj else 0, ptr(out),
                                              current_stream()), "dot_rows_noalloc")
    return out
