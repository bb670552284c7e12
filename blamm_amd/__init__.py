"""blamm_amd -- MI355X scan engine for `blamm scan`.

The product is native: ``lib/libblamm_hip.so`` (HIP kernels behind the C ABI of
``include/blamm_hip.h``), ``lib/libblamm_host.so`` (host restatements,
``include/blamm_host.h``) and ``bin/blamm`` (the `blamm dict|hist|scan`
command).  This module only locates and binds them with ctypes so tests and
benchmarks can drive the C ABI; it never scores anything itself and raises
when the native libraries are missing.
"""
from __future__ import annotations

import ctypes
import importlib.util
import os
from typing import Optional, Tuple

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_DIR = os.path.join(PKG_DIR, "lib")
BIN_DIR = os.path.join(PKG_DIR, "bin")
HIP_LIB = os.path.join(LIB_DIR, "libblamm_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libblamm_host.so")
BLAMM_BIN = os.path.join(BIN_DIR, "blamm")

# enum blamm_status / blamm_precision (include/blamm_hip.h)
BLAMM_OK = 0
BLAMM_ERR_ARG = -1
BLAMM_ERR_DEVICE = -2
BLAMM_ERR_NOMEM = -3
BLAMM_ERR_OVERFLOW = -4
BLAMM_ERR_STATE = -5
PREC_FP32 = 0
PREC_BF16X2 = 1
MAX_MOTIF_LEN = 32

HIT_DTYPE = np.dtype([("pos", "<u8"), ("col", "<u4"), ("score", "<f4")])


class BlammError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ScanStats(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint64), ("band_rescored", ctypes.c_uint64),
                ("hits", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
                ("filter_ms", ctypes.c_double)]


_hip = None
_host = None

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int)


def _need(path: str) -> None:
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run __graft_entry__.build() "
                           f"(make -C blamm_amd/csrc) first")


def _share_torch_hip_runtime() -> None:
    """A process must hold a single HIP runtime.  torch wheels bundle their own
    libamdhip64 (soname libamdhip64.so.7); importing torch first makes
    libblamm_hip.so bind to that copy instead of loading /opt/rocm's beside it,
    after which torch would find no GPU."""
    if importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401


def hip_lib() -> ctypes.CDLL:
    """The HIP engine (fails loudly when it is not built)."""
    global _hip
    if _hip is None:
        _need(HIP_LIB)
        _share_torch_hip_runtime()
        lib = ctypes.CDLL(HIP_LIB)
        vp = ctypes.c_void_p
        lib.blamm_hip_last_error.restype = ctypes.c_char_p
        lib.blamm_hip_abi_version.restype = ctypes.c_int
        lib.blamm_hip_device_count.argtypes = [_i32p]
        lib.blamm_hip_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        lib.blamm_hip_destroy.argtypes = [vp]
        lib.blamm_hip_set_library.argtypes = [vp, _f32p, ctypes.c_int, ctypes.c_int, _i32p, _f32p, ctypes.c_int]
        lib.blamm_hip_library_bounds.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        lib.blamm_hip_packed_words.argtypes = [ctypes.c_uint64]
        lib.blamm_hip_packed_words.restype = ctypes.c_uint64
        lib.blamm_hip_mask_words.argtypes = [ctypes.c_uint64]
        lib.blamm_hip_mask_words.restype = ctypes.c_uint64
        lib.blamm_hip_reserve.argtypes = [vp, ctypes.c_uint64]
        lib.blamm_hip_scan.argtypes = [vp, _u64p, _u64p, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                       vp, ctypes.c_uint64, _u64p]
        lib.blamm_hip_scan_device.argtypes = [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint64, ctypes.c_uint64,
                                              vp, ctypes.c_uint64, vp, ctypes.c_uint64, vp, vp]
        lib.blamm_hip_scan_finish.argtypes = [vp, _u64p]
        lib.blamm_hip_last_stats.argtypes = [vp, ctypes.POINTER(ScanStats)]
        _hip = lib
    return _hip


def host_lib() -> ctypes.CDLL:
    global _host
    if _host is None:
        _need(HOST_LIB)
        lib = ctypes.CDLL(HOST_LIB)
        lib.blamm_host_last_error.restype = ctypes.c_char_p
        lib.blamm_host_pwm.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.c_float, ctypes.c_int, _f32p]
        lib.blamm_host_max_score.argtypes = [_f32p, ctypes.c_int]
        lib.blamm_host_max_score.restype = ctypes.c_float
        lib.blamm_host_min_score.argtypes = [_f32p, ctypes.c_int]
        lib.blamm_host_min_score.restype = ctypes.c_float
        lib.blamm_host_theoretical_hist.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_uint64, ctypes.c_uint64,
                                                    _u64p, _f32p, _f32p]
        lib.blamm_host_pvalue_threshold.argtypes = [_f32p, ctypes.c_int, _u64p, ctypes.c_float, ctypes.c_uint64,
                                                    ctypes.c_uint64, ctypes.c_float, _f32p]
        lib.blamm_host_pack.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _u64p, _u64p, _i32p]
        _host = lib
    return _host


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def _check(rc: int) -> None:
    if rc != BLAMM_OK:
        raise BlammError(rc, hip_lib().blamm_hip_last_error().decode())


# ---- host restatements ------------------------------------------------------

def pwm(pfm: np.ndarray, bg_counts, pseudocount: float = 0.25, rc: bool = False) -> np.ndarray:
    """PFM (L x 4 counts) -> PWM (L x 4 float32) as `blamm` computes it."""
    pfm = np.ascontiguousarray(pfm, dtype=np.uint64)
    bg = np.ascontiguousarray(bg_counts, dtype=np.uint64)
    out = np.zeros((pfm.shape[0], 4), dtype=np.float32)
    if host_lib().blamm_host_pwm(_ptr(pfm, ctypes.c_uint64), pfm.shape[0], _ptr(bg, ctypes.c_uint64),
                                 ctypes.c_float(pseudocount), int(rc), _ptr(out, ctypes.c_float)) != 0:
        raise RuntimeError(host_lib().blamm_host_last_error().decode())
    return out


def pvalue_threshold(w: np.ndarray, bg_counts, pvalue: float, pseudocount: float = 0.25,
                     num_bins: int = 250, max_length: int = 10_000_000) -> float:
    """-pt threshold via the theoretical histogram `blamm hist` would write."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    bg = np.ascontiguousarray(bg_counts, dtype=np.uint64)
    out = ctypes.c_float()
    if host_lib().blamm_host_pvalue_threshold(_ptr(w, ctypes.c_float), w.shape[0], _ptr(bg, ctypes.c_uint64),
                                              ctypes.c_float(pseudocount), num_bins, max_length,
                                              ctypes.c_float(pvalue), ctypes.byref(out)) != 0:
        raise RuntimeError(host_lib().blamm_host_last_error().decode())
    return out.value


def pack(bases: bytes) -> Tuple[np.ndarray, Optional[np.ndarray]]:
    """Pack a filtered ACGTacgt byte string: (codes words, lowercase words or None)."""
    n = len(bases)
    codes = np.zeros((n + 31) // 32 + 1, dtype=np.uint64)
    lower = np.zeros((n + 63) // 64 + 1, dtype=np.uint64)
    has = ctypes.c_int()
    if host_lib().blamm_host_pack(bases, n, _ptr(codes, ctypes.c_uint64), _ptr(lower, ctypes.c_uint64),
                                  ctypes.byref(has)) != 0:
        raise RuntimeError(host_lib().blamm_host_last_error().decode())
    return codes[:(n + 31) // 32], (lower[:(n + 63) // 64] if has.value else None)


def pack_codes(codes: np.ndarray) -> np.ndarray:
    """Pack 2-bit codes (uint8 0..3) into 64-bit words, 32 bases per word."""
    n = codes.shape[0]
    pad = (-n) % 32
    c = np.concatenate([codes.astype(np.uint64), np.zeros(pad, np.uint64)]).reshape(-1, 32)
    shifts = (2 * np.arange(32, dtype=np.uint64))
    return np.bitwise_or.reduce(c << shifts, axis=1)


def build_pattern_matrix(pwms) -> Tuple[np.ndarray, np.ndarray, int]:
    """Stack per-column PWMs (L_c x 4) into the column-major P of the C ABI."""
    lmax = max(w.shape[0] for w in pwms)
    ldp = 4 * lmax
    P = np.zeros((len(pwms), ldp), dtype=np.float32)
    for c, w in enumerate(pwms):
        P[c, :4 * w.shape[0]] = np.asarray(w, dtype=np.float32).reshape(-1)
    lengths = np.array([w.shape[0] for w in pwms], dtype=np.int32)
    return P, lengths, ldp


# ---- the HIP engine ----------------------------------------------------------

class HipScanner:
    """One engine context on one device (thread-compatible)."""

    def __init__(self, device: int = 0):
        self.lib = hip_lib()
        self.ctx = ctypes.c_void_p()
        _check(self.lib.blamm_hip_create(device, ctypes.byref(self.ctx)))

    def close(self) -> None:
        if self.ctx:
            self.lib.blamm_hip_destroy(self.ctx)
            self.ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_library(self, P: np.ndarray, lengths: np.ndarray, thresholds: np.ndarray,
                    precision: int = PREC_BF16X2) -> None:
        P = np.ascontiguousarray(P, dtype=np.float32)
        lengths = np.ascontiguousarray(lengths, dtype=np.int32)
        thresholds = np.ascontiguousarray(thresholds, dtype=np.float32)
        _check(self.lib.blamm_hip_set_library(self.ctx, _ptr(P, ctypes.c_float), P.shape[1], P.shape[0],
                                              _ptr(lengths, ctypes.c_int), _ptr(thresholds, ctypes.c_float),
                                              precision))

    def bounds(self) -> Tuple[float, float]:
        band, err = ctypes.c_double(), ctypes.c_double()
        _check(self.lib.blamm_hip_library_bounds(self.ctx, ctypes.byref(band), ctypes.byref(err)))
        return band.value, err.value

    def scan(self, codes: np.ndarray, nbases: int, run_starts=None, lower: Optional[np.ndarray] = None,
             cap: int = 1 << 20) -> np.ndarray:
        """Scan a packed stream held in host memory; returns the hit records."""
        pw = int(self.lib.blamm_hip_packed_words(nbases))
        cbuf = np.zeros(pw, dtype=np.uint64)
        cbuf[:min(len(codes), pw)] = codes[:pw]
        mbuf = None
        if lower is not None:
            mw = int(self.lib.blamm_hip_mask_words(nbases))
            mbuf = np.zeros(mw, dtype=np.uint64)
            mbuf[:min(len(lower), mw)] = lower[:mw]
        runs = np.ascontiguousarray(run_starts if run_starts is not None else [], dtype=np.uint64)
        while True:
            hits = np.zeros(max(cap, 1), dtype=HIT_DTYPE)
            n = ctypes.c_uint64()
            rc = self.lib.blamm_hip_scan(self.ctx, _ptr(cbuf, ctypes.c_uint64),
                                         _ptr(mbuf, ctypes.c_uint64) if mbuf is not None else None,
                                         nbases, _ptr(runs, ctypes.c_uint64) if len(runs) else None, len(runs),
                                         hits.ctypes.data, cap, ctypes.byref(n))
            if rc == BLAMM_ERR_OVERFLOW and n.value > cap:
                cap = int(n.value)
                continue
            _check(rc)
            return hits[:n.value].copy()

    def stats(self) -> ScanStats:
        st = ScanStats()
        _check(self.lib.blamm_hip_last_stats(self.ctx, ctypes.byref(st)))
        return st

    # device-resident path (torch tensors as plain device memory)
    def scan_device(self, d_codes, d_lower, nbases: int, d_runs, nruns: int, d_hits, cap: int, d_nhits,
                    stream_ptr: int = 0) -> None:
        _check(self.lib.blamm_hip_scan_device(
            self.ctx, d_codes.data_ptr(), d_codes.numel(),
            d_lower.data_ptr() if d_lower is not None else None, d_lower.numel() if d_lower is not None else 0,
            nbases, d_runs.data_ptr() if nruns else None, nruns, d_hits.data_ptr(), cap,
            d_nhits.data_ptr() if d_nhits is not None else None, stream_ptr or None))

    def scan_finish(self) -> int:
        n = ctypes.c_uint64()
        _check(self.lib.blamm_hip_scan_finish(self.ctx, ctypes.byref(n)))
        return int(n.value)

    def reserve(self, n: int) -> None:
        _check(self.lib.blamm_hip_reserve(self.ctx, n))
