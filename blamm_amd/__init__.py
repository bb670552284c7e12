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
HIP_LIB = os.environ.get("BLAMM_HIP_LIB") or os.path.join(LIB_DIR, "libblamm_hip.so")
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
PREC_BF16X2_LO = 2
MAX_MOTIF_LEN = 64

HIT_DTYPE = np.dtype([("pos", "<u8"), ("col", "<u4"), ("score", "<f4")])


class BlammError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ScanStats(ctypes.Structure):
    _fields_ = [("candidates", ctypes.c_uint64), ("band_rescored", ctypes.c_uint64),
                ("hits", ctypes.c_uint64), ("kernel_ms", ctypes.c_double),
                ("filter_ms", ctypes.c_double)]


class Tuning(ctypes.Structure):
    """struct blamm_tuning: window tile, motif tile, K-padding."""
    _fields_ = [("chunk_tiles", ctypes.c_uint32), ("blocks_per_wave", ctypes.c_uint32),
                ("kpad_positions", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {"chunk_tiles": self.chunk_tiles, "blocks_per_wave": self.blocks_per_wave,
                "kpad_positions": self.kpad_positions}


class Request(ctypes.Structure):
    """struct blamm_request."""
    _fields_ = [("packed", ctypes.c_void_p), ("lcmask", ctypes.c_void_p), ("nbases", ctypes.c_uint64),
                ("run_starts", ctypes.c_void_p), ("nruns", ctypes.c_uint64), ("thresholds", ctypes.c_void_p),
                ("expected_hits", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_uint64), ("slots", ctypes.c_uint32),
                ("max_batch_requests", ctypes.c_uint32), ("max_batch_bases", ctypes.c_uint64)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("batch_requests", ctypes.c_uint32), ("nhits", ctypes.c_uint64),
                ("t_submit_ns", ctypes.c_uint64), ("t_start_ns", ctypes.c_uint64), ("t_done_ns", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("hits", ctypes.c_uint64), ("replays", ctypes.c_uint64), ("device_ms", ctypes.c_double),
                ("host_submit_ms", ctypes.c_double), ("host_wait_ms", ctypes.c_double)]


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
        lib.blamm_hip_scan_pinned.argtypes = [vp, _u64p, _u64p, ctypes.c_uint64, _u64p, ctypes.c_uint64,
                                              ctypes.POINTER(vp), _u64p]
        lib.blamm_hip_last_stats.argtypes = [vp, ctypes.POINTER(ScanStats)]
        lib.blamm_hip_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
        lib.blamm_hip_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
        lib.blamm_hip_autotune.argtypes = [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint64, ctypes.c_uint64, vp,
                                           ctypes.c_uint64, ctypes.c_char_p, ctypes.POINTER(Tuning),
                                           ctypes.POINTER(ctypes.c_double)]
        lib.blamm_hip_batch_create.argtypes = [vp, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(vp)]
        lib.blamm_hip_batch_destroy.argtypes = [vp]
        lib.blamm_hip_batch_device_bytes.argtypes = [vp]
        lib.blamm_hip_batch_device_bytes.restype = ctypes.c_uint64
        lib.blamm_hip_batch_submit.argtypes = [vp, ctypes.POINTER(Request), ctypes.c_uint32]
        lib.blamm_hip_batch_query.argtypes = [vp]
        lib.blamm_hip_batch_wait.argtypes = [vp, _u64p]
        lib.blamm_hip_batch_hits.argtypes = [vp, ctypes.c_uint32, ctypes.POINTER(vp), _u64p]
        lib.blamm_hip_batch_stats.argtypes = [vp, ctypes.POINTER(ScanStats)]
        lib.blamm_hip_server_create.argtypes = [vp, ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]
        lib.blamm_hip_server_submit.argtypes = [vp, ctypes.POINTER(Request), _u64p]
        lib.blamm_hip_server_wait.argtypes = [vp, ctypes.c_uint64, ctypes.c_int64]
        lib.blamm_hip_server_result.argtypes = [vp, ctypes.c_uint64, vp, ctypes.c_uint64,
                                                ctypes.POINTER(RequestInfo)]
        lib.blamm_hip_server_stats.argtypes = [vp, ctypes.POINTER(ServerStats)]
        lib.blamm_hip_server_release.argtypes = [vp, ctypes.c_uint64, ctypes.POINTER(RequestInfo)]
        lib.blamm_hip_server_destroy.argtypes = [vp]
        lib.blamm_hip_hist.argtypes = [vp, _u64p, _u64p, ctypes.c_uint64, _u64p, ctypes.c_uint64, ctypes.c_uint64,
                                       _f32p, _f32p, ctypes.c_uint32, _u64p]
        lib.blamm_hip_hist_device.argtypes = [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint64, ctypes.c_uint64, vp,
                                              ctypes.c_uint64, ctypes.c_uint64, _f32p, _f32p, ctypes.c_uint32, _u64p]
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


def max_score(w: np.ndarray) -> float:
    """Motif::getMaxScore of a PWM (L x 4 float32), as `blamm` computes it."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    return float(host_lib().blamm_host_max_score(_ptr(w, ctypes.c_float), w.shape[0]))


def min_score(w: np.ndarray) -> float:
    """Motif::getMinScore of a PWM (L x 4 float32), as `blamm` computes it."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    return float(host_lib().blamm_host_min_score(_ptr(w, ctypes.c_float), w.shape[0]))


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

    def scan_pinned(self, codes: np.ndarray, nbases: int, run_starts=None,
                    lower: Optional[np.ndarray] = None) -> np.ndarray:
        """Scan a host stream; the engine keeps the hits in its pinned buffer
        (returned here as a copy)."""
        pw = int(self.lib.blamm_hip_packed_words(nbases))
        cbuf = np.zeros(pw, dtype=np.uint64)
        cbuf[:min(len(codes), pw)] = codes[:pw]
        mbuf = None
        if lower is not None:
            mw = int(self.lib.blamm_hip_mask_words(nbases))
            mbuf = np.zeros(mw, dtype=np.uint64)
            mbuf[:min(len(lower), mw)] = lower[:mw]
        runs = np.ascontiguousarray(run_starts if run_starts is not None else [], dtype=np.uint64)
        p = ctypes.c_void_p()
        n = ctypes.c_uint64()
        _check(self.lib.blamm_hip_scan_pinned(self.ctx, _ptr(cbuf, ctypes.c_uint64),
                                              _ptr(mbuf, ctypes.c_uint64) if mbuf is not None else None,
                                              nbases, _ptr(runs, ctypes.c_uint64) if len(runs) else None,
                                              len(runs), ctypes.byref(p), ctypes.byref(n)))
        return _hits_array(p.value or 0, int(n.value))

    def hist(self, codes: np.ndarray, nbases: int, min_scores, max_scores, num_bins: int = 250, run_starts=None,
             lower: Optional[np.ndarray] = None, qend: Optional[int] = None) -> np.ndarray:
        """Empirical histograms (hist -e) of a host stream for every library
        column: (ncols, num_bins) uint64 counts."""
        pw = int(self.lib.blamm_hip_packed_words(nbases))
        cbuf = np.zeros(pw, dtype=np.uint64)
        cbuf[:min(len(codes), pw)] = codes[:pw]
        mbuf = None
        if lower is not None:
            mw = int(self.lib.blamm_hip_mask_words(nbases))
            mbuf = np.zeros(mw, dtype=np.uint64)
            mbuf[:min(len(lower), mw)] = lower[:mw]
        runs = np.ascontiguousarray(run_starts if run_starts is not None else [], dtype=np.uint64)
        mn = np.ascontiguousarray(min_scores, np.float32)
        mx = np.ascontiguousarray(max_scores, np.float32)
        counts = np.zeros((len(mn), num_bins), np.uint64)
        _check(self.lib.blamm_hip_hist(self.ctx, _ptr(cbuf, ctypes.c_uint64),
                                       _ptr(mbuf, ctypes.c_uint64) if mbuf is not None else None, nbases,
                                       _ptr(runs, ctypes.c_uint64) if len(runs) else None, len(runs),
                                       nbases if qend is None else qend, _ptr(mn, ctypes.c_float),
                                       _ptr(mx, ctypes.c_float), num_bins, _ptr(counts, ctypes.c_uint64)))
        return counts

    def hist_device(self, d_codes, d_lower, nbases: int, d_runs, nruns: int, min_scores, max_scores,
                    num_bins: int = 250, qend: Optional[int] = None) -> np.ndarray:
        """hist() on device-resident buffers (torch tensors)."""
        mn = np.ascontiguousarray(min_scores, np.float32)
        mx = np.ascontiguousarray(max_scores, np.float32)
        counts = np.zeros((len(mn), num_bins), np.uint64)
        _check(self.lib.blamm_hip_hist_device(
            self.ctx, d_codes.data_ptr(), d_codes.numel(),
            d_lower.data_ptr() if d_lower is not None else None, d_lower.numel() if d_lower is not None else 0,
            nbases, d_runs.data_ptr() if nruns else None, nruns, nbases if qend is None else qend,
            _ptr(mn, ctypes.c_float), _ptr(mx, ctypes.c_float), num_bins, _ptr(counts, ctypes.c_uint64)))
        return counts

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

    # tile shapes
    def tuning(self) -> Tuning:
        t = Tuning()
        _check(self.lib.blamm_hip_get_tuning(self.ctx, ctypes.byref(t)))
        return t

    def set_tuning(self, chunk_tiles: int, blocks_per_wave: int, kpad_positions: int) -> None:
        t = Tuning(chunk_tiles, blocks_per_wave, kpad_positions)
        _check(self.lib.blamm_hip_set_tuning(self.ctx, ctypes.byref(t)))

    def autotune(self, d_codes, d_lower, nbases: int, d_runs, nruns: int,
                 cache_path: Optional[str] = None) -> Tuple[Tuning, float]:
        """Pick the tile shapes on a device-resident sample (torch tensors)."""
        t = Tuning()
        ms = ctypes.c_double()
        _check(self.lib.blamm_hip_autotune(
            self.ctx, d_codes.data_ptr(), d_codes.numel(),
            d_lower.data_ptr() if d_lower is not None else None, d_lower.numel() if d_lower is not None else 0,
            nbases, d_runs.data_ptr() if nruns else None, nruns,
            cache_path.encode() if cache_path else None, ctypes.byref(t), ctypes.byref(ms)))
        return t, ms.value


class ScanRequest:
    """One stream and its thresholds, kept alive while the engine uses it."""

    def __init__(self, codes: np.ndarray, nbases: int, run_starts=None, lower: Optional[np.ndarray] = None,
                 thresholds: Optional[np.ndarray] = None, expected_hits: int = 0):
        nw = (nbases + 31) // 32
        self.codes = np.zeros(max(nw, 1), dtype=np.uint64)
        self.codes[:min(nw, len(codes))] = codes[:nw]
        self.lower = None
        if lower is not None:
            mw = (nbases + 63) // 64
            self.lower = np.zeros(max(mw, 1), dtype=np.uint64)
            self.lower[:min(mw, len(lower))] = lower[:mw]
        self.runs = np.ascontiguousarray(run_starts if run_starts is not None else [], dtype=np.uint64)
        self.thresholds = None if thresholds is None else np.ascontiguousarray(thresholds, dtype=np.float32)
        self.nbases = int(nbases)
        self.c = Request(self.codes.ctypes.data, self.lower.ctypes.data if self.lower is not None else None,
                         self.nbases, self.runs.ctypes.data if len(self.runs) else None, len(self.runs),
                         self.thresholds.ctypes.data if self.thresholds is not None else None, int(expected_hits))


def _hits_array(ptr: int, n: int) -> np.ndarray:
    if n == 0:
        return np.zeros(0, dtype=HIT_DTYPE)
    buf = (ctypes.c_char * (n * HIT_DTYPE.itemsize)).from_address(ptr)
    return np.frombuffer(buf, dtype=HIT_DTYPE).copy()


class Batch:
    """A batch slot: many requests, one captured launch sequence."""

    def __init__(self, scanner: HipScanner, max_bases: int = 1 << 26, max_requests: int = 256):
        self.scanner = scanner
        self.lib = scanner.lib
        self.b = ctypes.c_void_p()
        _check(self.lib.blamm_hip_batch_create(scanner.ctx, max_bases, max_requests, ctypes.byref(self.b)))
        self.nreq = 0

    def close(self) -> None:
        if self.b:
            self.lib.blamm_hip_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, reqs) -> None:
        arr = (Request * len(reqs))(*[r.c for r in reqs])
        _check(self.lib.blamm_hip_batch_submit(self.b, arr, len(reqs)))
        self.nreq = len(reqs)

    def done(self) -> bool:
        rc = self.lib.blamm_hip_batch_query(self.b)
        if rc < 0:
            _check(rc)
        return rc == 0

    def wait(self) -> int:
        n = ctypes.c_uint64()
        _check(self.lib.blamm_hip_batch_wait(self.b, ctypes.byref(n)))
        return int(n.value)

    def hits(self, req: int) -> np.ndarray:
        p = ctypes.c_void_p()
        n = ctypes.c_uint64()
        _check(self.lib.blamm_hip_batch_hits(self.b, req, ctypes.byref(p), ctypes.byref(n)))
        return _hits_array(p.value or 0, int(n.value))

    def stats(self) -> ScanStats:
        st = ScanStats()
        _check(self.lib.blamm_hip_batch_stats(self.b, ctypes.byref(st)))
        return st

    def device_bytes(self) -> int:
        return int(self.lib.blamm_hip_batch_device_bytes(self.b))


class Server:
    """Continuous-batching server over the scanner's resident library."""

    def __init__(self, scanner: HipScanner, hbm_budget_bytes: int = 0, slots: int = 0, max_batch_requests: int = 0,
                 max_batch_bases: int = 0):
        self.scanner = scanner
        self.lib = scanner.lib
        self.s = ctypes.c_void_p()
        cfg = ServerConfig(hbm_budget_bytes, slots, max_batch_requests, max_batch_bases)
        _check(self.lib.blamm_hip_server_create(scanner.ctx, ctypes.byref(cfg), ctypes.byref(self.s)))
        self._keep = {}

    def submit(self, req: ScanRequest) -> int:
        i = ctypes.c_uint64()
        _check(self.lib.blamm_hip_server_submit(self.s, ctypes.byref(req.c), ctypes.byref(i)))
        self._keep[i.value] = req
        return int(i.value)

    def wait(self, rid: int, timeout_ms: int = -1) -> bool:
        rc = self.lib.blamm_hip_server_wait(self.s, rid, timeout_ms)
        if rc < 0:
            _check(rc)
        return rc == 0

    def result(self, rid: int) -> Tuple[np.ndarray, RequestInfo]:
        info = RequestInfo()
        rc = self.lib.blamm_hip_server_result(self.s, rid, None, 0, ctypes.byref(info))
        if rc == BLAMM_ERR_OVERFLOW:
            hits = np.zeros(info.nhits, dtype=HIT_DTYPE)
            rc = self.lib.blamm_hip_server_result(self.s, rid, hits.ctypes.data, info.nhits, ctypes.byref(info))
        else:
            hits = np.zeros(0, dtype=HIT_DTYPE)
        self._keep.pop(rid, None)
        _check(rc)
        return hits, info

    def release(self, rid: int) -> RequestInfo:
        """Timing and hit count of a finished request; its hits are dropped."""
        info = RequestInfo()
        rc = self.lib.blamm_hip_server_release(self.s, rid, ctypes.byref(info))
        self._keep.pop(rid, None)
        _check(rc)
        return info

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(self.lib.blamm_hip_server_stats(self.s, ctypes.byref(st)))
        return st

    def close(self) -> None:
        if self.s:
            self.lib.blamm_hip_server_destroy(self.s)
            self.s = ctypes.c_void_p()
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
