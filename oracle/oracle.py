"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker / the timed CPU baseline -- never as part
of the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import Optional, Tuple

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
BIN = os.path.join(ORACLE_DIR, "_build", "blamm_oracle")

HIT_DTYPE = np.dtype([("pos", "<u8"), ("col", "<u4"), ("pad", "<u4"), ("score64", "<f8"),
                      ("score32", "<f4"), ("pad2", "<f4")])

_u64p = ctypes.POINTER(ctypes.c_uint64)
_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int)


def build() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def build_variant(define: str, out_dir: str) -> str:
    """Compile the oracle with -D<define> (a planted mistake) into out_dir."""
    path = os.path.join(out_dir, f"liboracle_{define}.so")
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
                    f"-D{define}", "-o", path, os.path.join(ORACLE_DIR, "blamm_oracle.c"),
                    "-lm", "-lpthread", "-ldl"], check=True)
    return path


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


class Oracle:
    """One loaded build of the oracle library."""

    def __init__(self, path: str = LIB):
        if path == LIB and not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(path)
        L.oracle_log2f.argtypes = [ctypes.c_float]
        L.oracle_log2f.restype = ctypes.c_float
        L.oracle_pfm2pwm.argtypes = [_u64p, ctypes.c_int, _u64p, ctypes.c_float, ctypes.c_int, _f32p]
        L.oracle_max_score.argtypes = [_f32p, ctypes.c_int]
        L.oracle_max_score.restype = ctypes.c_float
        L.oracle_min_score.argtypes = [_f32p, ctypes.c_int]
        L.oracle_min_score.restype = ctypes.c_float
        L.oracle_score_cutoff.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_float, ctypes.c_float, ctypes.c_float]
        L.oracle_score_cutoff.restype = ctypes.c_float
        L.oracle_revcompl_pfm.argtypes = [_u64p, ctypes.c_int]
        L.oracle_filter_fasta.argtypes = [ctypes.POINTER(ctypes.c_char_p), ctypes.c_int, ctypes.c_uint64,
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                          ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                          _u64p, _u64p, ctypes.c_char_p, ctypes.c_size_t]
        L.oracle_filter_fasta.restype = ctypes.c_int64
        L.oracle_scan_fp64.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _u64p, ctypes.c_uint64, _f32p,
                                       ctypes.c_int, _i32p, _f32p, ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_scan_fp64.restype = ctypes.c_int64
        L.oracle_scan_sgemm.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _u64p, ctypes.c_uint64, _f32p,
                                        ctypes.c_int, _i32p, _f32p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_scan_sgemm.restype = ctypes.c_int64
        L.oracle_blas_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_size_t]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_theoretical_hist.argtypes = [_f32p, ctypes.c_int, _f32p, ctypes.c_size_t, ctypes.c_size_t,
                                              _u64p, _f32p, _f32p]
        L.oracle_empirical_hist.argtypes = [ctypes.c_char_p, ctypes.c_uint64, _u64p, ctypes.c_uint64, _f32p,
                                            ctypes.c_int, _i32p, _f32p, _f32p, ctypes.c_int, ctypes.c_size_t,
                                            ctypes.c_uint64, ctypes.c_int, _u64p]
        L.oracle_load_motifs.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_size_t),
                                         ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
                                         ctypes.POINTER(ctypes.c_void_p)]
        L.oracle_is_permutation.argtypes = [ctypes.c_char_p]
        L.oracle_base_name.argtypes = [ctypes.c_char_p]
        L.oracle_base_name.restype = ctypes.c_void_p
        L.oracle_main.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_char_p)]
        self.L = L

    def load_motifs(self, path: str, perms: bool = True):
        """[(name, L x 4 counts)] as MotifContainer::load leaves them (sorted
        by length; permutations kept only with perms)."""
        n = ctypes.c_size_t()
        names, lens, pfms = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        if self.L.oracle_load_motifs(path.encode(), int(perms), ctypes.byref(n), ctypes.byref(names),
                                     ctypes.byref(lens), ctypes.byref(pfms)) != 0:
            raise RuntimeError(f"cannot open {path}")
        k = n.value
        name_ptrs = ctypes.cast(names, ctypes.POINTER(ctypes.c_void_p))
        len_arr = ctypes.cast(lens, ctypes.POINTER(ctypes.c_int))
        out, at = [], 0
        pf = ctypes.cast(pfms, ctypes.POINTER(ctypes.c_uint64))
        for i in range(k):
            nm = ctypes.string_at(name_ptrs[i]).decode("latin-1")
            L = len_arr[i]
            counts = np.array([pf[4 * at + j] for j in range(4 * L)], np.uint64).reshape(L, 4)
            at += L
            out.append((nm, counts))
            self.L.oracle_free(name_ptrs[i])
        for ptr in (names, lens, pfms):
            self.L.oracle_free(ptr)
        return out

    def is_permutation(self, name: str) -> bool:
        return bool(self.L.oracle_is_permutation(name.encode()))

    def base_name(self, name: str) -> str:
        ptr = self.L.oracle_base_name(name.encode())
        out = ctypes.string_at(ptr).decode()
        self.L.oracle_free(ptr)
        return out

    def main(self, args) -> int:
        """The oracle's command line (dict / hist / scan) in this process; it
        exits the process on input errors, so give it valid inputs only."""
        argv = (ctypes.c_char_p * (len(args) + 1))(b"blamm_oracle", *[a.encode() for a in args])
        return self.L.oracle_main(len(args) + 1, argv)

    def log2f(self, x: float) -> float:
        return self.L.oracle_log2f(x)

    def pfm2pwm(self, pfm: np.ndarray, bg, pseudocount: float = 0.25, rc: bool = False) -> np.ndarray:
        pfm = np.ascontiguousarray(pfm, dtype=np.uint64)
        bg = np.ascontiguousarray(bg, dtype=np.uint64)
        out = np.zeros((pfm.shape[0], 4), np.float32)
        self.L.oracle_pfm2pwm(_p(pfm, ctypes.c_uint64), pfm.shape[0], _p(bg, ctypes.c_uint64),
                              ctypes.c_float(pseudocount), int(rc), _p(out, ctypes.c_float))
        return out

    def revcompl_pfm(self, pfm: np.ndarray) -> np.ndarray:
        p = np.ascontiguousarray(pfm, dtype=np.uint64).copy()
        self.L.oracle_revcompl_pfm(_p(p, ctypes.c_uint64), p.shape[0])
        return p

    def max_score(self, w: np.ndarray) -> float:
        w = np.ascontiguousarray(w, np.float32)
        return self.L.oracle_max_score(_p(w, ctypes.c_float), w.shape[0])

    def min_score(self, w: np.ndarray) -> float:
        w = np.ascontiguousarray(w, np.float32)
        return self.L.oracle_min_score(_p(w, ctypes.c_float), w.shape[0])

    def score_cutoff(self, counts, min_score_: float, max_score_: float, pvalue: float) -> float:
        c = np.ascontiguousarray(counts, np.uint64)
        return self.L.oracle_score_cutoff(_p(c, ctypes.c_uint64), c.shape[0], ctypes.c_float(min_score_),
                                          ctypes.c_float(max_score_), ctypes.c_float(pvalue))

    def theoretical_hist(self, w: np.ndarray, bg_prob, num_bins: int = 250, max_length: int = 10_000_000):
        w = np.ascontiguousarray(w, np.float32)
        bg = np.ascontiguousarray(bg_prob, np.float32)
        counts = np.zeros(num_bins, np.uint64)
        mn, mx = ctypes.c_float(), ctypes.c_float()
        self.L.oracle_theoretical_hist(_p(w, ctypes.c_float), w.shape[0], _p(bg, ctypes.c_float), num_bins,
                                       max_length, _p(counts, ctypes.c_uint64), ctypes.byref(mn), ctypes.byref(mx))
        return counts, mn.value, mx.value

    def empirical_hist(self, stream: bytes, run_starts, pwms, min_scores, max_scores, num_bins: int = 250,
                       qend: Optional[int] = None, threads: int = 8) -> np.ndarray:
        """Empirical histograms (hist -e) of a filtered stream: (ncols, num_bins) uint64 counts."""
        P, lens, lmax = _stack(pwms)
        mn = np.ascontiguousarray(min_scores, np.float32)
        mx = np.ascontiguousarray(max_scores, np.float32)
        rs = np.ascontiguousarray(run_starts if run_starts is not None and len(run_starts) else [0], np.uint64)
        counts = np.zeros((len(pwms), num_bins), np.uint64)
        rc = self.L.oracle_empirical_hist(stream, len(stream), _p(rs, ctypes.c_uint64), len(rs),
                                          _p(P, ctypes.c_float), lmax, _p(lens, ctypes.c_int),
                                          _p(mn, ctypes.c_float), _p(mx, ctypes.c_float), len(pwms), num_bins,
                                          len(stream) if qend is None else qend, threads,
                                          _p(counts, ctypes.c_uint64))
        if rc != 0:
            raise RuntimeError("oracle empirical histogram failed")
        return counts

    def filter_fasta(self, files, max_len: int = (1 << 64) - 1):
        """(stream bytes, run starts, run records, run positions, record count)."""
        arr = (ctypes.c_char_p * len(files))(*[f.encode() for f in files])
        sp, rs, rq, rp = (ctypes.c_void_p() for _ in range(4))
        nr, ns = ctypes.c_uint64(), ctypes.c_uint64()
        err = ctypes.create_string_buffer(1024)
        n = self.L.oracle_filter_fasta(arr, len(files), max_len, ctypes.byref(sp), ctypes.byref(rs),
                                       ctypes.byref(rq), ctypes.byref(rp), ctypes.byref(nr), ctypes.byref(ns),
                                       err, 1024)
        if n < 0:
            raise RuntimeError(err.value.decode())

        def take(ptr, count, dtype, size):
            out = np.frombuffer(bytes((ctypes.c_char * (count * size)).from_address(ptr.value)), dtype=dtype).copy() \
                if count else np.zeros(0, dtype)
            self.L.oracle_free(ptr)
            return out
        stream = take(sp, n, np.uint8, 1).tobytes()
        return (stream, take(rs, nr.value, np.uint64, 8), take(rq, nr.value, np.uint64, 8),
                take(rp, nr.value, np.uint64, 8), int(ns.value))

    def _hits(self, ptr: ctypes.c_void_p, n: int) -> np.ndarray:
        if n <= 0:
            if ptr:
                self.L.oracle_free(ptr)
            return np.zeros(0, HIT_DTYPE)
        buf = (ctypes.c_char * (n * HIT_DTYPE.itemsize)).from_address(ptr.value)
        out = np.frombuffer(bytes(buf), dtype=HIT_DTYPE).copy()
        self.L.oracle_free(ptr)
        return out

    def scan_fp64(self, stream: bytes, run_starts, pwms, thresholds, threads: int = 8) -> np.ndarray:
        """Brute-force FP64 scan of a filtered stream (bytes ACGTacgt)."""
        P, lens, lmax = _stack(pwms)
        thr = np.ascontiguousarray(thresholds, np.float32)
        rs = np.ascontiguousarray(run_starts if run_starts is not None and len(run_starts) else [0], np.uint64)
        out = ctypes.c_void_p()
        n = self.L.oracle_scan_fp64(stream, len(stream), _p(rs, ctypes.c_uint64), len(rs), _p(P, ctypes.c_float),
                                    lmax, _p(lens, ctypes.c_int), _p(thr, ctypes.c_float), len(pwms), threads,
                                    ctypes.byref(out))
        if n < 0:
            raise RuntimeError("oracle scan failed")
        return self._hits(out, n)

    def blas_open(self, path: str = "") -> None:
        err = ctypes.create_string_buffer(1024)
        if self.L.oracle_blas_open(path.encode(), err, 1024) != 0:
            raise RuntimeError(err.value.decode())

    def scan_sgemm(self, stream: bytes, run_starts, pwms, thresholds, h: int = 1000, w: int = 250,
                   min_zero_area: int = 4096, threads: int = 8) -> np.ndarray:
        """The reference's sgemm scan restated (columns sorted by length)."""
        self.blas_open()
        P, lens, lmax = _stack(pwms)
        thr = np.ascontiguousarray(thresholds, np.float32)
        rs = np.ascontiguousarray(run_starts if run_starts is not None and len(run_starts) else [0], np.uint64)
        out = ctypes.c_void_p()
        n = self.L.oracle_scan_sgemm(stream, len(stream), _p(rs, ctypes.c_uint64), len(rs), _p(P, ctypes.c_float),
                                     lmax, _p(lens, ctypes.c_int), _p(thr, ctypes.c_float), len(pwms), h, w,
                                     min_zero_area, threads, ctypes.byref(out))
        if n < 0:
            raise RuntimeError(f"oracle sgemm scan failed ({n})")
        return self._hits(out, n)


def _stack(pwms):
    lmax = max(w.shape[0] for w in pwms)
    P = np.zeros((len(pwms), lmax, 4), np.float32)
    for c, w in enumerate(pwms):
        P[c, :w.shape[0]] = w
    return P, np.array([w.shape[0] for w in pwms], np.int32), lmax


_default = None


def default() -> Oracle:
    global _default
    if _default is None:
        _default = Oracle(LIB)
    return _default


def lib() -> ctypes.CDLL:
    return default().L


def log2f(x):
    return default().log2f(x)


def pfm2pwm(*a, **k):
    return default().pfm2pwm(*a, **k)


def max_score(w):
    return default().max_score(w)


def min_score(w):
    return default().min_score(w)


def score_cutoff(*a, **k):
    return default().score_cutoff(*a, **k)


def empirical_hist(*a, **k):
    return default().empirical_hist(*a, **k)


def theoretical_hist(*a, **k):
    return default().theoretical_hist(*a, **k)


def filter_fasta(*a, **k):
    return default().filter_fasta(*a, **k)


def scan_fp64(*a, **k):
    return default().scan_fp64(*a, **k)


def blas_open(path: str = "") -> None:
    default().blas_open(path)


def scan_sgemm(*a, **k):
    return default().scan_sgemm(*a, **k)


def run_cli(args, cwd: str, check: bool = True) -> subprocess.CompletedProcess:
    if not os.path.exists(BIN):
        build()
    return subprocess.run([BIN] + list(args), cwd=cwd, capture_output=True, text=True, check=check)
