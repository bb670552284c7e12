"""The scan entry point that allocates nothing (blamm_hip_scan_device_noalloc),
called through the C ABI and checked against blamm_hip_scan_device (needs the
GPU).

Both entry points run the same kernels; the no-alloc one keeps its counters
and filter records in the library's static device scratch.  Its hits must be
the same (position, column, score) records as scan_device's, on two calls in
a row, the second with a smaller input, and a scan whose filter records do not
fit the scratch must be rejected with BLAMM_ERR_OVERFLOW.
"""
import ctypes
import sys

import numpy as np
import pytest

import blamm_amd as ba
from blamm_amd import synth
from helpers import library, stream_case

pytestmark = pytest.mark.gpu

MAX_RECORDS = 4194304   # BLAMM_NOALLOC_MAX_RECORDS (include/blamm_hip.h)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def noalloc_fn():
    lib = ba.hip_lib()
    fn = lib.blamm_hip_scan_device_noalloc
    vp = ctypes.c_void_p
    fn.argtypes = [vp, vp, ctypes.c_uint64, vp, ctypes.c_uint64, ctypes.c_uint64, vp, ctypes.c_uint64, vp,
                   ctypes.c_uint64, vp, vp]
    fn.restype = ctypes.c_int
    return lib, fn


def device_stream(text, runs):
    import torch
    dev = torch.device("cuda:0")
    n = len(text)
    lib = ba.hip_lib()
    codes, lower = ba.pack(text)
    pw, mw = int(lib.blamm_hip_packed_words(n)), int(lib.blamm_hip_mask_words(n))
    hp = np.zeros(pw, np.uint64)
    hp[:codes.size] = codes
    d_codes = torch.from_numpy(hp.view(np.int64)).to(dev)
    d_low = None
    if lower is not None:
        hm = np.zeros(mw, np.uint64)
        hm[:lower.size] = lower
        d_low = torch.from_numpy(hm.view(np.int64)).to(dev)
    d_runs = torch.from_numpy(np.ascontiguousarray(runs, np.uint64).view(np.int64)).to(dev)
    return d_codes, d_low, d_runs


def run_both(sc, text, runs, cap):
    """(hits of scan_device, hits of scan_device_noalloc), each sorted."""
    import torch
    lib, fn = noalloc_fn()
    d_codes, d_low, d_runs = device_stream(text, runs)
    n, nr = len(text), len(runs)
    out = []
    for noalloc in (False, True):
        d_hits = torch.zeros(2 * cap, dtype=torch.int64, device="cuda:0")
        d_n = torch.zeros(1, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        if noalloc:
            rc = fn(sc.ctx, d_codes.data_ptr(), d_codes.numel(), d_low.data_ptr() if d_low is not None else None,
                    d_low.numel() if d_low is not None else 0, n, d_runs.data_ptr() if nr else None, nr,
                    d_hits.data_ptr(), cap, d_n.data_ptr(), None)
            assert rc == ba.BLAMM_OK, lib.blamm_hip_last_error().decode()
            nh = sc.scan_finish()
        else:
            sc.scan_device(d_codes, d_low, n, d_runs, nr, d_hits, cap, d_n)
            nh = sc.scan_finish()
        assert int(d_n.item()) == nh
        h = d_hits[: 2 * nh].cpu().numpy().view(ba.HIT_DTYPE)
        out.append(np.sort(h, order=["pos", "col"]))
    return out


@pytest.mark.parametrize("prec", [ba.PREC_BF16X2, ba.PREC_FP32])
def test_noalloc_matches_scan_device_twice(prec):
    """Two calls in a row, the second on a smaller stream: identical hit
    records to scan_device each time (the scratch keeps the first call's
    records and counters, which the second call must not see)."""
    text, runs, bg = stream_case(71, 600_000, lower_every=13)
    ms = synth.motifs(40, 5, 30, 0.5, seed=71)
    pwms, thr = library(ms, bg, pvalue=1e-3)
    P, L, _ = ba.build_pattern_matrix(pwms)
    with ba.HipScanner(0) as sc:
        sc.set_library(P, L, thr, prec)
        for n in (600_000, 150_001):
            sub = text[:n]
            rs = [int(r) for r in runs if r < n]
            say(f"noalloc: precision {prec}, {n} bases, {len(pwms)} columns")
            a, b = run_both(sc, sub, rs, cap=1 << 20)
            assert len(a) > 100
            assert np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["col"], b["col"])
            assert np.array_equal(a["score"].view(np.uint32), b["score"].view(np.uint32))


def test_noalloc_rejects_records_beyond_scratch():
    """Thresholds below every score make every lane of every tile a filter
    record: n * ncols / 16 records, above the scratch's maximum.  The call
    reports BLAMM_ERR_OVERFLOW with 0 hits and names the maximum; a following
    scan_device of the same stream still succeeds."""
    import torch
    rng = np.random.default_rng(5)
    n = 1 << 20
    text = synth.CODE2BASE[rng.integers(0, 4, n)].tobytes()
    pwms = [rng.normal(0, 1, size=(8, 4)).astype(np.float32) for _ in range(96)]
    thr = np.full(len(pwms), -1e6, np.float32)
    assert n * len(pwms) // 16 > MAX_RECORDS
    P, L, _ = ba.build_pattern_matrix(pwms)
    lib, fn = noalloc_fn()
    with ba.HipScanner(0) as sc:
        sc.set_library(P, L, thr, ba.PREC_BF16X2)
        d_codes, d_low, d_runs = device_stream(text, [])
        cap = 1 << 16
        d_hits = torch.zeros(2 * cap, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        say("noalloc overflow: 1 Mbp x 96 columns, every window a record")
        rc = fn(sc.ctx, d_codes.data_ptr(), d_codes.numel(), None, 0, n, None, 0, d_hits.data_ptr(), cap, None, None)
        assert rc == ba.BLAMM_OK, lib.blamm_hip_last_error().decode()
        nh = ctypes.c_uint64(123)
        rc = lib.blamm_hip_scan_finish(sc.ctx, ctypes.byref(nh))
        assert rc == ba.BLAMM_ERR_OVERFLOW
        assert nh.value == 0
        assert str(MAX_RECORDS) in lib.blamm_hip_last_error().decode()
        assert sc.stats().candidates > MAX_RECORDS
        # the ordinary entry point grows its own scratch and reports the real count
        d_small = torch.zeros(2 * 16, dtype=torch.int64, device="cuda:0")
        sc.scan_device(d_codes, None, n, d_runs, 0, d_small, 16, None)
        with pytest.raises(ba.BlammError) as e:
            sc.scan_finish()
        assert e.value.code == ba.BLAMM_ERR_OVERFLOW
