#!/usr/bin/env python3
"""Benchmark of the MI355X `blamm scan` engine.

Headline workload (the precision form): 2,000 motifs with lengths uniform in
5..30 (counts = round(100 Dirichlet(0.3))), forward + reverse complement =
4,000 PWM columns; 4 manifest groups of 62.5 Mbp each (GC uniform in
0.35..0.50), every group with its own background, PWMs and thresholds (the
theoretical histogram's cutoff at p = 1e-5).  Synthetic, seeded.

A step is one scan of all four groups, each against its own resident library
(one engine context per group): MFMA filter kernel + exact finalize (run
boundaries, FP64 band rescoring) + hit count read-back per group.  Inputs are
in HBM before timing starts.  `value` is the split-BF16 (hi-only filter)
throughput in Gbp*motifs/s; the FP32 scan of the same groups, the lazy-lo
form and an FP64-oracle slice check ride in the same line, with the plain-GPU
workload (100 Mbp x 2,000 columns), the batched and serving workloads, hist -e
and the reference's CPU configuration as side legs.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

PEAK_BF16_TFLOPS = 2500.0     # MI355X dense BF16 MFMA
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 MFMA
PEAK_HBM_GBS = 8000.0
PEAK_SCLK_MHZ = 2400.0        # MI355X peak engine clock, at which the MFMA peaks are quoted


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


SOURCE_FILES = ("blamm_amd/csrc/hip/scan_kernels.hpp", "blamm_amd/csrc/hip/blamm_hip.hip")


def source_hash() -> str:
    """SHA-256 (first 16 hex digits) of the kernel sources the profiles were
    taken on; a profile whose stamp differs describes another build."""
    import hashlib
    h = hashlib.sha256()
    for f in SOURCE_FILES:
        with open(os.path.join(REPO, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def profile_figures(name: str):
    """`key: value` lines of a committed rocprofv3 summary under profiles/, or
    None when it is absent or was taken on other kernel sources."""
    path = os.path.join(REPO, "profiles", name)
    out = {}
    try:
        for line in open(path):
            if ":" in line and not line.startswith(" "):
                k, v = line.split(":", 1)
                out[k.strip()] = v.strip()
    except OSError:
        return None
    if out.get("source_sha") != source_hash():
        return {"stale": True, "source_sha": out.get("source_sha")}
    return out


def library_for(ms, bg, pvalues):
    """fwd+rc PWM columns of a motif set and, per p-value, the thresholds
    `blamm scan -pt` derives from the theoretical histograms (fwd and rc share
    one histogram)."""
    import blamm_amd as ba
    from blamm_amd import synth
    pwms, thr = [], {p: [] for p in pvalues}
    for pfm in ms.pfms:
        fwd = ba.pwm(pfm, bg, 0.25, False)
        rc = ba.pwm(synth.reverse_complement_pfm(pfm), bg, 0.25, True)
        pwms += [fwd, rc]
        for p in pvalues:
            t = ba.pvalue_threshold(fwd, bg, p)
            thr[p] += [t, t]
    return pwms, {p: np.asarray(v, np.float32) for p, v in thr.items()}


def make_workload(seed: int, nbases: int, pvalue: float):
    import blamm_amd as ba
    from blamm_amd import synth
    t0 = time.time()
    stream, runs = synth.config1_sequence(seed=seed, n=nbases)
    ms = synth.config1_motifs(seed=seed)
    bg = synth.counts_of(stream)
    pwms, thr = library_for(ms, bg, [pvalue])
    thr = thr[pvalue]
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    codes = ba.pack_codes(stream)
    log(f"workload: {stream.size} bases, {len(runs)} runs, {len(pwms)} columns "
        f"(generated in {time.time() - t0:.1f}s)")
    return stream, runs, pwms, thr, P, lengths, codes


def to_device(codes, n, runs):
    """Packed words (with the engine's tail padding) and run starts as device
    tensors."""
    import torch
    import blamm_amd as ba
    pw = int(ba.hip_lib().blamm_hip_packed_words(n))
    host = np.zeros(pw, np.uint64)
    host[:min(codes.size, pw)] = codes[:pw]
    d_codes = torch.from_numpy(host.view(np.int64)).to("cuda:0")
    d_runs = torch.from_numpy(np.ascontiguousarray(runs, np.uint64).view(np.int64)).to("cuda:0")
    return d_codes, d_runs


class DeviceScan:
    """One engine context scanning a device-resident stream (the timed step)."""

    def __init__(self, P, lengths, thr, precision):
        import torch
        import blamm_amd as ba
        self.ba, self.torch = ba, torch
        self.sc = ba.HipScanner(0)
        self.sc.set_library(P, lengths, thr, precision)
        self.cap = 1 << 20
        self.d_hits = torch.empty(self.cap * 2, dtype=torch.int64, device="cuda:0")
        self.d_nhits = torch.zeros(1, dtype=torch.int64, device="cuda:0")

    def tune(self, d_codes, n, d_runs, nruns, cache):
        t, ms = self.sc.autotune(d_codes, None, n, d_runs, nruns, cache)
        return {**t.as_dict(), "sample_bases": int(n), "sample_ms": ms}

    def step(self, d_codes, n, d_runs, nruns):
        self.launch(d_codes, n, d_runs, nruns)
        return self.finish(d_codes, n, d_runs, nruns)

    def launch(self, d_codes, n, d_runs, nruns, stream_ptr=0):
        """Enqueue the scan (on `stream_ptr`, else the context's stream)."""
        self.sc.scan_device(d_codes, None, n, d_runs, nruns, self.d_hits, self.cap, self.d_nhits, stream_ptr)

    def finish(self, d_codes, n, d_runs, nruns):
        """Hit count of the enqueued scan; on an overflow (scratch or hit
        buffer too small) grow and scan again."""
        ba = self.ba
        while True:
            try:
                return self.sc.scan_finish()
            except ba.BlammError as e:
                if e.code != ba.BLAMM_ERR_OVERFLOW:
                    raise
                need = int(self.d_nhits.item())
                if need > self.cap:
                    self.cap = need + need // 8
                    self.d_hits = self.torch.empty(self.cap * 2, dtype=self.torch.int64, device="cuda:0")
                    self.torch.cuda.synchronize()
            self.launch(d_codes, n, d_runs, nruns)

    def hits(self, nh):
        return self.d_hits[: 2 * nh].cpu().numpy().view(self.ba.HIT_DTYPE)

    def close(self):
        self.sc.close()


def gpu_leg(args, stream, runs, P, lengths, thr, codes, precision, tune_sample):
    import torch
    n = int(stream.size)
    d_codes, d_runs = to_device(codes, n, runs)
    ds = DeviceScan(P, lengths, thr, precision)
    tuning = None
    if not args.no_tune:
        m = min(n, tune_sample)
        rs = runs[runs < m]
        dc, dr = to_device(codes, m, rs)
        tuning = ds.tune(dc, m, dr, int(rs.size), args.tune_cache)
        del dc, dr
    nruns = int(runs.size)
    torch.cuda.synchronize()
    for _ in range(max(1, args.warmup)):
        nh = ds.step(d_codes, n, d_runs, nruns)
    torch.cuda.synchronize()
    filt, kern = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nh = ds.step(d_codes, n, d_runs, nruns)
        st = ds.sc.stats()
        filt.append(st.filter_ms)
        kern.append(st.kernel_ms)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = ds.sc.stats()
    band, err = ds.sc.bounds()
    out = {"ms_per_step": dt * 1e3, "hits": nh, "candidates": int(st.candidates),
           "band_rescored": int(st.band_rescored), "filter_ms": float(np.mean(filt)),
           "kernel_ms": float(np.mean(kern)), "filter_band_max": band, "score_err_bound": err,
           "tuning": tuning or ds.sc.tuning().as_dict()}
    # keep a copy of the hits for the cross-precision identity check
    hits = ds.hits(nh)
    if precision == ds.ba.PREC_BF16X2:
        out["host_api"] = host_api_leg(ds, codes, n, runs, nh)
    ds.close()
    return out, hits


def host_api_leg(ds, codes, n, runs, nh, reps=3):
    """The same scan through the host-buffer entry point the CLI uses: the
    packed stream is uploaded and every hit read back inside the timed call."""
    import ctypes
    ba = ds.ba
    lib = ds.sc.lib
    pw = int(lib.blamm_hip_packed_words(n))
    cbuf = np.zeros(pw, np.uint64)
    cbuf[:min(codes.size, pw)] = codes[:pw]
    rr = np.ascontiguousarray(runs, np.uint64)
    cap = nh + 4096
    hits = np.empty(cap, dtype=ba.HIT_DTYPE)
    cnt = ctypes.c_uint64()

    def call():
        rc = lib.blamm_hip_scan(ds.sc.ctx, cbuf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), None, n,
                                rr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), int(rr.size),
                                hits.ctypes.data, cap, ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError(lib.blamm_hip_last_error().decode())

    hp = ctypes.c_void_p()

    def call_pinned():
        rc = lib.blamm_hip_scan_pinned(ds.sc.ctx, cbuf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), None, n,
                                       rr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), int(rr.size),
                                       ctypes.byref(hp), ctypes.byref(cnt))
        if rc != 0:
            raise RuntimeError(lib.blamm_hip_last_error().decode())

    out = {"hits": 0, "upload_bytes": int(pw * 8)}
    for name, fn in (("ms_per_scan", call), ("ms_per_scan_pinned", call_pinned)):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        out[name] = (time.perf_counter() - t0) / reps * 1e3
    out["hits"] = int(cnt.value)
    out["readback_bytes"] = int(cnt.value) * 16
    return out


def hit_keys(h, shift=16):
    return np.sort(h["pos"].astype(np.uint64) * np.uint64(1 << shift) + h["col"].astype(np.uint64))


def batched_leg(args):
    """Many uneven sequences in one persistent launch (captured in a hipGraph)
    against the same sequences scanned one launch each."""
    import torch
    import blamm_amd as ba
    from blamm_amd import synth
    t0 = time.time()
    ms, codes, starts, lengths = synth.config3(seed=3, nseq=args.batched_seqs)
    n = int(codes.size)
    bg = synth.counts_of(codes)
    pwms, thr = library_for(ms, bg, [1e-4])
    thr = thr[1e-4]
    P, L, _ = ba.build_pattern_matrix(pwms)
    ncols = len(pwms)
    words = ba.pack_codes(codes)
    log(f"batched workload: {len(lengths)} sequences, {n} bases, {ncols} columns ({time.time() - t0:.1f}s)")
    sc = ba.HipScanner(0)
    sc.set_library(P, L, thr, ba.PREC_BF16X2)
    tuning = None
    if not args.no_tune:
        d_codes, d_runs = to_device(words, n, starts)
        t, tms = sc.autotune(d_codes, None, n, d_runs, int(starts.size), args.tune_cache)
        tuning = t.as_dict()
        del d_codes, d_runs
    # one request holding every sequence (runs = sequence starts)
    req = ba.ScanRequest(words, n, starts, None, None)
    b = ba.Batch(sc, max_bases=n + 4096, max_requests=1)
    for _ in range(2):
        b.submit([req])
        nh_batch = b.wait()
    reps = max(1, args.steps // 2)
    dev, t_all = [], time.perf_counter()
    for _ in range(reps):
        b.submit([req])
        nh_batch = b.wait()
        dev.append(b.stats().kernel_ms)
    wall_batch = (time.perf_counter() - t_all) / reps
    # the library's first batch_hits call rebuilds blamm_hit from the read-back
    # records (timed here without the Python copy that b.hits() adds)
    import ctypes
    hp, hn = ctypes.c_void_p(), ctypes.c_uint64()
    t_exp = time.perf_counter()
    if sc.lib.blamm_hip_batch_hits(b.b, 0, ctypes.byref(hp), ctypes.byref(hn)) != 0:
        raise RuntimeError(sc.lib.blamm_hip_last_error().decode())
    expand_ms = (time.perf_counter() - t_exp) * 1e3
    bh = b.hits(0)
    b.close()
    # per-sequence launches over device-resident sequences (each sequence its
    # own word-aligned slice of one device buffer)
    woff = np.concatenate([[0], np.cumsum((lengths + 31) // 32)]).astype(np.int64)
    host = np.zeros(int(woff[-1]) + 16, np.uint64)
    for r in range(len(lengths)):
        s0 = int(starts[r])
        seg = codes[s0:s0 + int(lengths[r])]
        host[woff[r]:woff[r + 1]] = ba.pack_codes(seg)
    d_all = torch.from_numpy(host.view(np.int64)).to("cuda:0")
    cap = int(nh_batch) + 4096
    d_hits = torch.empty(cap * 2, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    lib, ctx = sc.lib, sc.ctx
    base = d_all.data_ptr()
    total_words = host.size
    import ctypes
    cnt = ctypes.c_uint64()

    def per_sequence_pass():
        """Each launch appends its hits after the previous launch's."""
        at, counts = 0, []
        for r in range(len(lengths)):
            for attempt in range(2):
                rc = lib.blamm_hip_scan_device(ctx, base + 8 * int(woff[r]), total_words - int(woff[r]), None, 0,
                                               int(lengths[r]), None, 0, d_hits.data_ptr() + 16 * at, cap - at,
                                               None, None)
                if rc != 0:
                    raise RuntimeError(lib.blamm_hip_last_error().decode())
                rc = lib.blamm_hip_scan_finish(ctx, ctypes.byref(cnt))
                if rc == ba.BLAMM_ERR_OVERFLOW and cnt.value == 0 and attempt == 0:
                    continue        # candidate scratch grown: run again
                if rc != 0:
                    raise RuntimeError(lib.blamm_hip_last_error().decode())
                break
            counts.append(int(cnt.value))
            at += int(cnt.value)
        return at, counts

    per_sequence_pass()
    t1 = time.perf_counter()
    nh_seq, counts = per_sequence_pass()
    wall_seq = time.perf_counter() - t1
    sh = d_hits[: 2 * nh_seq].cpu().numpy().view(ba.HIT_DTYPE).copy()
    sh["pos"] = sh["pos"] + np.repeat(starts.astype(np.uint64), counts)
    same = bool(len(bh) == len(sh) and np.array_equal(hit_keys(bh, 24), hit_keys(sh, 24)))
    sc.close()
    work = float(n) * ncols
    out = {
        "workload": f"batched uneven config: {len(lengths)} sequences lognormal (median 2 kbp, sigma 1, "
                    f"200 bp..200 kbp) GC 0.45, {ncols // 2} motifs (L 6-25) fwd+rc = {ncols} columns, p=1e-4",
        "bases": n, "columns": ncols, "sequences": int(len(lengths)),
        "graph_launch_ms": float(np.mean(dev)),
        "value": work / (np.mean(dev) / 1e3) / 1e9,
        "unit": "Gbp*motifs/s",
        "wall_ms_incl_transfers": wall_batch * 1e3,
        "hits_expand_ms": expand_ms,
        "per_sequence_launches_ms": wall_seq * 1e3,
        "per_sequence_value": work / wall_seq / 1e9,
        "speedup_vs_per_sequence": wall_seq / (np.mean(dev) / 1e3),
        "hits": int(nh_batch), "hits_per_sequence_launches": int(nh_seq),
        "identical_hit_set": same,
        "tuning": tuning,
    }
    del bh
    return out


def precision_workload(args):
    """The four groups of the precision form: per group its PWM columns,
    p = 1e-5 thresholds, pattern matrix, packed stream and device copies."""
    from blamm_amd import synth
    import blamm_amd as ba
    t0 = time.time()
    ms, recs = synth.config2(seed=2, n_each=args.precision_bases)
    groups = []
    for name, codes, gc in recs:
        bg = synth.counts_of(codes)
        pwms, thr = library_for(ms, bg, [1e-5])
        thr = thr[1e-5]
        P, L, _ = ba.build_pattern_matrix(pwms)
        n = int(codes.size)
        words = ba.pack_codes(codes)
        runs = np.zeros(1, np.uint64)
        d_codes, d_runs = to_device(words, n, runs)
        groups.append({"name": name, "gc": gc, "codes": codes, "n": n, "pwms": pwms, "thr": thr, "P": P, "L": L,
                       "d_codes": d_codes, "d_runs": d_runs})
    log(f"precision workload: {len(groups)} groups x {args.precision_bases} bases, "
        f"{len(groups[0]['pwms'])} columns ({time.time() - t0:.1f}s)")
    return ms, groups


class ClockSampler:
    """Samples the GPU's current shader clock (amdgpu sysfs `pp_dpm_sclk`,
    the level marked `*`) every 50 ms in a thread while running, so the
    roofline can be stated against the clock the chip actually held.  Only
    used when exactly one amdgpu card is visible; otherwise reports nothing."""

    def __init__(self):
        import glob
        cards = sorted(glob.glob("/sys/class/drm/card[0-9]*/device/pp_dpm_sclk"))
        self.path = cards[0] if len(cards) == 1 else None
        self.samples = []
        self._stop = None
        self._th = None

    def read(self):
        try:
            with open(self.path) as f:
                for ln in f:
                    if ln.rstrip().endswith("*"):
                        return float(ln.split(":")[1].strip().split()[0].lower().replace("mhz", ""))
        except (OSError, ValueError, IndexError):
            return None
        return None

    def __enter__(self):
        import threading
        if self.path:
            self._stop = threading.Event()

            def loop():
                while not self._stop.wait(0.05):
                    v = self.read()
                    if v:
                        self.samples.append(v)
            self._th = threading.Thread(target=loop, daemon=True)
            self._th.start()
        return self

    def __exit__(self, *exc):
        if self._th:
            self._stop.set()
            self._th.join()

    def summary(self):
        if not self.samples:
            return None
        v = np.asarray(self.samples)
        return {"source": self.path, "samples": int(v.size), "median_mhz": float(np.median(v)),
                "min_mhz": float(v.min()), "max_mhz": float(v.max())}


def groups_leg(args, groups, precision, steps, warmup, tune_cache=None, clock=None):
    """One context per group with its library resident; `steps` timed passes
    over all groups.  Returns timings and each group's hits."""
    import torch
    import blamm_amd as ba
    ctxs = [DeviceScan(g["P"], g["L"], g["thr"], precision) for g in groups]
    tuning = None
    if tune_cache is not None:
        g = groups[0]
        tuning = ctxs[0].tune(g["d_codes"], g["n"], g["d_runs"], 0, tune_cache)
        for c in ctxs[1:]:
            c.sc.set_tuning(tuning["chunk_tiles"], tuning["blocks_per_wave"], tuning["kpad_positions"])

    # The groups' scans are enqueued back to back on one stream, then
    # collected: the device runs them in order without a host round trip
    # between groups (the HIP events around each launch stay per scan).
    stream = torch.cuda.Stream()

    def one_pass():
        for c, g in zip(ctxs, groups):
            c.launch(g["d_codes"], g["n"], g["d_runs"], 0, stream.cuda_stream)
        res = []
        for c, g in zip(ctxs, groups):
            nh = c.finish(g["d_codes"], g["n"], g["d_runs"], 0)
            st = c.sc.stats()
            res.append((nh, st.filter_ms, st.kernel_ms, int(st.candidates), int(st.band_rescored)))
        return res

    for _ in range(max(1, warmup)):
        last = one_pass()
    torch.cuda.synchronize()
    filt, kern = [], []
    if clock is not None:
        clock.__enter__()
    t0 = time.perf_counter()
    for _ in range(steps):
        last = one_pass()
        filt += [r[1] for r in last]
        kern += [r[2] for r in last]
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    if clock is not None:
        clock.__exit__(None, None, None)
    out = {"ms_per_step": dt * 1e3, "hits": int(sum(r[0] for r in last)),
           "candidates": int(sum(r[3] for r in last)), "band_rescored": int(sum(r[4] for r in last)),
           "filter_ms_per_launch": float(np.mean(filt)), "kernel_ms_per_group": float(np.mean(kern)),
           "filter_band_max": max(c.sc.bounds()[0] for c in ctxs),
           "score_err_bound": max(c.sc.bounds()[1] for c in ctxs),
           "tuning": tuning or ctxs[0].sc.tuning().as_dict()}
    hits = [c.hits(r[0]) for c, r in zip(ctxs, last)]
    for c in ctxs:
        c.close()
    return out, hits


def fp64_slices(groups, hits, slice_bases):
    """The FP64 oracle on one slice of every group against the engine's hits
    there: hit-set difference and the largest score error."""
    import blamm_amd as ba
    from blamm_amd import synth
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    orc = oracle.default()
    rng = np.random.default_rng(7)
    diff, worst, nref, bases = 0, 0.0, 0, 0
    for g, h in zip(groups, hits):
        n = g["n"]
        a = int(rng.integers(0, max(1, n - slice_bases)))
        b = min(n, a + slice_bases)
        lmax = int(g["L"].max())
        text = synth.CODE2BASE[g["codes"][a:b]].tobytes()
        ref = orc.scan_fp64(text, [0], g["pwms"], g["thr"])
        keep = ref["pos"] + lmax <= b - a
        want = {(int(p) + a, int(c)): float(s) for p, c, s in zip(ref["pos"][keep], ref["col"][keep],
                                                                 ref["score64"][keep])}
        sel = (h["pos"] >= a) & (h["pos"] <= b - lmax)
        got = {(int(p), int(c)): float(s) for p, c, s in zip(h["pos"][sel], h["col"][sel], h["score"][sel])}
        diff += len(set(got) ^ set(want))
        nref += len(want)
        bases += b - a
        for k in set(got) & set(want):
            worst = max(worst, abs(got[k] - want[k]))
    return {"bases": bases, "slices": len(groups), "hits": nref, "hit_set_diff": diff,
            "max_abs_score_err_vs_fp64": worst}


def serve_leg(args):
    """Serving loop: resident library, Poisson arrivals, continuous batching;
    p50/p99 latency per offered load."""
    import threading
    import blamm_amd as ba
    from blamm_amd import synth
    t0 = time.time()
    ms = synth.config4_library(seed=4)
    rng = np.random.default_rng(4)
    bg = synth.counts_of(synth.bases(1_000_000, 0.41, rng))
    pvals = [1e-3, 1e-4, 1e-5, 1e-6]
    pwms, thr = library_for(ms, bg, pvals)
    P, L, _ = ba.build_pattern_matrix(pwms)
    ncols = len(pwms)
    draws = []
    for _ in range(args.serve_pool):
        codes, starts, subset, p = synth.config4_request(rng, len(ms.pfms), max_seqs=args.serve_max_seqs)
        if args.serve_pvalue:   # one p-value for every request (same pool otherwise)
            p = args.serve_pvalue
        draws.append((ba.pack_codes(codes), int(codes.size), starts, np.concatenate([2 * subset, 2 * subset + 1]), p))

    def make_pool(p_all=None):
        """Requests of the drawn pool, at their own p-values or all at p_all."""
        out = []
        for packed, n, starts, cols, p in draws:
            p = p_all or p
            t = np.full(ncols, np.inf, np.float32)
            t[cols] = thr[p][cols]
            # admission hint: the histograms' cutoffs pass a few times p per window
            hint = int(3 * p * n * cols.size) + 1024
            out.append((ba.ScanRequest(packed, n, starts, None, t, expected_hits=hint), float(n) * cols.size))
        return out

    pool = make_pool()
    log(f"serve workload: library {ncols} columns, pool of {len(pool)} requests, "
        f"{sum(r.nbases for r, _ in pool) / len(pool) / 1e6:.2f} Mbp per request ({time.time() - t0:.1f}s)")
    sc = ba.HipScanner(0)
    sc.set_library(P, L, thr[1e-4], ba.PREC_BF16X2)
    if args.serve_chunk_tiles:
        t = sc.tuning()
        sc.set_tuning(args.serve_chunk_tiles, t.blocks_per_wave, t.kpad_positions)
    srv = ba.Server(sc, hbm_budget_bytes=args.serve_budget_gib << 30, slots=args.serve_slots, max_batch_requests=64,
                    max_batch_bases=args.serve_batch_mbp * 1_000_000)

    def run(arrivals, pool):
        """Submit pool requests at the given offsets (s); returns latencies (ms),
        work and the span from first submit to last completion."""
        ids = [None] * len(arrivals)
        start = time.perf_counter()

        def submitter():
            for k, at in enumerate(arrivals):
                d = start + at - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[k] = srv.submit(pool[k % len(pool)][0])

        th = threading.Thread(target=submitter)
        th.start()
        lat, work, t_first, t_last, batch = [], 0.0, None, None, []
        for k in range(len(arrivals)):
            while ids[k] is None:
                time.sleep(0.0002)
            if not srv.wait(ids[k], 120_000):
                raise RuntimeError(f"serve: request {k} not finished after 120 s")
            info = srv.release(ids[k])
            lat.append((info.t_done_ns - info.t_submit_ns) / 1e6)
            work += pool[k % len(pool)][1]
            t_first = info.t_submit_ns if t_first is None else min(t_first, info.t_submit_ns)
            t_last = info.t_done_ns if t_last is None else max(t_last, info.t_done_ns)
            batch.append(info.batch_requests)
        th.join()
        span = (t_last - t_first) / 1e9
        return np.asarray(lat), work, span, float(np.mean(batch))

    run([0.0] * len(pool), pool)   # warm-up: captures the slot graphs, sizes scratch and pinned buffers
    n_closed = 4 * len(pool)
    s0 = srv.stats()
    lat, work, span, mb = run([0.0] * n_closed, pool)
    s1 = srv.stats()
    cap_rps = n_closed / span
    out = {
        "workload": f"serving config: library {ncols // 2} motifs (L 5-30, Dirichlet 0.3) fwd+rc = {ncols} columns; "
                    f"requests of 1-{args.serve_max_seqs} sequences (lognormal median 1 kbp, sigma 1, max 1 Mbp), "
                    f"motif subsets U{{10..2000}}, " +
                    (f"p = {args.serve_pvalue:g} for every request" if args.serve_pvalue else
                     "p in {1e-3,1e-4,1e-5,1e-6}"),
        "unit": "Gbp*motifs/s",
        "closed_loop": {"requests": n_closed, "rps": cap_rps, "value": work / span / 1e9,
                        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                        "mean_batch": mb, "span_ms": span * 1e3,
                        "device_ms": s1.device_ms - s0.device_ms,
                        "host_submit_ms": s1.host_submit_ms - s0.host_submit_ms,
                        "host_wait_ms": s1.host_wait_ms - s0.host_wait_ms,
                        "replays": int(s1.replays - s0.replays)},
        "poisson": [],
    }
    for f in args.serve_loads:
        qps = f * cap_rps
        nreq = max(16, int(qps * args.serve_seconds))
        arr = np.cumsum(rng.exponential(1.0 / qps, size=nreq))
        arr -= arr[0]
        lat, work, span, mb = run(list(arr), pool)
        out["poisson"].append({"offered_rps": qps, "requests": nreq, "value": work / span / 1e9,
                               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                               "mean_batch": mb})
        log(f"serve load {f:.2f}: {out['poisson'][-1]}")
    if not args.serve_pvalue:
        # the same requests, all at the headline scan's p-value (1e-5): what the
        # loop sustains when the output is as sparse as the headline's
        ph = make_pool(1e-5)
        run([0.0] * len(ph), ph)
        lat, work, span, mb = run([0.0] * n_closed, ph)
        cap_h = n_closed / span
        hp = {"closed_loop": {"requests": n_closed, "rps": cap_h, "value": work / span / 1e9,
                              "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                              "mean_batch": mb}}
        qps = 0.85 * cap_h
        nreq = max(16, int(qps * args.serve_seconds))
        arr = np.cumsum(rng.exponential(1.0 / qps, size=nreq))
        arr -= arr[0]
        lat, work, span, mb = run(list(arr), ph)
        hp["poisson_0.85"] = {"offered_rps": qps, "requests": nreq, "value": work / span / 1e9,
                              "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                              "mean_batch": mb}
        out["all_at_p1e-5"] = hp
        log(f"serve at p=1e-5: {hp}")
    st = srv.stats()
    out.update({"batches": int(st.batches), "replays": int(st.replays), "device_ms": st.device_ms,
                "host_submit_ms": st.host_submit_ms, "host_wait_ms": st.host_wait_ms, "hits": int(st.hits),
                "slots": args.serve_slots, "batch_mbp": args.serve_batch_mbp,
                "pool_mbp_per_request": sum(r.nbases for r, _ in pool) / len(pool) / 1e6,
                "pool_gbp_motifs_per_request": sum(w for _, w in pool) / len(pool) / 1e9})
    srv.close()
    sc.close()
    return out


def hist_leg(args, stream, runs):
    """Empirical histograms (`blamm hist -e`): the first 10 Mbp of the stream
    (hist's default -l) against the 1,000 forward motifs, 250 bins.  GPU time
    per call with the stream resident; the oracle (FP32 position-order scores,
    16 threads) on the first 1 Mbp for equality and as the CPU figure."""
    import torch
    import blamm_amd as ba
    from blamm_amd import synth
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    n = min(int(stream.size), 10_000_000)
    sub = stream[:n]
    rs = runs[runs < n]
    ms = synth.config1_motifs(seed=args.seed)
    bg = synth.counts_of(sub)
    pwms = [ba.pwm(p, bg, 0.25, False) for p in ms.pfms]
    mn = np.array([ba.min_score(w) for w in pwms], np.float32)
    mx = np.array([ba.max_score(w) for w in pwms], np.float32)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    nbins = 250
    sc = ba.HipScanner(0)
    try:
        sc.set_library(P, lengths, np.zeros(len(pwms), np.float32), ba.PREC_BF16X2)
        d_codes, d_runs = to_device(ba.pack_codes(sub), n, rs)
        torch.cuda.synchronize()
        for _ in range(max(1, args.warmup)):
            counts = sc.hist_device(d_codes, None, n, d_runs, len(rs), mn, mx, nbins)
        reps = max(3, args.steps)
        kern, exact = [], 0
        t0 = time.perf_counter()
        for _ in range(reps):
            counts = sc.hist_device(d_codes, None, n, d_runs, len(rs), mn, mx, nbins)
            st = sc.stats()
            kern.append(st.kernel_ms)
            exact = int(st.band_rescored)
        ms_call = (time.perf_counter() - t0) / reps * 1e3
        m = 1_000_000
        text = np.frombuffer(b"ACGT", np.uint8)[sub[:m]].tobytes()
        rsm = rs[rs < m]
        d_c2, d_r2 = to_device(ba.pack_codes(sub[:m]), m, rsm)
        g_small = sc.hist_device(d_c2, None, m, d_r2, len(rsm), mn, mx, nbins)
    finally:
        sc.close()
    cores = min(16, os.cpu_count() or 1)
    t1 = time.perf_counter()
    o_small = oracle.empirical_hist(text, rsm, pwms, mn, mx, nbins, threads=cores)
    t_cpu = time.perf_counter() - t1
    windows = int(counts.sum())
    kms = float(np.median(kern))
    return {
        "workload": f"hist -e: first {n} bases x {len(pwms)} forward motifs, {nbins} bins",
        "ms_per_call": ms_call, "kernel_ms": kms,
        "value": n * len(pwms) / (ms_call / 1e3) / 1e9, "unit": "Gbp*motifs/s",
        "value_kernels": n * len(pwms) / (kms / 1e3) / 1e9,
        "windows": windows, "windows_binned_from_fp32_sum": exact,
        "identical_to_oracle_first_1Mbp": bool(np.array_equal(g_small, o_small)),
        "cpu_oracle": {"value": m * len(pwms) / t_cpu / 1e9, "unit": "Gbp*motifs/s", "cores": cores,
                       "kind": "port", "sample": f"first {m} bases x {len(pwms)} motifs, {t_cpu:.2f} s"},
    }


def cli_leg(args):
    """The `blamm` command line end to end on the same workload written as
    files (FASTA with 60-column lines, jaspar motifs, one-group manifest):
    dict, theoretical hist, `scan -rc -pt <p>` into occurrences.txt and into
    /dev/null (formatting without the file write)."""
    import shutil
    import subprocess
    import tempfile
    import blamm_amd as ba
    from blamm_amd import synth
    d = tempfile.mkdtemp(prefix="blamm_cli_")
    try:
        synth.write_fasta(os.path.join(d, "chr.fa"), [synth.config1_record(seed=args.seed, n=args.bases)])
        synth.config1_motifs(seed=args.seed).write_jaspar(os.path.join(d, "motifs.jaspar"))
        with open(os.path.join(d, "sequences.mf"), "w") as f:
            f.write("grp\tchr.fa\n")
        times, phases = {}, {}
        pt = f"{args.pvalue:g}"
        for name, cmd in (("dict", ["dict", "sequences.mf"]),
                          ("hist", ["hist", "motifs.jaspar", "sequences.mf"]),
                          ("scan", ["scan", "-rc", "-pt", pt, "motifs.jaspar", "sequences.mf"]),
                          ("scan_devnull", ["scan", "-rc", "-pt", pt, "-o", "/dev/null", "motifs.jaspar",
                                            "sequences.mf"])):
            t0 = time.perf_counter()
            r = subprocess.run([ba.BLAMM_BIN] + cmd, cwd=d, capture_output=True, text=True,
                               env={**os.environ, "BLAMM_TIMING": "1"})
            times[name] = time.perf_counter() - t0
            if r.returncode != 0:
                raise RuntimeError(f"blamm {' '.join(cmd)} failed: {r.stderr[-500:]}")
            if name.startswith("scan"):
                phases[name] = {ln.split("] ", 1)[1].rsplit(" ", 2)[0]: float(ln.rsplit(" ", 2)[1])
                                for ln in r.stderr.splitlines() if ln.startswith("[blamm timing]")}
            log(f"cli {name}: {times[name]:.2f} s")
        out = os.path.join(d, "occurrences.txt")
        nlines = 0
        with open(out, "rb") as f:
            while True:
                b = f.read(1 << 24)
                if not b:
                    break
                nlines += b.count(b"\n")
        size = os.path.getsize(out)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"workload": f"blamm dict/hist/scan -rc -pt {pt} on a {args.bases // 1_000_000} Mbp FASTA x "
                        f"1,000 jaspar motifs", "seconds": times, "scan_phases_ms": phases,
            "occurrence_lines": nlines,
            "occurrence_bytes": size, "threads": os.cpu_count()}


def cpu_info():
    """Host cores this process may use (its affinity set), the CPU share it
    uses (OMP_NUM_THREADS when set, the box's share of cores for one GPU) and
    the CPU model."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    share = affinity
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        share = min(share, int(env))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return affinity, share, model


def cpu_leg(stream, runs, pwms, thr, budget_s: float, what: str):
    """The reference's CPU path -- strip matrices of h = 1000 rows x w = 250
    offsets, column tiles of the length-sorted pattern matrix, one OpenBLAS
    sgemm per tile and offset (BLAS threads 1), one worker per core
    (pwmscan.cpp:223-294) -- restated by the oracle, timed on a bounded prefix
    of the stream."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    affinity, cores, model = cpu_info()
    text = np.frombuffer(b"ACGT", np.uint8)[stream]
    block = 1000 * 250

    def run(nb):
        m = min(stream.size, nb)
        sub = text[:m].tobytes()
        rs = runs[runs < m]
        t0 = time.perf_counter()
        oracle.scan_sgemm(sub, rs, pwms, thr, threads=cores)
        return m, time.perf_counter() - t0

    # the reference orders motifs by length before building P (motif.cpp:436)
    order = sorted(range(len(pwms)), key=lambda c: pwms[c].shape[0])
    pwms = [pwms[c] for c in order]
    thr = np.asarray(thr, np.float32)[order]
    m1, t1 = run(cores * block)
    rounds = max(1, int(budget_s / max(t1, 1e-3)))
    m, t = run(cores * block * rounds) if rounds > 1 else (m1, t1)
    value = m * len(pwms) / t / 1e9
    return {"value": value, "unit": "Gbp*motifs/s", "cores": cores, "kind": "port",
            "affinity_cores": affinity, "cpu_model": model,
            "sample": f"first {m} bases of {what} x {len(pwms)} columns, sgemm path "
                      f"(h=1000, w=250, OpenBLAS sgemm, BLAS threads 1, {cores} workers), {t:.1f} s"}


def config0_leg(args):
    """The reference's CPU configuration (64 motifs L 6-20 fwd+rc = 128
    columns; 4 records of 250 kbp, 1% N-runs; p = 1e-4): the CPU path timed
    whole on the host's share of cores, the GPU scan of the same stream, and
    both hit sets against the FP64 oracle."""
    import torch
    import blamm_amd as ba
    from blamm_amd import synth
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    ms, recs = synth.config0(seed=0)
    stream, runs = synth.filtered(recs)
    bg = synth.counts_of(stream)
    pwms, thr = library_for(ms, bg, [1e-4])
    thr = thr[1e-4]
    n = int(stream.size)
    text = np.frombuffer(b"ACGT", np.uint8)[stream].tobytes()
    orc = oracle.default()
    ref = orc.scan_fp64(text, runs, pwms, thr)
    want = set(zip(ref["pos"].tolist(), ref["col"].tolist()))
    affinity, cores, model = cpu_info()
    order = sorted(range(len(pwms)), key=lambda c: pwms[c].shape[0])
    reps, t_cpu = 0, 0.0
    while t_cpu < 2.0 or reps < 3:
        t0 = time.perf_counter()
        cpu_hits = oracle.scan_sgemm(text, runs, [pwms[c] for c in order], thr[order], threads=cores)
        t_cpu += time.perf_counter() - t0
        reps += 1
    t_cpu /= reps
    cpu_set = {(int(p), int(order[c])) for p, c in zip(cpu_hits["pos"], cpu_hits["col"])}
    P, L, _ = ba.build_pattern_matrix(pwms)
    ds = DeviceScan(P, L, thr, ba.PREC_BF16X2)
    d_codes, d_runs = to_device(ba.pack_codes(stream), n, runs)
    for _ in range(3):
        nh = ds.step(d_codes, n, d_runs, int(runs.size))
    torch.cuda.synchronize()
    k = 20
    t0 = time.perf_counter()
    for _ in range(k):
        nh = ds.step(d_codes, n, d_runs, int(runs.size))
    t_gpu = (time.perf_counter() - t0) / k
    h = ds.hits(nh)
    ds.close()
    gpu_set = set(zip(h["pos"].tolist(), h["col"].tolist()))
    work = float(n) * len(pwms)
    return {"workload": f"reference CPU config: {len(pwms) // 2} motifs (L 6-20) fwd+rc = {len(pwms)} columns, "
                        f"4 records x 250 kbp GC 0.41 1% N-runs, p=1e-4",
            "bases": n, "columns": len(pwms), "hits_fp64": len(want),
            "cpu": {"value": work / t_cpu / 1e9, "ms": t_cpu * 1e3, "cores": cores, "cpu_model": model,
                    "kind": "port", "hit_set_diff_vs_fp64": len(cpu_set ^ want),
                    "note": "the FP32 sgemm path; windows within FP32 rounding of a threshold may differ"},
            "gpu": {"value": work / t_gpu / 1e9, "ms": t_gpu * 1e3, "hit_set_diff_vs_fp64": len(gpu_set ^ want)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bases", type=int, default=100_000_000, help="bases of the plain-GPU workload")
    ap.add_argument("--pvalue", type=float, default=1e-4, help="p-value of the plain-GPU workload")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-lo", action="store_true", help="skip the lazy-lo form of the split-BF16 scan")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64-oracle slices")
    ap.add_argument("--tune-cache", default=os.path.join(os.path.expanduser("~"), ".cache", "blamm_amd", "tune.txt"))
    ap.add_argument("--sections", default="plain,batched,serve,hist,config0",
                    help="side workloads after the headline: plain, batched, serve, hist, config0, cli "
                         "(comma list, or none; cli writes ~2 GB of occurrences to a temporary directory)")
    ap.add_argument("--batched-seqs", type=int, default=20_000)
    ap.add_argument("--precision-bases", type=int, default=62_500_000)
    ap.add_argument("--slice-bases", type=int, default=150_000, help="bases per group of the FP64-oracle check")
    ap.add_argument("--serve-pool", type=int, default=32)
    ap.add_argument("--serve-max-seqs", type=int, default=5000)
    ap.add_argument("--serve-budget-gib", type=int, default=32)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-batch-mbp", type=int, default=16, help="bases per serving batch (Mbp)")
    ap.add_argument("--serve-slots", type=int, default=3, help="serving batches in flight")
    ap.add_argument("--serve-chunk-tiles", type=int, default=0,
                    help="window tile (32-window tiles per work item) of the serving batches (0: the default)")
    ap.add_argument("--serve-pvalue", type=float, default=None,
                    help="give every serving request this p-value (1e-3..1e-6) instead of drawing it")
    ap.add_argument("--serve-loads", type=lambda v: [float(x) for x in v.split(",")], default=[0.3, 0.6, 0.85])
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES (for scaling FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(a multi-GPU scan would shard the stream by range with an Lmax-1 halo and no collective)", file=sys.stderr)
        sys.exit(2)
    import blamm_amd as ba
    if args.tune_cache:
        os.makedirs(os.path.dirname(args.tune_cache), exist_ok=True)
    ms, groups = precision_workload(args)
    ncols = len(groups[0]["pwms"])
    nbp = sum(g["n"] for g in groups)
    lsum = float(np.sum(groups[0]["L"], dtype=np.float64))
    flops_step = float(nbp) * 8.0 * lsum
    flops_launch = flops_step / len(groups)
    tune = None if args.no_tune else args.tune_cache
    log("split-bf16 (headline)")
    clock = ClockSampler()
    bf, bf_hits = groups_leg(args, groups, ba.PREC_BF16X2, args.steps, args.warmup, tune, clock)
    log(f"split-bf16: {bf}")
    sclk = clock.summary()
    log(f"shader clock during the timed scans: {sclk}")
    keys_bf = [hit_keys(h) for h in bf_hits]
    value = nbp * ncols / (bf["ms_per_step"] / 1e3) / 1e9
    achieved = flops_launch / (bf["filter_ms_per_launch"] / 1e3) / 1e12
    prof = profile_figures("filter_persistent_config2_traffic.txt") or {}
    traffic = None
    if prof and not prof.get("stale") and "traffic_bytes_per_launch" in prof:
        traffic = float(prof["traffic_bytes_per_launch"])
    pbytes = 4 * 4 * int(groups[0]["L"].max()) * ncols
    alg_launch = 0.25 * nbp / len(groups) + 16.0 * bf["hits"] / len(groups) + pbytes
    hbm = {"algorithmic_bytes_per_group": alg_launch,
           "algorithmic_bytes_formula": "0.25 B/base (2-bit stream) + 16 B/hit + the FP32 pattern matrix once",
           "achieved_GBs": alg_launch / (bf["kernel_ms_per_group"] / 1e3) / 1e9, "peak_GBs": PEAK_HBM_GBS}
    if prof and not prof.get("stale") and "scan_traffic_bytes_per_group" in prof:
        hbm["measured_scan_traffic_per_group"] = float(prof["scan_traffic_bytes_per_group"])
        hbm["traffic_over_algorithmic"] = float(prof["scan_traffic_bytes_per_group"]) / alg_launch
    line = {
        "metric": "Gbp*motifs/s (split-BF16 scan vs FP32 GPU scan, FP64-exact hit set)",
        "value": value,
        "unit": "Gbp*motifs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": bf["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x2 (hi terms on MFMA)->f32 (+f32/f64 rescoring)",
        "data": "synthetic (seeded)",
        "config": {"workload": f"precision config: {ncols // 2} motifs (L 5-30, Dirichlet 0.3) fwd+rc = {ncols} "
                               f"columns, {len(groups)} groups x {args.precision_bases / 1e6:g} Mbp GC U(0.35,0.50), "
                               f"own background/PWMs/thresholds per group, p=1e-5",
                   "bases": nbp, "columns": ncols, "groups": len(groups), "pvalue": 1e-5,
                   "precision": "split-bf16, hi-only filter"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic,
                     "kernel": "filter_persistent_kernel<true,false,false> (one launch per group)",
                     "algorithmic_flops_per_launch": flops_launch,
                     "traffic_source": "profiles/filter_persistent_config2_traffic.txt"
                                       + ("" if traffic is not None else " (absent or taken on other sources)"),
                     "source_sha": source_hash()},
        # the published peak assumes the 2.4 GHz peak engine clock; under this
        # sustained MFMA load the chip holds a lower clock (sampled above)
        "clock": None if sclk is None else dict(
            sclk, peak_at_clock_TFLOPs=PEAK_BF16_TFLOPS * sclk["median_mhz"] / PEAK_SCLK_MHZ,
            frac_at_clock=achieved / (PEAK_BF16_TFLOPS * sclk["median_mhz"] / PEAK_SCLK_MHZ)),
        "hbm": hbm,
        "scan": bf,
    }
    if not args.no_lo:
        log("split-bf16 lazy-lo")
        lo, lo_hits = groups_leg(args, groups, ba.PREC_BF16X2_LO, max(2, args.steps // 2), 1)
        lo["value"] = nbp * ncols / (lo["ms_per_step"] / 1e3) / 1e9
        lach = flops_launch / (lo["filter_ms_per_launch"] / 1e3) / 1e12
        lo.update({"roofline_achieved_TFLOPs": lach, "roofline_frac": lach / PEAK_BF16_TFLOPS,
                   "identical_hit_set": all(np.array_equal(a, hit_keys(b)) for a, b in zip(keys_bf, lo_hits))})
        del lo_hits
        line["split_lazy_lo"] = lo
        log(f"lazy lo: {lo}")
    if not args.no_fp32:
        log("fp32")
        fp, fp_hits = groups_leg(args, groups, ba.PREC_FP32, 2, 1)
        fp["value"] = nbp * ncols / (fp["ms_per_step"] / 1e3) / 1e9
        fach = flops_launch / (fp["filter_ms_per_launch"] / 1e3) / 1e12
        fp.update({"roofline_achieved_TFLOPs": fach, "roofline_frac": fach / PEAK_FP32_TFLOPS,
                   "identical_hit_set": all(np.array_equal(a, hit_keys(b)) for a, b in zip(keys_bf, fp_hits))})
        del fp_hits
        line["fp32"] = fp
        line["speedup_vs_fp32"] = value / fp["value"]
        log(f"fp32: {fp}")
    if not args.no_check:
        log("FP64 oracle slices")
        line["fp64_slices"] = fp64_slices(groups, bf_hits, args.slice_bases)
        log(f"fp64 slices: {line['fp64_slices']}")
    del bf_hits, keys_bf
    for g in groups:
        del g["d_codes"], g["d_runs"]
    sections = [x for x in args.sections.split(",") if x and x != "none"]
    if "plain" in sections or "hist" in sections:
        log("plain-GPU workload")
        stream, runs, pwms, thr, P, lengths, codes = make_workload(args.seed, args.bases, args.pvalue)
        if "plain" in sections:
            p_nbp = int(stream.size)
            p_flops = float(p_nbp) * 8.0 * float(np.sum(lengths, dtype=np.float64))
            pb, pb_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_BF16X2, tune_sample=p_nbp)
            pb["value"] = p_nbp * len(pwms) / (pb["ms_per_step"] / 1e3) / 1e9
            pach = p_flops / (pb["filter_ms"] / 1e3) / 1e12
            pb.update({"roofline_achieved_TFLOPs": pach, "roofline_frac": pach / PEAK_BF16_TFLOPS,
                       "workload": f"plain-GPU config: {len(pwms) // 2} motifs (L 5-30) fwd+rc = {len(pwms)} columns, "
                                   f"{args.bases // 1_000_000} Mbp GC 0.41 CpG 0.2x 0.5% N-runs, p={args.pvalue:g}",
                       "bases": p_nbp, "columns": len(pwms)})
            if "host_api" in pb:
                pb["host_api"]["value"] = p_nbp * len(pwms) / (pb["host_api"]["ms_per_scan"] / 1e3) / 1e9
                pb["host_api"]["value_pinned"] = p_nbp * len(pwms) / (pb["host_api"]["ms_per_scan_pinned"] / 1e3) / 1e9
            if not args.no_fp32:
                pf, pf_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_FP32, tune_sample=16_000_000)
                pf["value"] = p_nbp * len(pwms) / (pf["ms_per_step"] / 1e3) / 1e9
                pf["identical_hit_set"] = bool(np.array_equal(hit_keys(pb_hits), hit_keys(pf_hits)))
                pb["fp32"] = pf
                pb["speedup_vs_fp32"] = pb["value"] / pf["value"]
                del pf_hits
            del pb_hits
            line["plain_gpu"] = pb
            log(f"plain: {pb}")
        if "hist" in sections:
            log("hist leg")
            line["hist"] = hist_leg(args, stream, runs)
            log(f"hist: {line['hist']}")
        del stream, runs, codes
    if "batched" in sections:
        log("batched leg")
        line["batched"] = batched_leg(args)
        log(f"batched: {line['batched']}")
    if "serve" in sections:
        log("serve leg")
        line["serve"] = serve_leg(args)
        log(f"serve: {line['serve']}")
    if "config0" in sections:
        log("reference CPU config leg")
        line["config0"] = config0_leg(args)
        log(f"config0: {line['config0']}")
    if "cli" in sections:
        log("cli leg")
        line["cli"] = cli_leg(args)
        log(f"cli: {line['cli']}")
    if not args.no_cpu:
        log("cpu baseline leg")
        g = groups[0]
        line["cpu_baseline"] = cpu_leg(g["codes"], np.zeros(1, np.uint64), g["pwms"], g["thr"], args.cpu_budget,
                                       f"group {g['name']} of the precision workload")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
