#!/usr/bin/env python3
"""Benchmark of the MI355X `blamm scan` engine.

Workload (the plain-GPU drop-in configuration): 1,000 motifs with lengths
uniform in 5..30 (counts = round(100 Dirichlet(0.5))), forward + reverse
complement = 2,000 PWM columns; one 100 Mbp record at GC 0.41 with the CG
dinucleotide at 0.2x and 0.5% of positions in N-runs of 10..500; thresholds
from the theoretical histogram at p = 1e-4.  Synthetic, seeded.

A step is one scan of the whole resident stream against the whole resident
library: MFMA filter kernels + exact finalize (run boundaries, FP64 band
rescoring) + hit count read-back.  Inputs are in HBM before timing starts.
Prints one JSON line.
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


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per filter launch from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (None when absent)."""
    path = os.path.join(REPO, "profiles", "filter_persistent_config1_traffic.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


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
        ba = self.ba
        while True:
            self.sc.scan_device(d_codes, None, n, d_runs, nruns, self.d_hits, self.cap, self.d_nhits)
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
    cap = 1 << 20
    d_hits = torch.empty(cap * 2, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    lib, ctx = sc.lib, sc.ctx
    base = d_all.data_ptr()
    total_words = host.size
    import ctypes
    cnt = ctypes.c_uint64()

    def per_sequence_pass():
        tot = 0
        for r in range(len(lengths)):
            rc = lib.blamm_hip_scan_device(ctx, base + 8 * int(woff[r]), total_words - int(woff[r]), None, 0,
                                           int(lengths[r]), None, 0, d_hits.data_ptr(), cap, None, None)
            if rc != 0:
                raise RuntimeError(lib.blamm_hip_last_error().decode())
            if lib.blamm_hip_scan_finish(ctx, ctypes.byref(cnt)) != 0:
                raise RuntimeError(lib.blamm_hip_last_error().decode())
            tot += cnt.value
        return tot

    per_sequence_pass()
    t1 = time.perf_counter()
    nh_seq = per_sequence_pass()
    wall_seq = time.perf_counter() - t1
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
        "per_sequence_launches_ms": wall_seq * 1e3,
        "per_sequence_value": work / wall_seq / 1e9,
        "speedup_vs_per_sequence": wall_seq / (np.mean(dev) / 1e3),
        "hits": int(nh_batch), "hits_per_sequence_launches": int(nh_seq),
        "same_hit_count": bool(nh_batch == nh_seq),
        "tuning": tuning,
    }
    del bh
    return out


def precision_leg(args, orc_slice_bases=500_000):
    """Split-BF16 vs FP32 on 4 manifest groups (own background, PWMs and
    thresholds each), p = 1e-5; identical hit sets, and the max-abs score error
    against the FP64 oracle on a slice."""
    import torch
    import blamm_amd as ba
    from blamm_amd import synth
    t0 = time.time()
    ms, recs = synth.config2(seed=2, n_each=args.precision_bases)
    log(f"precision workload: {len(recs)} groups x {args.precision_bases} bases ({time.time() - t0:.1f}s)")
    res = {"bf16_ms": 0.0, "bf16lo_ms": 0.0, "fp32_ms": 0.0, "groups": [], "identical_hit_sets": True}
    total_work = 0.0
    slice_check = None
    for gi, (name, codes, gc) in enumerate(recs):
        bg = synth.counts_of(codes)
        pwms, thr = library_for(ms, bg, [1e-5])
        thr = thr[1e-5]
        P, L, _ = ba.build_pattern_matrix(pwms)
        n = int(codes.size)
        words = ba.pack_codes(codes)
        runs = np.zeros(1, np.uint64)
        d_codes, d_runs = to_device(words, n, runs)
        g = {"group": name, "gc": gc, "bases": n}
        keys = {}
        for prec, tag in ((ba.PREC_BF16X2, "bf16"), (ba.PREC_BF16X2_LO, "bf16lo"), (ba.PREC_FP32, "fp32")):
            ds = DeviceScan(P, L, thr, prec)
            nh = ds.step(d_codes, n, d_runs, 1)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(max(1, args.steps // 3)):
                nh = ds.step(d_codes, n, d_runs, 1)
            dt = (time.perf_counter() - t1) / max(1, args.steps // 3)
            res[f"{tag}_ms"] += dt * 1e3
            g[f"{tag}_ms"] = dt * 1e3
            g["hits"] = int(nh)
            keys[tag] = hit_keys(ds.hits(nh))
            if tag == "bf16":
                g["score_err_bound"] = ds.sc.bounds()[1]
            ds.close()
        g["identical"] = bool(np.array_equal(keys["bf16"], keys["fp32"]) and np.array_equal(keys["bf16"], keys["bf16lo"]))
        res["identical_hit_sets"] &= g["identical"]
        total_work += float(n) * len(pwms)
        res["groups"].append(g)
        if gi == 0 and orc_slice_bases:
            # FP64 oracle on a slice of the first group
            sys.path.insert(0, os.path.join(REPO, "oracle"))
            import oracle
            m = min(n, orc_slice_bases)
            text = np.frombuffer(b"ACGT", np.uint8)[codes[:m]].tobytes()
            ref = oracle.default().scan_fp64(text, [0], pwms, thr)
            sc = ba.HipScanner(0)
            sc.set_library(P, L, thr, ba.PREC_BF16X2)
            got = sc.scan(ba.pack_codes(codes[:m]), m, [0], None)
            sc.close()
            a = np.lexsort((got["col"], got["pos"]))
            bb = np.lexsort((ref["col"], ref["pos"]))
            same = bool(len(got) == len(ref) and np.array_equal(got["pos"][a], ref["pos"][bb])
                        and np.array_equal(got["col"][a], ref["col"][bb]))
            err = float(np.max(np.abs(got["score"][a].astype(np.float64) - ref["score64"][bb]))) if same and len(got) else None
            slice_check = {"bases": m, "hits": int(len(ref)), "hit_set_diff": 0 if same else
                           int(len(set(zip(got["pos"], got["col"])) ^ set(zip(ref["pos"], ref["col"])))),
                           "max_abs_score_err_vs_fp64": err}
        del d_codes, d_runs
    ncols = 2 * len(ms.pfms)
    res.update({
        "workload": f"precision config: {len(ms.pfms)} motifs (L 5-30, Dirichlet 0.3) fwd+rc = {ncols} columns, "
                    f"{len(recs)} groups x {args.precision_bases // 1_000_000} Mbp GC U(0.35,0.50), p=1e-5",
        "value": total_work / (res["bf16_ms"] / 1e3) / 1e9,
        "fp32_value": total_work / (res["fp32_ms"] / 1e3) / 1e9,
        "lazy_lo_value": total_work / (res["bf16lo_ms"] / 1e3) / 1e9,
        "unit": "Gbp*motifs/s",
        "speedup_vs_fp32": res["fp32_ms"] / res["bf16_ms"],
        "fp64_slice": slice_check,
    })
    return res


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
    pool = []
    for _ in range(args.serve_pool):
        codes, starts, subset, p = synth.config4_request(rng, len(ms.pfms), max_seqs=args.serve_max_seqs)
        t = np.full(ncols, np.inf, np.float32)
        cols = np.concatenate([2 * subset, 2 * subset + 1])
        t[cols] = thr[p][cols]
        # admission hint: the histograms' cutoffs pass a few times p per window
        hint = int(3 * p * codes.size * cols.size) + 1024
        req = ba.ScanRequest(ba.pack_codes(codes), int(codes.size), starts, None, t, expected_hits=hint)
        pool.append((req, float(codes.size) * cols.size))
    log(f"serve workload: library {ncols} columns, pool of {len(pool)} requests, "
        f"{sum(r.nbases for r, _ in pool) / len(pool) / 1e6:.2f} Mbp per request ({time.time() - t0:.1f}s)")
    sc = ba.HipScanner(0)
    sc.set_library(P, L, thr[1e-4], ba.PREC_BF16X2)
    srv = ba.Server(sc, hbm_budget_bytes=args.serve_budget_gib << 30, slots=2, max_batch_requests=64,
                    max_batch_bases=args.serve_batch_mbp * 1_000_000)

    def run(arrivals):
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

    run([0.0] * len(pool))   # warm-up: captures the slot graphs, sizes scratch and pinned buffers
    n_closed = 2 * len(pool)
    s0 = srv.stats()
    lat, work, span, mb = run([0.0] * n_closed)
    s1 = srv.stats()
    cap_rps = n_closed / span
    out = {
        "workload": f"serving config: library {ncols // 2} motifs (L 5-30, Dirichlet 0.3) fwd+rc = {ncols} columns; "
                    f"requests of 1-{args.serve_max_seqs} sequences (lognormal median 1 kbp, sigma 1, max 1 Mbp), "
                    f"motif subsets U{{10..2000}}, p in {{1e-3,1e-4,1e-5,1e-6}}",
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
        lat, work, span, mb = run(list(arr))
        out["poisson"].append({"offered_rps": qps, "requests": nreq, "value": work / span / 1e9,
                               "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                               "mean_batch": mb})
        log(f"serve load {f:.2f}: {out['poisson'][-1]}")
    st = srv.stats()
    out.update({"batches": int(st.batches), "replays": int(st.replays), "device_ms": st.device_ms,
                "host_submit_ms": st.host_submit_ms, "host_wait_ms": st.host_wait_ms})
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


def cpu_leg(stream, runs, pwms, thr, budget_s: float):
    """The reference's sgemm scan (oracle restatement, OpenBLAS, BLAS threads 1)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    cores = min(16, os.cpu_count() or 1)
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
            "sample": f"first {m} bases of the same stream x {len(pwms)} columns, sgemm path "
                      f"(h=1000, w=250, OpenBLAS sgemm, BLAS threads 1), {t:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bases", type=int, default=100_000_000)
    ap.add_argument("--pvalue", type=float, default=1e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--no-lo", action="store_true", help="skip the lazy-lo form of the split-BF16 scan")
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile shapes")
    ap.add_argument("--tune-cache", default=os.path.join(os.path.expanduser("~"), ".cache", "blamm_amd", "tune.txt"))
    ap.add_argument("--sections", default="batched,precision,serve,hist",
                    help="extra workloads after the headline scan: batched, precision, serve, hist, cli "
                         "(comma list, or none; cli writes ~2 GB of occurrences to a temporary directory)")
    ap.add_argument("--batched-seqs", type=int, default=20_000)
    ap.add_argument("--precision-bases", type=int, default=62_500_000)
    ap.add_argument("--serve-pool", type=int, default=32)
    ap.add_argument("--serve-max-seqs", type=int, default=5000)
    ap.add_argument("--serve-budget-gib", type=int, default=32)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-batch-mbp", type=int, default=16, help="bases per serving batch (Mbp)")
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
    stream, runs, pwms, thr, P, lengths, codes = make_workload(args.seed, args.bases, args.pvalue)
    nbp = int(stream.size)
    ncols = len(pwms)
    flops = float(nbp) * 8.0 * float(np.sum(lengths, dtype=np.float64))
    log("split-bf16 leg")
    bf, bf_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_BF16X2, tune_sample=nbp)
    log(f"split-bf16: {bf}")
    lo = None
    if not args.no_lo:
        log("split-bf16 lazy-lo leg")
        lo, lo_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_BF16X2_LO, tune_sample=nbp)
        log(f"split-bf16 lazy lo: {lo}")
        a = np.sort(bf_hits["pos"].astype(np.uint64) * np.uint64(1 << 16) + bf_hits["col"])
        b = np.sort(lo_hits["pos"].astype(np.uint64) * np.uint64(1 << 16) + lo_hits["col"])
        lo["identical_hit_set"] = bool(a.size == b.size and np.array_equal(a, b))
        del lo_hits
    fp = None
    if not args.no_fp32:
        log("fp32 leg")
        fp, fp_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_FP32, tune_sample=16_000_000)
        log(f"fp32: {fp}")
        a = np.sort(bf_hits["pos"].astype(np.uint64) * np.uint64(1 << 16) + bf_hits["col"])
        b = np.sort(fp_hits["pos"].astype(np.uint64) * np.uint64(1 << 16) + fp_hits["col"])
        fp["identical_hit_set"] = bool(a.size == b.size and np.array_equal(a, b))
    value = nbp * ncols / (bf["ms_per_step"] / 1e3) / 1e9
    traffic = measured_traffic()
    achieved = flops / (bf["filter_ms"] / 1e3) / 1e12
    alg_bytes = 0.25 * nbp + 16.0 * bf["hits"]
    line = {
        "metric": "Gbp*motifs/s (split-BF16 scan, FP64-exact hit set)",
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
        "config": {"workload": f"plain-GPU config: {ncols // 2} motifs (L 5-30) fwd+rc = {ncols} columns, "
                               f"{args.bases // 1_000_000} Mbp GC 0.41 CpG 0.2x 0.5% N-runs, p={args.pvalue:g}",
                   "bases": nbp, "columns": ncols, "pvalue": args.pvalue,
                   "precision": "split-bf16, hi-only filter"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic,
                     "kernel": "filter_persistent_kernel<true,false,false> (one launch per step)",
                     "algorithmic_flops": flops},
        "hbm": {"algorithmic_bytes": alg_bytes, "achieved_GBs": alg_bytes / (bf["kernel_ms"] / 1e3) / 1e9,
                "peak_GBs": PEAK_HBM_GBS},
        "scan": bf,
    }
    if "host_api" in bf:
        bf["host_api"]["value"] = nbp * ncols / (bf["host_api"]["ms_per_scan"] / 1e3) / 1e9
        bf["host_api"]["value_pinned"] = nbp * ncols / (bf["host_api"]["ms_per_scan_pinned"] / 1e3) / 1e9
    if lo is not None:
        lo["value"] = nbp * ncols / (lo["ms_per_step"] / 1e3) / 1e9
        lach = flops / (lo["filter_ms"] / 1e3) / 1e12
        lo.update({"roofline_achieved_TFLOPs": lach, "roofline_frac": lach / PEAK_BF16_TFLOPS})
        line["split_lazy_lo"] = lo
    if fp is not None:
        fval = nbp * ncols / (fp["ms_per_step"] / 1e3) / 1e9
        fach = flops / (fp["filter_ms"] / 1e3) / 1e12
        fp.update({"value": fval, "roofline_achieved_TFLOPs": fach, "roofline_frac": fach / PEAK_FP32_TFLOPS})
        line["fp32"] = fp
        line["speedup_vs_fp32"] = value / fval
    del bf_hits
    if fp is not None:
        del fp_hits
    sections = [x for x in args.sections.split(",") if x and x != "none"]
    if "batched" in sections:
        log("batched leg")
        line["batched"] = batched_leg(args)
        log(f"batched: {line['batched']}")
    if "precision" in sections:
        log("precision leg")
        line["precision_form"] = precision_leg(args)
        log(f"precision: {line['precision_form']}")
    if "serve" in sections:
        log("serve leg")
        line["serve"] = serve_leg(args)
        log(f"serve: {line['serve']}")
    if "cli" in sections:
        log("cli leg")
        line["cli"] = cli_leg(args)
        log(f"cli: {line['cli']}")
    if "hist" in sections:
        log("hist leg")
        line["hist"] = hist_leg(args, stream, runs)
        log(f"hist: {line['hist']}")
    if not args.no_cpu:
        log("cpu baseline leg")
        line["cpu_baseline"] = cpu_leg(stream, runs, pwms, thr, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
