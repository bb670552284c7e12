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


def make_workload(seed: int, nbases: int, pvalue: float):
    import blamm_amd as ba
    from blamm_amd import synth
    t0 = time.time()
    stream, runs = synth.config1_sequence(seed=seed, n=nbases)
    ms = synth.config1_motifs(seed=seed)
    bg = synth.counts_of(stream)
    pwms, thr = [], []
    for pfm in ms.pfms:
        fwd = ba.pwm(pfm, bg, 0.25, False)
        rc = ba.pwm(synth.reverse_complement_pfm(pfm), bg, 0.25, True)
        t = ba.pvalue_threshold(fwd, bg, pvalue)   # fwd and rc share the histogram
        pwms += [fwd, rc]
        thr += [t, t]
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    codes = ba.pack_codes(stream)
    log(f"workload: {stream.size} bases, {len(runs)} runs, {len(pwms)} columns "
        f"(generated in {time.time() - t0:.1f}s)")
    return stream, runs, pwms, np.asarray(thr, np.float32), P, lengths, codes


def gpu_leg(args, stream, runs, P, lengths, thr, codes, precision):
    import torch
    import blamm_amd as ba
    n = int(stream.size)
    dev = torch.device("cuda:0")
    pw = int(ba.hip_lib().blamm_hip_packed_words(n))
    host = np.zeros(pw, np.uint64)
    host[:codes.size] = codes
    d_codes = torch.from_numpy(host.view(np.int64)).to(dev)
    d_runs = torch.from_numpy(np.ascontiguousarray(runs, np.uint64).view(np.int64)).to(dev)
    d_nhits = torch.zeros(1, dtype=torch.int64, device=dev)
    sc = ba.HipScanner(0)
    sc.set_library(P, lengths, thr, precision)
    cap = 1 << 20
    d_hits = torch.empty(cap * 2, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()

    def step():
        nonlocal cap, d_hits
        while True:
            sc.scan_device(d_codes, None, n, d_runs, int(runs.size), d_hits, cap, d_nhits)
            try:
                return sc.scan_finish()
            except ba.BlammError as e:
                if e.code != ba.BLAMM_ERR_OVERFLOW:
                    raise
                need = int(d_nhits.item())
                if need > cap:
                    cap = need + need // 8
                    d_hits = torch.empty(cap * 2, dtype=torch.int64, device=dev)
                    torch.cuda.synchronize()

    for _ in range(max(1, args.warmup)):
        nh = step()
    torch.cuda.synchronize()
    filt, kern = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        nh = step()
        st = sc.stats()
        filt.append(st.filter_ms)
        kern.append(st.kernel_ms)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    st = sc.stats()
    band, err = sc.bounds()
    out = {"ms_per_step": dt * 1e3, "hits": nh, "candidates": int(st.candidates),
           "band_rescored": int(st.band_rescored), "filter_ms": float(np.mean(filt)),
           "kernel_ms": float(np.mean(kern)), "filter_band_max": band, "score_err_bound": err}
    # keep a copy of the hits for the cross-precision identity check
    hits = d_hits[: 2 * nh].cpu().numpy().view(ba.HIT_DTYPE)
    sc.close()
    return out, hits


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
              f"(the scan shards by stream range with no collective; see the design notes)", file=sys.stderr)
        sys.exit(2)
    import blamm_amd as ba
    stream, runs, pwms, thr, P, lengths, codes = make_workload(args.seed, args.bases, args.pvalue)
    nbp = int(stream.size)
    ncols = len(pwms)
    flops = float(nbp) * 8.0 * float(np.sum(lengths, dtype=np.float64))
    log("split-bf16 leg")
    bf, bf_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_BF16X2)
    log(f"split-bf16: {bf}")
    fp = None
    if not args.no_fp32:
        log("fp32 leg")
        fp, fp_hits = gpu_leg(args, stream, runs, P, lengths, thr, codes, ba.PREC_FP32)
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
        "dtype": "bf16x2->f32 (+f64 band)",
        "data": "synthetic (seeded)",
        "config": {"workload": f"plain-GPU config: {ncols // 2} motifs (L 5-30) fwd+rc = {ncols} columns, "
                               f"{args.bases // 1_000_000} Mbp GC 0.41 CpG 0.2x 0.5% N-runs, p={args.pvalue:g}",
                   "bases": nbp, "columns": ncols, "pvalue": args.pvalue, "precision": "split-bf16"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic,
                     "kernel": "filter_persistent_kernel<true,2,false> (one launch per step)",
                     "algorithmic_flops": flops},
        "hbm": {"algorithmic_bytes": alg_bytes, "achieved_GBs": alg_bytes / (bf["kernel_ms"] / 1e3) / 1e9,
                "peak_GBs": PEAK_HBM_GBS},
        "scan": bf,
    }
    if fp is not None:
        fval = nbp * ncols / (fp["ms_per_step"] / 1e3) / 1e9
        fach = flops / (fp["filter_ms"] / 1e3) / 1e12
        fp.update({"value": fval, "roofline_achieved_TFLOPs": fach, "roofline_frac": fach / PEAK_FP32_TFLOPS})
        line["fp32"] = fp
        line["speedup_vs_fp32"] = value / fval
    if not args.no_cpu:
        log("cpu baseline leg")
        line["cpu_baseline"] = cpu_leg(stream, runs, pwms, thr, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
