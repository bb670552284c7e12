"""The precision, batched and serving workloads at their full sizes against the
FP64 oracle (needs the GPU).

- precision form: 2,000 motifs (L 5-30, Dirichlet 0.3) fwd+rc = 4,000 columns,
  4 groups of 62.5 Mbp (GC U(0.35, 0.50)), each with its own background, PWMs
  and p = 1e-5 thresholds -- the per-group loop of `blamm scan`
  (pwmscan.cpp:589-640);
- batched form: 20,000 sequences (lognormal, median 2 kbp) x 1,500 motifs
  fwd+rc = 3,000 columns in one captured launch sequence;
- serving form: a server over the 2,000-motif library, requests of 1-5,000
  sequences with motif subsets and p in {1e-3, 1e-4, 1e-5, 1e-6}.

The hit set of a window is decided by its FP64 left-to-right sum against the
float threshold (the oracle predicate); every check below compares hit SETS
(position, column) bit for bit, and scores against the engine's stated bound.
"""
import hashlib
import sys

import numpy as np
import pytest

import blamm_amd as ba
from blamm_amd import synth
from helpers import library

pytestmark = pytest.mark.gpu

PRECS = [ba.PREC_BF16X2, ba.PREC_BF16X2_LO, ba.PREC_FP32]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def keys(h, shift=16):
    return np.sort(np.asarray(h["pos"], np.uint64) * np.uint64(1 << shift) + np.asarray(h["col"], np.uint64))


def fp64_score(w, codes, q):
    """Left-to-right FP64 sum of a window (the oracle predicate's score)."""
    s = 0.0
    for p in range(w.shape[0]):
        s += float(w[p, codes[q + p]])
    return s


def check_sampled(h, codes, pwms, thr, err, rng, n=1500):
    """Sampled hits: FP64 score >= threshold and reported score within err."""
    if len(h) == 0:
        return 0.0
    worst = 0.0
    for i in rng.choice(len(h), size=min(n, len(h)), replace=False):
        q, c = int(h["pos"][i]), int(h["col"][i])
        s64 = fp64_score(pwms[c], codes, q)
        assert s64 >= float(thr[c]), (q, c, s64, thr[c])
        worst = max(worst, abs(s64 - float(h["score"][i])))
    assert worst <= err, (worst, err)
    return worst


# ---- precision form (4 groups x 62.5 Mbp x 4,000 columns, p = 1e-5) ----------

@pytest.fixture(scope="module")
def config2():
    say("precision workload: generating 4 x 62.5 Mbp")
    ms, recs = synth.config2(seed=2, n_each=62_500_000)
    return ms, recs


@pytest.fixture(scope="module")
def scanner():
    with ba.HipScanner(0) as sc:
        yield sc


@pytest.mark.parametrize("group", [0, 1, 2, 3])
def test_precision_config_group(scanner, orc, config2, group):
    """One manifest group of the precision workload: the three precision modes
    give the same hit set over all 62.5 Mbp; a slice equals the FP64 oracle
    exactly; sampled hits re-score at or above the threshold within the
    stated bound; sampled near-threshold non-hits (windows that a threshold
    0.25 lower admits) re-score below the threshold."""
    ms, recs = config2
    name, codes, gc = recs[group]
    bg = synth.counts_of(codes)
    pwms, thr = library(ms, bg, pvalue=1e-5)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    n = int(codes.size)
    packed = ba.pack_codes(codes)
    out = {}
    for prec in PRECS:
        say(f"precision group {group} ({name}, GC {gc:.3f}): scan prec={prec}")
        scanner.set_library(P, lengths, thr, prec)
        out[prec] = scanner.scan(packed, n, [0], None, cap=24_000_000)
    h = out[ba.PREC_BF16X2]
    kb = keys(h)
    assert len(kb) > 1_000_000
    for prec in (ba.PREC_BF16X2_LO, ba.PREC_FP32):
        assert np.array_equal(kb, keys(out[prec])), f"group {group}: prec {prec} hit set differs"
    del out
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    _, err = scanner.bounds()
    rng = np.random.default_rng(100 + group)
    # an exact slice against the FP64 oracle
    lmax = int(lengths.max())
    a = int(rng.integers(0, n - 200_000))
    b = a + 120_000
    say(f"precision group {group}: oracle slice [{a}, {b})")
    text = synth.CODE2BASE[codes[a:b]].tobytes()
    ref = orc.scan_fp64(text, [0], pwms, thr)
    want = {(int(p) + a, int(c)) for p, c in zip(ref["pos"], ref["col"]) if p + lmax <= b - a}
    sel = (h["pos"] >= a) & (h["pos"] <= b - lmax)
    got = set(zip(h["pos"][sel].tolist(), h["col"][sel].tolist()))
    assert got == want and len(want) > 100, (group, len(got ^ want))
    worst = check_sampled(h, codes, pwms, thr, err, rng)
    # near-threshold non-hits over the whole group
    say(f"precision group {group}: near-threshold windows")
    lowered = (thr - np.float32(0.25)).astype(np.float32)
    scanner.set_library(P, lengths, lowered, ba.PREC_BF16X2)
    sup = scanner.scan(packed, n, [0], None, cap=48_000_000)
    ks = keys(sup)
    near = np.setdiff1d(ks, kb, assume_unique=True)
    assert np.all(np.isin(kb, ks)), "a hit is missing from the lowered-threshold superset"
    assert near.size > 1000
    for k in rng.choice(near, size=min(1500, near.size), replace=False):
        q, c = int(k >> np.uint64(16)), int(k & np.uint64(0xFFFF))
        s64 = fp64_score(pwms[c], codes, q)
        assert float(lowered[c]) <= s64 < float(thr[c]), (q, c, s64, lowered[c], thr[c])
    say(f"precision group {group}: {len(kb)} hits, max sampled score error {worst:.3g} (bound {err:.3g})")


# ---- batched form (20,000 sequences x 3,000 columns) ------------------------------

def per_sequence_hits(sc, codes, starts, lengths, cap_total):
    """One scan_device launch per sequence over device-resident, word-aligned
    copies of the sequences; hits of all launches in one device buffer.
    Returns global (position, column) keys and scores."""
    import torch
    lib = sc.lib
    woff = np.concatenate([[0], np.cumsum((lengths + 31) // 32)]).astype(np.int64)
    host = np.zeros(int(woff[-1]) + 16, np.uint64)
    for r in range(len(lengths)):
        s0 = int(starts[r])
        host[woff[r]:woff[r + 1]] = ba.pack_codes(codes[s0:s0 + int(lengths[r])])
    d_all = torch.from_numpy(host.view(np.int64)).to("cuda:0")
    d_hits = torch.empty(2 * cap_total, dtype=torch.int64, device="cuda:0")
    torch.cuda.synchronize()
    import ctypes
    cnt = ctypes.c_uint64()
    at, owner = 0, []
    for r in range(len(lengths)):
        for attempt in range(2):
            rc = lib.blamm_hip_scan_device(sc.ctx, d_all.data_ptr() + 8 * int(woff[r]), host.size - int(woff[r]),
                                           None, 0, int(lengths[r]), None, 0, d_hits.data_ptr() + 16 * at,
                                           cap_total - at, None, None)
            assert rc == 0, lib.blamm_hip_last_error().decode()
            rc = lib.blamm_hip_scan_finish(sc.ctx, ctypes.byref(cnt))
            # a candidate-scratch overflow grows the scratch: run once more
            if rc == ba.BLAMM_ERR_OVERFLOW and cnt.value == 0 and attempt == 0:
                continue
            assert rc == 0, lib.blamm_hip_last_error().decode()
            break
        owner.append((at, int(cnt.value)))
        at += int(cnt.value)
    h = d_hits[: 2 * at].cpu().numpy().view(ba.HIT_DTYPE).copy()
    pos = h["pos"].astype(np.uint64)
    for r, (o, k) in enumerate(owner):
        pos[o:o + k] += np.uint64(int(starts[r]))
    h["pos"] = pos
    return h


def test_batched_config(scanner, orc):
    """The batched workload: all 20,000 sequences as one request whose runs
    are the sequences (one captured launch sequence) and one launch per
    sequence give the same hit set and scores; the first 1,000 sequences as
    1,000 requests of one batch give the same hits per sequence; 25 sampled
    sequences equal the FP64 oracle."""
    say("batched workload: generating")
    ms, codes, starts, lengths = synth.config3(seed=3, nseq=20_000)
    n = int(codes.size)
    bg = synth.counts_of(codes)
    pwms, thr = library(ms, bg, pvalue=1e-4)
    P, L, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, L, thr, ba.PREC_BF16X2)
    _, err = scanner.bounds()
    words = ba.pack_codes(codes)
    say(f"batched: one request, {len(lengths)} runs, {n} bases")
    b = ba.Batch(scanner, max_bases=n + 4096, max_requests=1000)
    b.submit([ba.ScanRequest(words, n, starts, None, None)])
    b.wait()
    h_one = b.hits(0)
    k_one = keys(h_one, 20)
    assert len(k_one) > 1_000_000
    say("batched: one launch per sequence")
    h_seq = per_sequence_hits(scanner, codes, starts, lengths, len(h_one) + 4096)
    assert np.array_equal(k_one, keys(h_seq, 20))
    o1 = np.lexsort((h_one["col"], h_one["pos"]))
    o3 = np.lexsort((h_seq["col"], h_seq["pos"]))
    assert np.array_equal(h_one["score"][o1], h_seq["score"][o3])
    say("batched: 1,000 sequences as 1,000 requests of one batch")
    nreq = 1000
    reqs = []
    for r in range(nreq):
        s0 = int(starts[r])
        reqs.append(ba.ScanRequest(ba.pack_codes(codes[s0:s0 + int(lengths[r])]), int(lengths[r]), [0], None, None))
    b.submit(reqs)
    total = b.wait()
    parts = []
    for r in range(nreq):
        hr = b.hits(r)
        hr["pos"] = hr["pos"] + np.uint64(int(starts[r]))
        parts.append(hr)
    b.close()
    h_many = np.concatenate(parts)
    end = int(starts[nreq])
    sel = h_one["pos"] < end
    assert len(h_many) == total and total == int(sel.sum())
    assert np.array_equal(keys(h_one[sel], 20), keys(h_many, 20))
    rng = np.random.default_rng(33)
    for r in rng.choice(len(lengths), size=25, replace=False):
        s0, ln = int(starts[r]), int(lengths[r])
        text = synth.CODE2BASE[codes[s0:s0 + ln]].tobytes()
        ref = orc.scan_fp64(text, [0], pwms, thr)
        want = {(int(p) + s0, int(c)) for p, c in zip(ref["pos"], ref["col"])}
        sel = (h_one["pos"] >= s0) & (h_one["pos"] < s0 + ln)
        got = set(zip(h_one["pos"][sel].tolist(), h_one["col"][sel].tolist()))
        assert got == want, (r, len(got ^ want))
    check_sampled(h_one, codes, pwms, thr, err, rng)


# ---- serving form (2,000-motif library, mixed requests) -----------------------------

def digest(h):
    h = h[np.lexsort((h["col"], h["pos"]))]
    return hashlib.sha256(h.tobytes()).hexdigest(), len(h)


def test_serving_config(scanner, orc):
    """A server over the resident 2,000-motif library (4,000 columns) serves 12
    requests drawn from the serving distribution (seeded; a draw whose
    expected hit list exceeds 4M hits is redrawn to bound host memory) in a
    mix: each result is bit-identical to the same request served alone, and
    its hits on a prefix of its sequences equal the FP64 oracle under the
    request's own column subset and p-value."""
    ms = synth.config4_library(seed=4)
    rng = np.random.default_rng(44)
    bg = synth.counts_of(synth.bases(1_000_000, 0.41, rng))
    pvals = [1e-3, 1e-4, 1e-5, 1e-6]
    pwms, thr0 = library(ms, bg, pvalue=1e-4)
    thr_p = {p: library(ms, bg, pvalue=p)[1] for p in pvals}
    P, L, _ = ba.build_pattern_matrix(pwms)
    ncols = len(pwms)
    reqs = []
    while len(reqs) < 12:
        codes, starts, subset, p = synth.config4_request(rng, len(ms.pfms))
        cols = np.sort(np.concatenate([2 * subset, 2 * subset + 1]))
        if 5 * p * codes.size * cols.size > 4_000_000:
            continue
        t = np.full(ncols, np.inf, np.float32)
        t[cols] = thr_p[p][cols]
        reqs.append((codes, starts, cols, p, t))
    say(f"serving: {len(reqs)} requests, {sum(r[0].size for r in reqs)} bases, p {[r[3] for r in reqs]}")
    scanner.set_library(P, L, thr0, ba.PREC_BF16X2)
    _, err = scanner.bounds()

    def request(r):
        codes, starts, cols, p, t = r
        return ba.ScanRequest(ba.pack_codes(codes), int(codes.size), starts, None, t)

    srv = ba.Server(scanner, hbm_budget_bytes=2 << 30, slots=2, max_batch_requests=8,
                    max_batch_bases=8_000_000)
    try:
        say("serving: mix")
        objs = [request(r) for r in reqs]
        ids = [srv.submit(o) for o in objs]
        mixed, in_batch = [], []
        for k, rid in enumerate(ids):
            assert srv.wait(rid, 120_000)
            h, info = srv.result(rid)
            assert info.status == 0
            in_batch.append(info.batch_requests)
            mixed.append(digest(h))
            codes, starts, cols, p, t = reqs[k]
            assert set(np.unique(h["col"]).tolist()) <= set(cols.tolist())
            # oracle on the whole sequences of a prefix of at most 20 kbp
            # (windows never span a sequence start), or on the first 20 kbp
            # of the first sequence
            ends = np.append(starts[1:], codes.size).astype(np.int64)
            whole = bool(np.any(ends <= 20_000))
            m = int(ends[ends <= 20_000].max()) if whole else 20_000
            text = synth.CODE2BASE[codes[:m]].tobytes()
            rs = [int(s) for s in starts if s < m]
            ref = orc.scan_fp64(text, rs, [pwms[c] for c in cols], t[cols])
            want = {(int(q), int(cols[c])) for q, c in zip(ref["pos"], ref["col"])}
            lmax = int(L.max())
            if whole:       # every window of these sequences lies inside the prefix
                sel = h["pos"] < m
            else:           # a cut inside the first sequence: windows clear of the cut
                sel = h["pos"] + lmax <= m
                want = {x for x in want if x[0] + lmax <= m}
            got = set(zip(h["pos"][sel].tolist(), h["col"][sel].tolist()))
            assert got == want, (k, p, len(got ^ want))
            if len(h):
                check_sampled(h, codes, pwms, t, err, rng, n=300)
        assert max(in_batch) > 1
        say("serving: alone")
        for k, r in enumerate(reqs):
            o = request(r)
            rid = srv.submit(o)
            assert srv.wait(rid, 120_000)
            h, info = srv.result(rid)
            assert info.status == 0 and info.batch_requests == 1
            assert digest(h) == mixed[k], f"request {k} differs alone vs in the mix"
    finally:
        srv.close()
