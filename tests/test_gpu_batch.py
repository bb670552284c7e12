"""Batched scans, tile-shape tuning and the serving loop against the oracle
(needs the GPU).

A batch packs many requests (streams of uneven length, each with its own
thresholds and column subset) into one captured launch sequence.  Every
request must give the hit set the FP64 oracle gives for that request alone,
scores within the engine's stated bound, and the same result alone as in a
mix.
"""
import os
import sys

import numpy as np
import pytest

import blamm_amd as ba
from blamm_amd import synth
from helpers import hit_set, library, stream_case

pytestmark = pytest.mark.gpu

PRECS = [ba.PREC_BF16X2, ba.PREC_BF16X2_LO, ba.PREC_FP32]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def scanner():
    with ba.HipScanner(0) as sc:
        yield sc


def make_library(seed, nmotifs=30, lmin=1, lmax=64, pvalue=1e-3):
    text, runs, bg = stream_case(seed, 20_000)
    ms = synth.motifs(nmotifs, lmin, lmax, 0.5, seed=seed)
    pwms, thr = library(ms, bg, pvalue=pvalue)
    return pwms, thr, bg


def make_requests(seed, pwms, base_thr, sizes, lower_every=(0, 0, 17)):
    """Requests of the given sizes; some lowercase, some with a column subset
    (+inf elsewhere), some with shifted thresholds, one with the library's."""
    rng = np.random.default_rng(seed)
    out = []
    for i, n in enumerate(sizes):
        le = lower_every[i % len(lower_every)]
        if n == 0:
            text, runs = b"", []
        else:
            text, runs, _ = stream_case(seed * 100 + i, n, lower_every=le)
        kind = i % 4
        if kind == 0:
            thr = None                          # the library's thresholds
        elif kind == 1:
            thr = base_thr.copy()
            keep = rng.random(len(thr)) < 0.3
            thr[~keep] = np.inf                 # column subset
        elif kind == 2:
            thr = (base_thr - np.float32(1.5)).astype(np.float32)   # looser threshold
        else:
            thr = base_thr.copy()
            thr[: len(thr) // 2] = np.inf
        out.append((text, runs, thr))
    return out


def to_request(text, runs, thr):
    packed, lower = ba.pack(text)
    return ba.ScanRequest(packed, len(text), runs, lower, thr)


def oracle_hits(orc, text, runs, pwms, thr):
    return orc.scan_fp64(text, runs, pwms, thr)


def assert_matches(hits, ref, err, what):
    got, want = hit_set(hits), hit_set(ref)
    assert got == want, (f"{what}: {len(got - want)} extra {sorted(got - want)[:4]}, "
                         f"{len(want - got)} missing {sorted(want - got)[:4]}")
    assert len(hits) == len(got), f"{what}: duplicate hits"
    if len(hits):
        a = np.lexsort((hits["col"], hits["pos"]))
        b = np.lexsort((ref["col"], ref["pos"]))
        diff = np.abs(hits["score"][a].astype(np.float64) - ref["score64"][b])
        assert diff.max() <= err, (what, diff.max(), err)


def sorted_hits(h):
    return h[np.lexsort((h["col"], h["pos"]))]


SIZES = [0, 1, 31, 64, 65, 1000, 4999, 20_000, 33, 50_000, 7, 12_345]


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("tuning", [(512, 2, 4), (64, 1, 8)])
def test_batch_matches_oracle_per_request(scanner, orc, prec, tuning):
    pwms, thr, _ = make_library(61)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, prec)
    scanner.set_tuning(*tuning)
    try:
        _, err = scanner.bounds()
        reqs = make_requests(62, pwms, thr, SIZES)
        b = ba.Batch(scanner, max_bases=1 << 16, max_requests=4)
        say(f"batch prec={prec} tuning={tuning} requests={len(reqs)} bases={sum(len(t) for t, _, _ in reqs)}")
        b.submit([to_request(*r) for r in reqs])
        total = b.wait()
        seen = 0
        for i, (text, runs, t) in enumerate(reqs):
            ref = oracle_hits(orc, text, runs, pwms, thr if t is None else t)
            h = b.hits(i)
            assert_matches(h, ref, err, f"request {i} (n={len(text)})")
            seen += len(h)
        assert seen == total
        assert total > 1000
        b.close()
    finally:
        scanner.set_tuning(512, 2, 4)


def test_batch_graph_replays_match_eager(scanner, orc):
    """The captured sequence is replayed for batches of different shapes;
    every replay equals the eager single-stream path, with eager scans and
    synchronisations in between."""
    pwms, thr, _ = make_library(71)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    b = ba.Batch(scanner, max_bases=1 << 20, max_requests=64)
    sets = [make_requests(72, pwms, thr, [3000, 40_000, 0, 900]),
            make_requests(73, pwms, thr, [25_000, 128, 60_000, 5, 7000, 31]),
            make_requests(74, pwms, thr, [10_000])]
    eager = []
    for reqs in sets:
        per = []
        for text, runs, t in reqs:
            scanner.set_library(P, lengths, thr if t is None else t, ba.PREC_BF16X2)
            packed, lower = ba.pack(text)
            per.append(sorted_hits(scanner.scan(packed, len(text), runs, lower)))
        eager.append(per)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    for rep in range(3):
        for k, reqs in enumerate(sets):
            say(f"replay {rep} set {k}")
            b.submit([to_request(*r) for r in reqs])
            b.wait()
            for i in range(len(reqs)):
                h = sorted_hits(b.hits(i))
                e = eager[k][i]
                assert np.array_equal(h["pos"], e["pos"]) and np.array_equal(h["col"], e["col"]), (rep, k, i)
                assert np.array_equal(h["score"], e["score"]), (rep, k, i)
            # eager work between replays
            text, runs, _ = reqs[0]
            packed, lower = ba.pack(text)
            scanner.scan(packed, len(text), runs, lower)
    b.close()


def test_batch_buffers_grow(scanner, orc):
    """A slot created tiny grows every buffer (and replays after candidate or
    hit overflow) without changing the result."""
    pwms, thr, _ = make_library(81, pvalue=1e-2)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    _, err = scanner.bounds()
    b = ba.Batch(scanner, max_bases=64, max_requests=1)
    before = b.device_bytes()
    reqs = make_requests(82, pwms, thr, [200_000, 3, 100_000])
    say("batch growth")
    b.submit([to_request(*r) for r in reqs])
    b.wait()
    for i, (text, runs, t) in enumerate(reqs):
        assert_matches(b.hits(i), oracle_hits(orc, text, runs, pwms, thr if t is None else t), err, f"request {i}")
    assert b.stats().candidates > 1 << 16
    assert b.device_bytes() > before
    b.close()


@pytest.mark.parametrize("prec", PRECS)
def test_tunings_agree_with_oracle(scanner, orc, prec):
    text, runs, bg = stream_case(91, 70_000, lower_every=29)
    ms = synth.motifs(45, 1, 64, 0.5, seed=91)
    pwms, thr = library(ms, bg, pvalue=1e-3)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, prec)
    ref = orc.scan_fp64(text, runs, pwms, thr)
    packed, lower = ba.pack(text)
    nbs = [1, 2] if prec != ba.PREC_FP32 else [1, 2, 4]
    try:
        for nb in nbs:
            for kp in (4, 8):
                for ct in (1, 7, 512):
                    say(f"tuning prec={prec} nb={nb} kpad={kp} chunk={ct}")
                    scanner.set_tuning(ct, nb, kp)
                    _, err = scanner.bounds()
                    assert_matches(scanner.scan(packed, len(text), runs, lower), ref, err, f"nb={nb} kp={kp} ct={ct}")
    finally:
        scanner.set_tuning(512, 2, 4)


def test_invalid_tunings_rejected(scanner):
    pwms, thr, _ = make_library(95)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    for bad in [(0, 2, 4), (512, 3, 4), (512, 4, 4), (512, 2, 6)]:
        with pytest.raises(ba.BlammError):
            scanner.set_tuning(*bad)
    t = scanner.tuning()
    assert (t.chunk_tiles, t.blocks_per_wave, t.kpad_positions) == (512, 2, 4)


def test_autotune_uses_cache(scanner, orc, tmp_path):
    import torch
    text, runs, bg = stream_case(97, 400_000)
    ms = synth.motifs(64, 5, 30, 0.5, seed=97)
    pwms, thr = library(ms, bg, pvalue=1e-4)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    packed, _ = ba.pack(text)
    n = len(text)
    pw = int(scanner.lib.blamm_hip_packed_words(n))
    buf = np.zeros(pw, np.uint64)
    buf[:len(packed)] = packed
    d_codes = torch.from_numpy(buf.view(np.int64)).cuda()
    d_runs = torch.from_numpy(np.asarray(runs, np.uint64).view(np.int64)).cuda()
    cache = str(tmp_path / "tune.txt")
    try:
        say("autotune (timing)")
        t1, ms1 = scanner.autotune(d_codes, None, n, d_runs, len(runs), cache)
        lines = open(cache).read().splitlines()
        assert len(lines) == 1 and lines[0].split()[1:4] == [str(t1.chunk_tiles), str(t1.blocks_per_wave),
                                                              str(t1.kpad_positions)]
        assert ms1 > 0
        scanner.set_tuning(512, 2, 4)
        say("autotune (cache)")
        t2, ms2 = scanner.autotune(d_codes, None, n, d_runs, len(runs), cache)
        assert t2.as_dict() == t1.as_dict() and abs(ms2 - ms1) <= 1e-3 * ms1
        assert len(open(cache).read().splitlines()) == 1
        assert scanner.tuning().as_dict() == t1.as_dict()
        _, err = scanner.bounds()
        assert_matches(scanner.scan(packed, n, runs, None), orc.scan_fp64(text, runs, pwms, thr), err, "tuned")
    finally:
        scanner.set_tuning(512, 2, 4)


def test_server_alone_equals_mix(scanner, orc):
    """Continuous batching: each request's result in a mix (several batches
    in flight, requests admitted as slots free up) is bit-identical to the
    same request served alone, and matches the oracle."""
    pwms, thr, _ = make_library(101)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    _, err = scanner.bounds()
    sizes = [5000, 120, 30_000, 0, 800, 64, 9000, 45_000, 3, 2500, 15_000, 700, 60_000, 1, 4096, 20_000]
    reqs = make_requests(102, pwms, thr, sizes)
    # a small budget forces several batches
    srv = ba.Server(scanner, hbm_budget_bytes=2 * 1_200_000, slots=2, max_batch_requests=5)
    try:
        say("server: alone")
        alone = []
        for r in reqs:
            rid = srv.submit(to_request(*r))
            assert srv.wait(rid, 60_000)
            h, info = srv.result(rid)
            assert info.status == 0 and info.batch_requests == 1
            alone.append(sorted_hits(h))
        say("server: mix")
        ids = [srv.submit(to_request(*r)) for r in reqs]
        mixed, sizes_in_batch = [], []
        for rid in ids:
            assert srv.wait(rid, 60_000)
            h, info = srv.result(rid)
            assert info.status == 0 and info.t_done_ns >= info.t_start_ns >= info.t_submit_ns
            sizes_in_batch.append(info.batch_requests)
            mixed.append(sorted_hits(h))
        st = srv.stats()
        assert st.completed == 2 * len(reqs) and st.batches >= len(reqs) + 3
        assert max(sizes_in_batch) > 1
        for i, (a, m) in enumerate(zip(alone, mixed)):
            assert np.array_equal(a, m), f"request {i} differs alone vs mix"
        for i in (0, 2, 7, 12):
            text, runs, t = reqs[i]
            assert_matches(mixed[i], oracle_hits(orc, text, runs, pwms, thr if t is None else t), err, f"request {i}")
    finally:
        srv.close()


def test_server_rejects_request_over_budget(scanner):
    pwms, thr, _ = make_library(111)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    srv = ba.Server(scanner, hbm_budget_bytes=200_000, slots=2)
    try:
        text, runs, _ = stream_case(112, 1_000_000)
        with pytest.raises(ba.BlammError) as e:
            srv.submit(to_request(text, runs, None))
        assert e.value.code == ba.BLAMM_ERR_ARG
    finally:
        srv.close()


def random_request(rng, pwms, base_thr, n):
    """One request of n bases: random GC, N-run density (dense for some, so
    run starts fall inside most windows), lowercase share, and a column subset
    or shifted thresholds."""
    if n == 0:
        text, runs = b"", []
    else:
        text, runs, _ = stream_case(int(rng.integers(1 << 30)), n,
                                    lower_every=int(rng.choice([0, 0, 3, 29])),
                                    nfrac=float(rng.choice([0.0, 0.01, 0.3])), gc=float(rng.uniform(0.3, 0.6)))
    r = rng.random()
    if r < 0.25:
        thr = None
    elif r < 0.6:
        thr = base_thr.copy()
        thr[rng.random(len(thr)) < rng.uniform(0.1, 0.9)] = np.inf
    else:
        thr = (base_thr - np.float32(rng.uniform(0.0, 3.0))).astype(np.float32)
    return text, runs, thr


BATCH_SWEEP = int(os.environ.get("BLAMM_BATCH_SWEEP", "6"))


@pytest.mark.parametrize("case", range(BATCH_SWEEP))
def test_batch_random_compositions(scanner, orc, case):
    """Random libraries (lengths 1-64), random request mixes (sizes from 0 to
    60 kbp, dense and sparse runs, lowercase, subsets, shifted thresholds) and
    random tile shapes, three batches in a row on one slot (so buffers grow
    and the captured sequence is replayed for other shapes): every request
    equals the FP64 oracle for that request alone."""
    rng = np.random.default_rng(9000 + case)
    prec = PRECS[case % 3]
    pwms, thr, _ = make_library(9100 + case, nmotifs=int(rng.integers(5, 40)), lmin=1,
                                lmax=int(rng.choice([12, 40, 64])), pvalue=float(rng.choice([1e-2, 1e-3])))
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, prec)
    nbmax = 4 if prec == ba.PREC_FP32 else 2
    tuning = (int(rng.choice([1, 16, 512])), int(rng.choice([b for b in (1, 2, 4) if b <= nbmax])),
              int(rng.choice([4, 8])))
    scanner.set_tuning(*tuning)
    b = ba.Batch(scanner, max_bases=1 << 12, max_requests=2)
    try:
        _, err = scanner.bounds()
        for rnd in range(3):
            sizes = [int(rng.choice([0, 1, 17, 64, 500, 5000, 60_000])) for _ in range(int(rng.integers(1, 10)))]
            reqs = [random_request(rng, pwms, thr, n) for n in sizes]
            say(f"batch sweep case={case} round={rnd} prec={prec} tuning={tuning} sizes={sizes}")
            b.submit([to_request(*r) for r in reqs])
            total = b.wait()
            seen = 0
            for i, (text, runs, t) in enumerate(reqs):
                h = b.hits(i)
                assert_matches(h, oracle_hits(orc, text, runs, pwms, thr if t is None else t), err,
                               f"case {case} round {rnd} request {i} (n={len(text)})")
                seen += len(h)
            assert seen == total
    finally:
        b.close()
        scanner.set_tuning(512, 2, 4)


def test_batch_with_more_requests_than_the_lds_tables(scanner, orc):
    """1,100 requests in one batch: more than finalize and the grouping kernels
    count in LDS (1,024), so they count with global atomics; every request
    equals the oracle and the single-request batch's hits."""
    pwms, thr, _ = make_library(131, nmotifs=12, lmin=4, lmax=20, pvalue=1e-2)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    rng = np.random.default_rng(132)
    reqs = []
    for i in range(1100):
        text, runs, _ = stream_case(5000 + i, int(rng.integers(1, 400)))
        reqs.append((text, runs, None if i % 3 else (thr - np.float32(0.5)).astype(np.float32)))
    b = ba.Batch(scanner, max_bases=1 << 20, max_requests=1100)
    try:
        _, err = scanner.bounds()
        say("batch of 1,100 requests")
        b.submit([to_request(*r) for r in reqs])
        total = b.wait()
        assert total > 1000
        seen = 0
        sel = list(range(0, 1100, 137))
        in_batch = {}
        for i, (text, runs, t) in enumerate(reqs):
            h = b.hits(i)
            seen += len(h)
            if i in sel:
                in_batch[i] = sorted_hits(h)
            if i % 50 == 0:
                assert_matches(h, oracle_hits(orc, text, runs, pwms, thr if t is None else t), err, f"request {i}")
        assert seen == total
        # the same requests one per batch give the same hits
        for i in sel:
            b.submit([to_request(*reqs[i])])
            b.wait()
            assert np.array_equal(sorted_hits(b.hits(0)), in_batch[i]), f"request {i} alone vs in the batch"
    finally:
        b.close()
