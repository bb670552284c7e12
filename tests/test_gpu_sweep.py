"""Randomised sweep of small cases against the oracle (needs the GPU).

Each seed draws a stream length (including sizes below one tile and around
tile and word boundaries), dense run breaks (runs shorter than the motifs
included), lowercase density, motif lengths 1..64, thresholds from far below
to above every score and +inf (a skipped column), then checks the scan in both
precisions and the empirical histograms count for count.
"""
import os
import sys

import numpy as np
import pytest

import blamm_amd as ba

pytestmark = pytest.mark.gpu

# 32 seeds by default; BLAMM_SWEEP_SEEDS=first:count runs another range
_first, _count = (int(x) for x in os.environ.get("BLAMM_SWEEP_SEEDS", "0:32").split(":"))
SEEDS = list(range(_first, _first + _count))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def scanner():
    with ba.HipScanner(0) as sc:
        yield sc


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    if rng.random() < 0.5:
        n = int(rng.choice([1, 2, 31, 32, 33, 63, 64, 65, 500, 1024, 4095, 4096, 4097, 16_384]))
    else:
        n = int(rng.integers(1, 60_000))
    codes = rng.integers(0, 4, size=n)
    text = bytearray(np.frombuffer(b"ACGT", np.uint8)[codes].tobytes())
    low = float(rng.choice([0.0, 0.05, 0.5, 1.0]))
    if low:
        for i in np.nonzero(rng.random(n) < low)[0]:
            text[i] |= 0x20
    nbreak = int(rng.integers(0, max(1, n // int(rng.choice([3, 20, 200, 5000])) + 1)))
    runs = np.unique(np.concatenate([[0], rng.integers(0, max(n, 1), size=nbreak)])).astype(np.uint64)
    ncols = int(rng.integers(1, 90))
    pwms = []
    lmax = 65 if seed % 3 == 0 else 33
    for _ in range(ncols):
        L = int(rng.integers(1, lmax))
        kind = rng.integers(0, 3)
        if kind == 0:
            w = rng.normal(0, 1.5, size=(L, 4))
        elif kind == 1:
            w = rng.integers(-16, 17, size=(L, 4)) / 8.0       # exact sums
        else:
            w = np.log2((rng.dirichlet(np.full(4, 0.5), size=L) * 100 + 0.25) / 25.25)
        pwms.append(w.astype(np.float32))
    thr = []
    for w in pwms:
        mx, mn = float(w.max(1).sum()), float(w.min(1).sum())
        k = rng.integers(0, 5)
        thr.append(np.inf if k == 0 else mn - 1.0 if k == 1 else mx + 1.0 if k == 2 else
                   mn + (mx - mn) * float(rng.uniform(0.5, 0.98)))
    return bytes(text), runs, pwms, np.asarray(thr, np.float32), rng


@pytest.mark.parametrize("seed", SEEDS)
def test_sweep_scan_and_hist(scanner, orc, seed):
    text, runs, pwms, thr, rng = random_case(seed)
    say(f"sweep seed={seed} n={len(text)} runs={len(runs)} cols={len(pwms)}")
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    packed, lower = ba.pack(text)
    want = orc.scan_fp64(text, runs, pwms, thr, threads=4)
    want_set = set(zip(want["pos"].tolist(), want["col"].tolist()))
    for prec in (ba.PREC_BF16X2, ba.PREC_BF16X2_LO, ba.PREC_FP32):
        scanner.set_library(P, lengths, thr, prec)
        got = scanner.scan(packed, len(text), runs, lower)
        got_set = set(zip(got["pos"].tolist(), got["col"].tolist()))
        assert got_set == want_set, (seed, prec, len(got_set - want_set), len(want_set - got_set))
        assert len(got) == len(got_set)
        _, err = scanner.bounds()
        if len(got):
            a = np.lexsort((got["col"], got["pos"]))
            b = np.lexsort((want["col"], want["pos"]))
            assert np.abs(got["score"][a].astype(np.float64) - want["score64"][b]).max() <= err
    mn = np.array([ba.min_score(w) for w in pwms], np.float32)
    mx = np.array([ba.max_score(w) for w in pwms], np.float32)
    nbins = int(rng.choice([2, 7, 250, 300, 700]))
    scanner.set_library(P, lengths, np.zeros(len(pwms), np.float32), ba.PREC_BF16X2)
    qend = int(rng.integers(0, len(text) + 1)) if rng.random() < 0.3 else None
    got_h = scanner.hist(packed, len(text), mn, mx, nbins, runs, lower, qend=qend)
    want_h = orc.empirical_hist(text, runs, pwms, mn, mx, nbins, qend=qend, threads=4)
    assert (got_h == want_h).all(), (seed, nbins, np.nonzero((got_h != want_h).any(axis=1))[0][:5])
