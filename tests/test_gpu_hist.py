"""Empirical histograms (`blamm hist -e`) on the GPU against the oracle.

The engine's histograms (libblamm_hip.so through ctypes, or the `blamm`
executable) must equal the oracle's count for count: every valid window's
FP32 position-order score binned with the reference's float arithmetic.
"""
import filecmp
import os
import sys

import numpy as np
import pytest

import blamm_amd as ba
from blamm_amd import synth
from helpers import EXAMPLE, copy_example, run_blamm, run_oracle, stream_case

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def scanner():
    with ba.HipScanner(0) as sc:
        yield sc


def ranges(pwms):
    return (np.array([ba.min_score(w) for w in pwms], np.float32),
            np.array([ba.max_score(w) for w in pwms], np.float32))


def gpu_vs_oracle(sc, orc, text, runs, pwms, mn, mx, nbins, qend=None):
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    sc.set_library(P, lengths, np.zeros(len(pwms), np.float32), ba.PREC_BF16X2)
    packed, lower = ba.pack(text)
    got = sc.hist(packed, len(text), mn, mx, nbins, runs, lower, qend=qend)
    want = orc.empirical_hist(text, runs, pwms, mn, mx, nbins, qend=qend)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} columns differ, first {bad[:5]}: " \
                          f"{np.nonzero(got[bad[0]] != want[bad[0]])[0][:8]}"
    return got, sc.stats()


@pytest.mark.parametrize("seed,lower,lmin,lmax,nbins", [
    (1, 0, 5, 30, 250),      # LDS histograms, two column blocks per item
    (2, 13, 1, 32, 250),
    (3, 7, 4, 16, 300),      # LDS histograms, one block per item
    (4, 0, 6, 24, 1000),     # histograms in HBM
    (5, 101, 1, 8, 2),
    (6, 5, 30, 64, 250),     # motifs longer than 32 positions (their own launch)
    (7, 0, 1, 64, 1000),
])
def test_hist_matches_oracle(scanner, orc, seed, lower, lmin, lmax, nbins):
    say(f"hist seed={seed} lower/{lower} L={lmin}..{lmax} bins={nbins}")
    text, runs, bg = stream_case(seed, 150_000, lower_every=lower)
    ms = synth.motifs(70, lmin, lmax, 0.5, seed=seed)
    pwms = [ba.pwm(p, bg) for p in ms.pfms]
    mn, mx = ranges(pwms)
    got, st = gpu_vs_oracle(scanner, orc, text, runs, pwms, mn, mx, nbins)
    assert int(got.sum()) > 0
    say(f"  windows binned from their FP32 sum: {st.band_rescored}")


def test_hist_qend_and_empty(scanner, orc):
    """Windows starting at or after qend are not counted; an empty stream adds nothing."""
    text, runs, bg = stream_case(11, 40_000, lower_every=9)
    pwms = [ba.pwm(p, bg) for p in synth.motifs(20, 3, 25, 0.5, seed=11).pfms]
    mn, mx = ranges(pwms)
    for qend in (0, 1, 31, 32, 33, 12_345, len(text) - 3):
        gpu_vs_oracle(scanner, orc, text, runs, pwms, mn, mx, 250, qend=qend)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, np.zeros(len(pwms), np.float32), ba.PREC_BF16X2)
    got = scanner.hist(np.zeros(0, np.uint64), 0, mn, mx, 250)
    assert int(got.sum()) == 0


def test_hist_constructed_edges(scanner, orc):
    """Windows whose FP32 position-order sum and exact sum fall on opposite
    sides of a bin edge, a zero-width histogram (every window in bin 0),
    scores outside [min, max] (clamped) and dyadic weights whose scores sit
    exactly on bin edges."""
    rng = np.random.default_rng(7)
    n = 20_000
    codes = rng.integers(0, 4, size=n)
    codes[1000:1005] = 0
    text = bytearray(np.frombuffer(b"ACGT", np.uint8)[codes].tobytes())
    for i in rng.choice(n, size=800, replace=False):
        text[i] |= 0x20
    text = bytes(text)
    runs = np.array([0, 5000, 5001, 5005, 19_990], np.uint64)
    w_order = np.zeros((5, 4), np.float32)
    w_order[0, 0] = np.float32(1.0) - np.float32(2.0 ** -24)
    w_order[1:, 0] = np.float32(2.0 ** -26)
    w_flat = np.array([[0.5, -0.5, 0.25, 0.0]] * 3, np.float32)
    dyadic = [(rng.integers(-16, 17, size=(int(L), 4)) / 8.0).astype(np.float32) for L in (1, 4, 9, 17, 32)]
    pwms = [w_order, w_flat] + dyadic
    mn = np.array([0.0, 0.25] + [-3.0] * len(dyadic), np.float32)
    mx = np.array([2.0, 0.25] + [5.0] * len(dyadic), np.float32)
    got, _ = gpu_vs_oracle(scanner, orc, text, runs, pwms, mn, mx, 16)
    assert got[1, 0] == got[1].sum() > 0
    # the constructed FP32-order window at 1000 lands in bin 0 of width 1
    got2, _ = gpu_vs_oracle(scanner, orc, text[1000:1005], [0], [w_order], [0.0], [2.0], 2)
    assert got2[0].tolist() == [1, 0]


def test_hist_totals_at_scale(scanner, orc):
    """A larger stream: every column holds one count per valid window
    (closed form), and a sample of columns equals the oracle."""
    text, runs, bg = stream_case(21, 4_000_000, lower_every=0, nfrac=0.005)
    ms = synth.motifs(300, 5, 30, 0.5, seed=21)
    pwms = [ba.pwm(p, bg) for p in ms.pfms]
    mn, mx = ranges(pwms)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, np.zeros(len(pwms), np.float32), ba.PREC_BF16X2)
    packed, lower = ba.pack(text)
    got = scanner.hist(packed, len(text), mn, mx, 250, runs, lower)
    st = scanner.stats()
    n = len(text)
    starts = np.unique(np.concatenate([[0], np.asarray(runs, np.int64), [n]]))
    seglen = np.diff(starts)
    for c, w in enumerate(pwms):
        L = w.shape[0]
        assert int(got[c].sum()) == int(np.maximum(0, seglen - L + 1).sum()), c
    sample = [0, 1, 77, 150, 299]
    want = orc.empirical_hist(text, runs, [pwms[c] for c in sample], mn[sample], mx[sample], 250)
    assert (got[sample] == want).all()
    say(f"  4 Mbp x 300 columns: filter {st.filter_ms:.2f} ms, all {st.kernel_ms:.2f} ms, "
        f"windows binned from their FP32 sum: {st.band_rescored}")


def test_hist_needs_split_library(scanner):
    pwms = [np.ones((4, 4), np.float32)]
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, np.zeros(1, np.float32), ba.PREC_FP32)
    packed, _ = ba.pack(b"ACGTACGT")
    with pytest.raises(ba.BlammError):
        scanner.hist(packed, 8, [0.0], [4.0], 10)


def test_cli_hist_empirical_matches_oracle(tmp_path):
    """`blamm hist -e` writes the same .dat files as the oracle's command line
    (example inputs, default and small bin counts, a length limit)."""
    for args in (["-e"], ["-e", "-b", "40", "-l", "123456"]):
        a = copy_example(str(tmp_path / ("gpu" + "_".join(args))))
        b = copy_example(str(tmp_path / ("orc" + "_".join(args))))
        for d in (a, b):
            os.makedirs(os.path.join(d, "h"), exist_ok=True)
        run_blamm(["dict", "sequences.mf"], a)
        run_oracle(["dict", "sequences.mf"], b)
        run_blamm(["hist"] + args + ["-H", "h", "motifs.jaspar", "sequences.mf"], a)
        run_oracle(["hist"] + args + ["-H", "h", "motifs.jaspar", "sequences.mf"], b)
        fa = sorted(f for f in os.listdir(os.path.join(a, "h")) if f.endswith(".dat"))
        fb = sorted(f for f in os.listdir(os.path.join(b, "h")) if f.endswith(".dat"))
        assert fa == fb and len(fa) > 0
        for f in fa:
            assert filecmp.cmp(os.path.join(a, "h", f), os.path.join(b, "h", f), shallow=False), (args, f)
