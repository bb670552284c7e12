"""Checks that pin the oracle to the reference's semantics.

Each check takes an oracle build (oracle.Oracle) and raises AssertionError when
it disagrees with a closed form, a brute force or a hand-derived expectation.
They are run on the real oracle (tests/test_oracle_pinning.py) and on builds
with planted mistakes (tests/test_oracle_mutations.py), where each must fail.
"""
from __future__ import annotations

import ctypes
import itertools
import math
import os
import tempfile
from fractions import Fraction

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_libm.log2f.argtypes = [ctypes.c_float]
_libm.log2f.restype = ctypes.c_float

F = np.float32


# ---- PFM -> PWM (motif.cpp:194-223) ------------------------------------------

def check_pwm_values(orc) -> None:
    """Uniform PFM under a uniform background gives 0 bits exactly; a column
    (n,0,0,0) gives log2(((n+pc)/(n+4pc)) / ((m+pc)/(4m+4pc))) to float
    accuracy; unequal backgrounds change the weights the expected way."""
    pc = 0.25
    w = orc.pfm2pwm(np.full((3, 4), 7, np.uint64), [1000, 1000, 1000, 1000], pc)
    assert np.all(w == 0.0), w
    pfm = np.array([[100, 0, 0, 0], [0, 0, 0, 40], [10, 20, 30, 40]], np.uint64)
    bg = [400, 100, 200, 300]
    w = orc.pfm2pwm(pfm, bg, pc)
    tot_bg = sum(bg) + 4 * pc
    for p in range(3):
        tot = float(pfm[p].sum()) + 4 * pc
        for b in range(4):
            exact = math.log2(((float(pfm[p, b]) + pc) / tot) / ((bg[b] + pc) / tot_bg))
            assert abs(float(w[p, b]) - exact) <= 4e-7 * max(1.0, abs(exact)), (p, b, w[p, b], exact)


def check_pwm_bitwise(orc) -> None:
    """The float pipeline of PFM2PWM with the host log2f, evaluated in numpy
    float32 step by step (pins the rounding of each operation)."""
    rng = np.random.default_rng(11)
    for _ in range(20):
        L = int(rng.integers(1, 12))
        pfm = rng.integers(0, 300, size=(L, 4)).astype(np.uint64)
        bg = rng.integers(1, 10**9, size=4).astype(np.uint64)
        pc = F(rng.choice([0.25, 0.5, 1.0, 0.01]))
        for rc in (False, True):
            w = orc.pfm2pwm(pfm, bg, float(pc), rc)
            tot = F(int(bg.sum())) + F(4.0) * pc
            prob = [(F(int(x)) + pc) / tot for x in bg]
            if rc:
                prob = prob[::-1]
            for p in range(L):
                ct = F(int(pfm[p].sum())) + F(4.0) * pc
                for b in range(4):
                    ppm = (F(int(pfm[p, b])) + pc) / ct
                    want = _libm.log2f(float(ppm / prob[b]))
                    assert F(want) == w[p, b], (p, b, want, w[p, b])


def check_pwm_revcompl(orc) -> None:
    """W_rc[p][b] == W[L-1-p][3-b] bit for bit (motif.cpp:204-207, 267-286)."""
    rng = np.random.default_rng(5)
    for _ in range(20):
        L = int(rng.integers(1, 20))
        pfm = rng.integers(0, 200, size=(L, 4)).astype(np.uint64)
        bg = rng.integers(1, 10**7, size=4).astype(np.uint64)
        fwd = orc.pfm2pwm(pfm, bg, 0.25, False)
        rcp = orc.revcompl_pfm(pfm)
        assert np.array_equal(rcp, pfm[::-1, ::-1])
        rc = orc.pfm2pwm(rcp, bg, 0.25, True)
        assert np.array_equal(rc, fwd[::-1, ::-1]), (fwd, rc)


def check_min_max_bruteforce(orc) -> None:
    """max/min score equal the best/worst window score over all 4^L words."""
    rng = np.random.default_rng(3)
    for _ in range(10):
        L = int(rng.integers(1, 6))
        w = rng.normal(0, 2, size=(L, 4)).astype(np.float32)
        scores = [sum(float(w[p, b]) for p, b in enumerate(word)) for word in itertools.product(range(4), repeat=L)]
        assert abs(orc.max_score(w) - max(scores)) <= 1e-5 * L
        assert abs(orc.min_score(w) - min(scores)) <= 1e-5 * L


# ---- getScoreCutoff (motif.cpp:47-69) -----------------------------------------

def check_cutoff_closed_forms(orc) -> None:
    """One occupied bin i with N counts: cutoff = (i + 1 - p) * width + min
    (the bin coordinate rounded to float first); p beyond the mass returns max;
    uniform histograms interpolate linearly from the top."""
    B, mn, mx = 250, -10.0, 15.0
    width = F(mx - mn) / F(B)
    for i in (0, 17, 100, 249):
        for p in (1e-4, 0.25, 0.5, 1.0):
            counts = np.zeros(B, np.uint64)
            counts[i] = 123456
            got = orc.score_cutoff(counts, mn, mx, p)
            frac = float(F(p))
            want = F(F(frac * i + (1.0 - frac) * (i + 1)) * width + F(mn))
            assert got == want, (i, p, got, want)
    # walking down a uniform histogram: exact rational bin coordinate
    counts = np.full(B, 1000, np.uint64)
    for p in (0.001, 0.01, 0.3, 0.999):
        got = orc.score_cutoff(counts, mn, mx, p)
        x = Fraction(float(F(p))) * B            # bins consumed from the top
        coord = B - x                            # continuous bin coordinate
        want = float(coord) * float(width) + mn
        assert abs(got - want) <= 2e-5 * abs(mx - mn), (p, got, want)
    # an empty histogram gives 0/0 in the interpolation: NaN (no window passes)
    counts = np.zeros(B, np.uint64)
    assert math.isnan(orc.score_cutoff(counts, mn, mx, 0.5))


def check_cutoff_float_coordinate(orc) -> None:
    """The bin coordinate is rounded to float before scaling (motif.cpp:63-64).
    These histograms were found by search so that the float-coordinate and
    the double-coordinate evaluations round to different floats."""
    cases = [(-24.96710777282715, 11.744668006896973, 76, 409736, 1.2560315099108266e-06),
             (-15.270136833190918, 27.818889617919922, 125, 6066358, 0.023425258696079254),
             (-12.842764854431152, 5.839639186859131, 191, 7296554, 1.1276573786744848e-05)]
    for mn, mx, i, cnt, p in cases:
        counts = np.zeros(250, np.uint64)
        counts[i] = cnt
        got = orc.score_cutoff(counts, mn, mx, p)
        width = F(F(mx) - F(mn)) / F(250)
        frac = float(F(p)) * cnt / cnt
        coord = frac * i + (1.0 - frac) * (i + 1)
        want = F(F(coord) * width + F(mn))
        alt = F(coord * float(width) + float(F(mn)))
        assert want != alt
        assert got == want, (mn, mx, i, cnt, p, got, want)


# ---- theoretical histogram (motif.cpp:151-192, hist.cpp:162-175) ---------------
# Every grid score s sits on the boundary of bin s, so which of bins s-1, s
# it lands in is decided by float rounding; these checks hold for either.

def _grid(w, bg, nb):
    """Integer grid scores and their probabilities (exact enumeration)."""
    L = w.shape[0]
    mn = sum(min(float(x) for x in row) for row in w)
    mx = sum(max(float(x) for x in row) for row in w)
    a = nb / (mx - mn)
    b = -a * mn / L
    g = [[int(math.copysign(math.floor(abs(a * float(w[p, k]) + b) + 0.5), a * float(w[p, k]) + b))
          for k in range(4)] for p in range(L)]
    dist = {}
    for word in itertools.product(range(4), repeat=L):
        s = sum(g[p][k] for p, k in enumerate(word))
        pr = 1.0
        for k in word:
            pr *= float(bg[k])
        dist[s] = dist.get(s, 0.0) + pr
    return dist


def check_hist_support_and_counts(orc) -> None:
    """Nonzero bins lie at grid scores s (bin s or s-1); each holds
    (size_t)(maxLength * p_s) of a grid score that can reach it."""
    rng = np.random.default_rng(12)
    for _ in range(12):
        L = int(rng.integers(1, 4))
        w = np.round(rng.normal(0, 2, size=(L, 4)), 1).astype(np.float32)
        bg = np.array(rng.dirichlet(np.ones(4) * 2), np.float32)
        nb = int(rng.choice([20, 50, 250]))
        counts, mn, mx = orc.theoretical_hist(w, bg, nb, 10_000_000)
        dist = _grid(w, bg, nb)
        for j in np.nonzero(counts)[0]:
            cands = [int(10_000_000 * dist[s]) for s in dist if min(nb - 1, max(0, s)) == j or
                     min(nb - 1, max(0, s - 1)) == j]
            assert any(abs(int(counts[j]) - c) <= 1 for c in cands), (j, counts[j], cands)
        for s, pr in dist.items():
            if pr * 10_000_000 >= 2:
                near = {min(nb - 1, max(0, s)), min(nb - 1, max(0, s - 1)), min(nb - 1, max(0, s + 1))}
                assert any(counts[j] > 0 for j in near), (s, pr)


def check_hist_store_not_add(orc) -> None:
    """With 2 bins and many grid scores per bin the counts keep only one
    score's mass per bin (motif.h:109-115): far below maxLength in total."""
    w = np.array([[0.0, 1.0, 2.0, 3.0], [0.0, 1.0, 2.0, 3.0], [0.0, 0.5, 1.0, 1.5]], np.float32)
    bg = np.array([0.25, 0.25, 0.25, 0.25], np.float32)
    counts, mn, mx = orc.theoretical_hist(w, bg, 2, 1_000_000)
    dist = _grid(w, bg, 2)
    assert counts.sum() <= 2 * int(1_000_000 * max(dist.values())) + 2, counts
    assert counts.sum() < 990_000, counts   # adding would keep ~all of the 1e6 mass


# ---- FASTA filtering (sequence.cpp:142-293) -------------------------------------

FASTA_CASES = [
    # (name, [file texts], max_len, stream, [(start, record, pos)], records)
    ("gap_and_lines", [b">s1 desc\nACGTNNacgt\nAC\n>s2\nNNNA\n"], None,
     b"ACGTacgtACA", [(0, 0, 0), (4, 0, 6), (10, 1, 3)], 2),
    ("crlf", [b">s\r\nAC\r\nGT\r\n"], None, b"ACGT", [(0, 0, 0), (2, 0, 3)], 1),
    ("empty_lines", [b">s\n\nAC\n\nGT\n"], None, b"ACGT", [(0, 0, 0)], 1),
    ("empty_record", [b">a\nNNN\n>b\nAC\n"], None, b"AC", [(0, 1, 0)], 2),
    ("records_adjacent", [b">a\nAC\n>b\nGT\n"], None, b"ACGT", [(0, 0, 0), (2, 1, 0)], 2),
    # the next record's first base has the position the previous run would
    # continue with: still a new run (different record)
    ("record_same_position", [b">a\nAC\n>b\nNNGT\n"], None, b"ACGT", [(0, 0, 0), (2, 1, 2)], 2),
    ("truncate", [b">s\nACGTA\n"], 3, b"ACG", [(0, 0, 0)], 1),
    ("two_files", [b">a\nAC", b">b\nGT\n"], None, b"ACGT", [(0, 0, 0), (2, 1, 0)], 2),
    ("iupac", [b">s\nACRYGT\nAAA\n"], None, b"ACGTAAA", [(0, 0, 0), (2, 0, 4)], 1),
    ("lower_valid", [b">s\nacgtNA\n"], None, b"acgtA", [(0, 0, 0), (4, 0, 5)], 1),
]


def check_fasta_cases(orc) -> None:
    with tempfile.TemporaryDirectory() as d:
        for name, texts, max_len, stream, runs, nrec in FASTA_CASES:
            files = []
            for i, t in enumerate(texts):
                fn = os.path.join(d, f"{name}_{i}.fa")
                with open(fn, "wb") as f:
                    f.write(t)
                files.append(fn)
            s, rs, rq, rp, ns = orc.filter_fasta(files, max_len if max_len is not None else (1 << 64) - 1)
            got = list(zip(rs.tolist(), rq.tolist(), rp.tolist()))
            assert s == stream, (name, s, stream)
            assert got == runs, (name, got, runs)
            assert ns == nrec, (name, ns, nrec)
        fn = os.path.join(d, "bad.fa")
        with open(fn, "wb") as f:
            f.write(b"ACGT\n>s\nAC\n")
        try:
            orc.filter_fasta([fn])
        except RuntimeError as e:
            assert "fasta" in str(e)
        else:
            raise AssertionError("sequence before the first record must fail")


# ---- the scan predicate ----------------------------------------------------------

def brute_force_hits(stream: bytes, run_starts, pwms, thr):
    """Pure-Python FP64 scan: window at q for column c is scored iff it ends
    before the next run start; uppercase bases add their weight, lowercase 0;
    hit iff score >= threshold."""
    idx = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}
    n = len(stream)
    starts = sorted(set(int(x) for x in run_starts) | {0})
    nxt = [n] * n
    j = len(starts)
    nb = n
    for q in range(n - 1, -1, -1):
        while j > 0 and starts[j - 1] > q:
            nb = starts[j - 1]
            j -= 1
        nxt[q] = nb
    out = set()
    for c, w in enumerate(pwms):
        L = w.shape[0]
        for q in range(n):
            if q + L > nxt[q]:
                continue
            s = 0.0
            for p in range(L):
                b = idx.get(stream[q + p])
                if b is not None:
                    s += float(w[p, b])
            if s >= float(thr[c]):
                out.add((q, c))
    return out


def _random_case(seed, n=1500, ncols=6, lower=True):
    rng = np.random.default_rng(seed)
    codes = rng.integers(0, 4, size=n)
    text = bytearray(np.frombuffer(b"ACGT", np.uint8)[codes].tobytes())
    if lower:
        for i in rng.choice(n, size=n // 10, replace=False):
            text[i] |= 0x20
    runs = np.unique(np.concatenate([[0], rng.choice(n, size=12, replace=False)])).astype(np.uint64)
    pwms = [rng.normal(0, 1.5, size=(int(rng.integers(1, 12)), 4)).astype(np.float32) for _ in range(ncols)]
    thr = [float(np.quantile([w.max(1).sum() for w in pwms], 0.1)) - 4.0 for _ in range(ncols)]
    thr = [F(w.max(1).sum() * 0.55) for w in pwms]
    return bytes(text), runs, pwms, np.asarray(thr, np.float32)


def check_scan_bruteforce(orc) -> None:
    for seed in range(4):
        text, runs, pwms, thr = _random_case(seed)
        want = brute_force_hits(text, runs, pwms, thr)
        got = orc.scan_fp64(text, runs, pwms, thr, threads=3)
        assert set(zip(got["pos"].tolist(), got["col"].tolist())) == want, seed


def check_scan_equality_threshold(orc) -> None:
    """Weights on a 1/8 grid make window sums exact: a threshold equal to a
    window's sum must report it (>=), one ulp above must not."""
    rng = np.random.default_rng(2)
    n = 400
    codes = rng.integers(0, 4, size=n)
    text = np.frombuffer(b"ACGT", np.uint8)[codes].tobytes()
    w = (rng.integers(-16, 17, size=(7, 4)) / 8.0).astype(np.float32)
    q = 123
    exact = sum(float(w[p, codes[q + p]]) for p in range(7))
    hits = orc.scan_fp64(text, [0], [w], np.array([exact], np.float32), threads=2)
    assert q in set(hits["pos"].tolist())
    above = np.nextafter(np.float32(exact), np.float32(np.inf))
    hits = orc.scan_fp64(text, [0], [w], np.array([above], np.float32), threads=2)
    assert q not in set(hits["pos"].tolist())


def check_scan_revcompl_symmetry(orc) -> None:
    """The rc column at q scores the reverse complement of the window, i.e.
    the forward column at the mirrored position of the rc stream."""
    rng = np.random.default_rng(4)
    n = 3000
    codes = rng.integers(0, 4, size=n)
    text = np.frombuffer(b"ACGT", np.uint8)[codes].tobytes()
    rc_text = np.frombuffer(b"ACGT", np.uint8)[3 - codes[::-1]].tobytes()
    pfm = rng.integers(0, 100, size=(9, 4)).astype(np.uint64)
    bg = np.bincount(codes, minlength=4).astype(np.uint64)
    fwd = orc.pfm2pwm(pfm, bg, 0.25, False)
    rcw = orc.pfm2pwm(orc.revcompl_pfm(pfm), bg, 0.25, True)
    thr = np.array([fwd.max(1).sum() * 0.4], np.float32)
    a = orc.scan_fp64(text, [0], [rcw], thr, threads=2)
    b = orc.scan_fp64(rc_text, [0], [fwd], thr, threads=2)
    L = 9
    mirrored = set((n - L - int(p)) for p in b["pos"])
    assert set(a["pos"].tolist()) == mirrored


def check_scan_lowercase_zero(orc) -> None:
    """A lowercase base contributes 0 (all-zero one-hot row), not its weight."""
    w = np.array([[1.0, 2.0, 3.0, 4.0]] * 3, np.float32)
    text = b"AcG"
    hits = orc.scan_fp64(text, [0], [w], np.array([4.0], np.float32), threads=1)
    assert set(hits["pos"].tolist()) == {0}      # 1 + 0 + 3 = 4 >= 4
    hits = orc.scan_fp64(text, [0], [w], np.array([4.5], np.float32), threads=1)
    assert len(hits) == 0                        # would be 6 if c counted


# ---- empirical histograms (hist -e; hist.cpp:70-138, motif.h:58-66, 96-102) ------

def _valid_windows(n, run_starts, L, qend=None):
    """Window starts q whose L bases end at or before the next run start."""
    starts = sorted(set(int(x) for x in run_starts) | {0}) + [n]
    out = []
    for a, b in zip(starts[:-1], starts[1:]):
        out.extend(range(a, max(a, b - L + 1)))
    return [q for q in out if qend is None or q < qend]


def check_ehist_single_base(orc) -> None:
    """L = 1: every base is a window; uppercase bases land in the bin of their
    weight, lowercase ones in the bin of 0 (closed form from base counts)."""
    rng = np.random.default_rng(31)
    n = 5000
    codes = rng.integers(0, 4, size=n)
    text = bytearray(np.frombuffer(b"ACGT", np.uint8)[codes].tobytes())
    low = rng.choice(n, size=400, replace=False)
    for i in low:
        text[i] |= 0x20
    w = np.array([[-3.0, -1.0, 1.0, 3.0]], np.float32)
    nb = 8                                        # width 1 on [-4, 4]: bins 1, 3, 5, 7; zero -> bin 4
    got = orc.empirical_hist(bytes(text), [0, 17, 4000], [w], [-4.0], [4.0], nb, threads=1)[0]
    upper = np.ones(n, bool)
    upper[low] = False
    want = np.zeros(nb, np.uint64)
    for b, bin_ in zip(range(4), (1, 3, 5, 7)):
        want[bin_] = np.count_nonzero((codes == b) & upper)
    want[4] = len(low)
    assert got.tolist() == want.tolist(), (got, want)


def check_ehist_dyadic_bruteforce(orc) -> None:
    """Weights on a 1/8 grid and a histogram of width 1/2: every sum, bin
    coordinate and edge is exact (coordinates fall on quarters, so truncation
    and rounding differ), and the bins follow from exact rational arithmetic
    (pure Python).  Includes N-run breaks, lowercase bases and bins beyond the
    histogram range on both sides (clamped)."""
    idx = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}
    for seed in range(3):
        rng = np.random.default_rng(40 + seed)
        n = 1200
        codes = rng.integers(0, 4, size=n)
        text = bytearray(np.frombuffer(b"ACGT", np.uint8)[codes].tobytes())
        for i in rng.choice(n, size=n // 8, replace=False):
            text[i] |= 0x20
        text = bytes(text)
        runs = np.unique(np.concatenate([[0], rng.choice(n, size=10, replace=False)])).astype(np.uint64)
        pwms = [(rng.integers(-16, 17, size=(int(rng.integers(1, 11)), 4)) / 8.0).astype(np.float32)
                for _ in range(5)]
        # ranges narrower than the score range so both clamps occur
        lo = [Fraction(int(rng.integers(-24, -4)), 8) for _ in pwms]
        nbs = [int(rng.integers(5, 30)) for _ in pwms]
        his = [lo[c] + Fraction(nbs[c], 2) for c in range(len(pwms))]
        qend = n - 37
        for c, w in enumerate(pwms):
            got = orc.empirical_hist(text, runs, [w], [float(lo[c])], [float(his[c])], nbs[c], qend=qend,
                                     threads=1)[0]
            want = [0] * nbs[c]
            L = w.shape[0]
            for q in _valid_windows(n, runs, L, qend):
                s = Fraction(0)
                for p in range(L):
                    b = idx.get(text[q + p])
                    if b is not None:
                        s += Fraction(float(w[p, b]))
                k = math.floor((s - lo[c]) * 2)     # int() truncates; negative values clamp to 0 anyway
                want[min(max(k, 0), nbs[c] - 1)] += 1
            assert got.tolist() == want, (seed, c)


def check_ehist_total_windows(orc) -> None:
    """Each histogram holds exactly one count per valid window: the sum over
    runs of max(0, run length - L + 1), for windows starting before qend."""
    rng = np.random.default_rng(52)
    n = 30_000
    codes = rng.integers(0, 4, size=n)
    text = np.frombuffer(b"ACGT", np.uint8)[codes].tobytes()
    runs = np.unique(np.concatenate([[0], rng.choice(n, size=300, replace=False),
                                     np.arange(1000, 1030)])).astype(np.uint64)
    pwms = [rng.normal(0, 1.5, size=(L, 4)).astype(np.float32) for L in (1, 2, 7, 16, 30)]
    mn = [orc.min_score(w) for w in pwms]
    mx = [orc.max_score(w) for w in pwms]
    for qend in (n, 12_345):
        got = orc.empirical_hist(text, runs, pwms, mn, mx, 250, qend=qend, threads=2)
        for c, w in enumerate(pwms):
            assert int(got[c].sum()) == len(_valid_windows(n, runs, w.shape[0], qend)), (qend, c)


def check_ehist_fp32_order(orc) -> None:
    """The score binned is the FP32 sum in position order: for weights
    1 - 2^-24, 2^-26, 2^-26, 2^-26, 2^-26 each FP32 add rounds back to
    1 - 2^-24 (bin 0 of width 1), while the exact sum is 1.0 (bin 1)."""
    w = np.full((5, 4), 0.0, np.float32)
    w[0, 0] = np.float32(1.0) - np.float32(2.0 ** -24)
    w[1:, 0] = np.float32(2.0 ** -26)
    got = orc.empirical_hist(b"AAAAA", [0], [w], [0.0], [2.0], 2, threads=1)[0]
    assert got.tolist() == [1, 0], got


def check_ehist_width_zero(orc) -> None:
    """min == max gives width 0: (score - min) / 0 is +-inf or NaN, which the
    x86 int conversion turns into INT_MIN, clamped to bin 0."""
    w = np.array([[0.5, -0.5, 0.25, 0.0]] * 3, np.float32)
    got = orc.empirical_hist(b"ACGTTGCAAC", [0], [w], [0.25], [0.25], 10, threads=1)[0]
    assert got.tolist() == [8] + [0] * 9, got
