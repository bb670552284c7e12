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
# it lands in is decided by float rounding; the two invariant checks below
# hold for either.  check_hist_dyadic_exact fixes the bin and the count of
# every grid score on designs where every float operation is exact.

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


def _round_half_away(x: Fraction) -> int:
    """std::round: nearest integer, halves away from zero."""
    f = math.floor(abs(x) + Fraction(1, 2))
    return f if x >= 0 else -f


def _exact_theoretical(w, bg, nb, max_length):
    """The theoretical histogram from its definition in exact arithmetic, for
    designs where every float step of the reference is exact: the affine map
    a = nb / (max - min), b = -a * min / L rounds each weight to a grid
    integer (halves away from zero), the grid sums' probabilities are the
    convolution of the background, a grid score s maps back to the score
    (s - L b) / a, whose bin coordinate (score - min) / width is exact; bins
    are clamped to [0, nb - 1], a later (higher) score overwrites an earlier
    one in its bin, and the count is floor(max_length * probability)."""
    L = len(w)
    W = [[Fraction(float(x)) for x in row] for row in w]
    B = [Fraction(float(x)) for x in bg]
    mn = sum(min(r) for r in W)
    mx = sum(max(r) for r in W)
    a = Fraction(nb) / (mx - mn)
    b = -a * mn / L
    g = [[_round_half_away(a * x + b) for x in row] for row in W]
    dist = {}
    for word in itertools.product(range(4), repeat=L):
        s = sum(g[p][k] for p, k in enumerate(word))
        pr = Fraction(1)
        for k in word:
            pr *= B[k]
        dist[s] = dist.get(s, 0) + pr
    width = (mx - mn) / nb
    counts = [0] * nb
    # max_length enters the product as a float (size_t * float converts the
    # integer); the designs keep that product exact
    mf = Fraction(float(np.float32(max_length)))
    for s in sorted(dist):
        score = (s - L * b) / a
        coord = (score - mn) / width
        j = min(nb - 1, max(0, math.floor(coord)))
        prod = mf * dist[s]
        assert Fraction(float(np.float32(float(prod)))) == prod, "design leaves the float product inexact"
        counts[j] = math.floor(prod)
    return counts, float(mn), float(mx)


# Designs whose float steps are all exact: weights on a quarter grid, ranges
# and backgrounds dyadic, nb / range and min / L dyadic, max_length * p exact
# in float (or, for the last case, the float product differs from the double
# one).  (weights, background, bins, max_length, expected counts by hand)
HIST_DYADIC = [
    # a = 1, b = 0: grid = weights; s = 8 clamps into bin 7 and overwrites
    # s = 7 (store, not add); counts truncate 15.625, 31.25, 46.875, 62.5 ...
    ([[0, 1, 2, 4], [0, 1, 3, 4]], [0.125, 0.125, 0.25, 0.5], 8, 1000,
     [15, 31, 46, 62, 156, 187, 125, 250]),
    # a = 2, b = 1: a w + b = -0.5, 2.5 and 4.5 round away from zero
    # (-1, 3, 5; ties-to-even would give 0, 2, 4)
    ([[-1, -0.75, 0.75, 1], [0, 0.25, 1.75, 2]], [0.125, 0.125, 0.25, 0.5], 8, 1000,
     [31, 31, 0, 0, 281, 93, 0, 562]),
    # (float)100000007 = 100000008: the float product gives 12500001 where
    # the double product truncates 12500000.875 to 12500000
    ([[0, 1, 2, 4]], [0.125, 0.125, 0.25, 0.5], 4, 100000007,
     [12500001, 12500001, 25000002, 50000004]),
]


def check_hist_dyadic_exact(orc) -> None:
    """Bin and count of every grid score fixed by hand (HIST_DYADIC) and by
    exact-arithmetic brute force on random dyadic designs (L = 1, 2, 4;
    8 to 64 bins; clamping, collisions and half-integer grid points)."""
    for w, bg, nb, M, want in HIST_DYADIC:
        w = np.array(w, np.float32)
        counts, mn, mx = orc.theoretical_hist(w, np.array(bg, np.float32), nb, M)
        assert counts.tolist() == want, (w.tolist(), counts.tolist(), want)
        ex, emn, emx = _exact_theoretical(w, bg, nb, M)
        assert ex == want and (mn, mx) == (emn, emx), (ex, want)
    rng = np.random.default_rng(77)
    done = 0
    while done < 40:
        L = int(rng.choice([1, 2, 4]))
        k = int(rng.integers(1, 4))              # range = 2^k
        nb = int(rng.choice([8, 16, 32, 64]))
        if Fraction(nb, 2 ** k).denominator != 1 and nb < 2 ** k:
            continue
        # per position: min and max on the quarter grid, the spans summing to 2^k
        spans = rng.multinomial(4 * 2 ** k - L, np.ones(L) / L) + 1   # quarters, each >= 1
        rows = []
        for p in range(L):
            lo = int(rng.integers(-8, 5))
            hi = lo + int(spans[p])
            mid = rng.integers(lo, hi + 1, size=2)
            rows.append([lo / 4, hi / 4, float(mid[0]) / 4, float(mid[1]) / 4])
            rng.shuffle(rows[-1])
        mn = sum(Fraction(min(r)).limit_denominator(4) for r in rows)
        if (mn / L * Fraction(nb, 2 ** k)).denominator > 2 ** 20:
            continue
        bgq = rng.multinomial(16 - 4, np.ones(4) / 4) + 1          # sixteenths, each >= 1
        bg = [q / 16 for q in bgq]
        w = np.array(rows, np.float32)
        counts, _, _ = orc.theoretical_hist(w, np.array(bg, np.float32), nb, 1000)
        ex, _, _ = _exact_theoretical(w, bg, nb, 1000)
        assert counts.tolist() == ex, (rows, bg, nb, counts.tolist(), ex)
        done += 1


# ---- motif loaders (motif.cpp:323-437, motif.h:256-281) ---------------------------

# The reference's example/motifs.jaspar (tests/golden/example), typed from the
# file: rows A, C, G, T, one count per position.
EXAMPLE_JASPAR = {
    "MA0001.2": [[22, 16, 25, 27, 0, 0, 82, 40, 56, 35, 65, 25, 64, 0, 33, 52, 45, 21],
                 [29, 9, 8, 16, 92, 79, 1, 4, 0, 0, 1, 4, 0, 0, 14, 5, 23, 24],
                 [34, 29, 4, 13, 0, 0, 2, 3, 1, 0, 4, 3, 28, 92, 15, 7, 13, 26],
                 [10, 41, 58, 39, 3, 16, 10, 48, 38, 60, 25, 63, 3, 3, 33, 31, 14, 24]],
    "MA0002.2": [[287, 234, 123, 57, 0, 87, 0, 17, 10, 131, 500],
                 [496, 485, 1072, 0, 75, 127, 0, 42, 400, 463, 158],
                 [696, 467, 149, 7, 1872, 70, 1987, 1848, 251, 81, 289],
                 [521, 814, 656, 1936, 53, 1716, 13, 93, 1339, 1325, 1053]],
}


def _example_jaspar_path():
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "example", "motifs.jaspar")


def check_load_jaspar_example(orc) -> None:
    """The reference's example file: both motifs, shortest first (sorted by
    length, motif.cpp:436), counts as typed from the file (every position of
    MA0001.2 sums to 95)."""
    got = orc.load_motifs(_example_jaspar_path())
    assert [n for n, _ in got] == ["MA0002.2", "MA0001.2"], [n for n, _ in got]
    for name, pfm in got:
        want = np.array(EXAMPLE_JASPAR[name], np.uint64).T
        assert pfm.shape == want.shape and np.array_equal(pfm, want), name
    assert set(got[1][1].sum(axis=1).tolist()) == {95}


JASPAR_LAYOUTS = (">M1 first motif\nA [ 1 2 3 ]\nC [ 4 5 6 ]\nG [ 7 8 9 ]\nT [ 10 11 12 ]\n\n"
                  ">M2\tx\nA  [1 2 3 ]\nC  [4 5 6]\nG  [ 7 8 9]\nT  [10 11 12 ]\n"
                  ">M3\nA [ 5 ]\nC [ 0 ]\nG [ 0 ]\nT [ 95 ]")


def check_load_jaspar_layouts(orc) -> None:
    """Header names are the first token without '>'; each count row skips two
    tokens ("A" and "[", so "[1" swallows a count) and reads counts until the
    first token that is not one ("3]" still reads 3); the motif length is the
    A row's count; blank lines between motifs and a missing final newline
    are accepted."""
    want = {"M1": [[1, 4, 7, 10], [2, 5, 8, 11], [3, 6, 9, 12]],
            "M2": [[2, 5, 7, 11], [3, 6, 8, 12]],
            "M3": [[5, 0, 0, 95]]}
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "m.jaspar")
        with open(fn, "w") as f:
            f.write(JASPAR_LAYOUTS)
        got = orc.load_motifs(fn)
    assert [n for n, _ in got] == ["M3", "M2", "M1"], [n for n, _ in got]
    for name, pfm in got:
        assert pfm.tolist() == want[name], (name, pfm.tolist())


# (name, is a permutation, base name) by motif.h:256-281: a trailing run of
# digits (possibly empty) preceded by "__", at index >= 1
PERMUTATION_NAMES = [
    ("MA0001.2", False, "MA0001.2"), ("M__1", True, "M"), ("M__", True, "M"), ("M_1", False, "M_1"),
    ("__5", True, ""), ("_5", False, "_5"), ("5", False, "5"), ("", False, ""), ("A__B__12", True, "A__B"),
    ("A__12x", False, "A__12x"), ("M___7", True, "M_"), ("__", True, ""), ("X__007", True, "X"),
]


def check_permutation_names(orc) -> None:
    for name, perm, base in PERMUTATION_NAMES:
        assert orc.is_permutation(name) == perm, name
        assert orc.base_name(name) == base, name


CLUSTER_BUSTER = (">A1 first motif\n1 2 3 4\n5 6 7 8\n\n>A1__12\n0 0 0 9\n>B__\n1 1 1 1\n2 2 2 2\n3 3 3 3\n"
                  ">X_1\n9 8 7 6\n>__5\n4 4 4 4\n>_5\n3 3 3 3\n>Q99\n1 0 0 0\n")


def check_load_cluster_buster(orc) -> None:
    """Cluster-buster files (any name not ending in .jaspar): a '>' line names
    a motif with the whole rest of the line; each other non-empty line adds
    one position of four counts; blank lines are skipped; hist loads the
    motifs without permutations (perms = false), scan with them."""
    text = CLUSTER_BUSTER
    want = {"A1 first motif": [[1, 2, 3, 4], [5, 6, 7, 8]], "A1__12": [[0, 0, 0, 9]],
            "B__": [[1, 1, 1, 1], [2, 2, 2, 2], [3, 3, 3, 3]], "X_1": [[9, 8, 7, 6]], "__5": [[4, 4, 4, 4]],
            "_5": [[3, 3, 3, 3]], "Q99": [[1, 0, 0, 0]]}
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "motifs.cb")
        with open(fn, "w") as f:
            f.write(text)
        for perms, names in ((True, set(want)), (False, {"A1 first motif", "X_1", "_5", "Q99"})):
            got = orc.load_motifs(fn, perms)
            assert {n for n, _ in got} == names and len(got) == len(names), (perms, [n for n, _ in got])
            lens = [p.shape[0] for _, p in got]
            assert lens == sorted(lens), lens
            for name, pfm in got:
                assert pfm.tolist() == want[name], name


# ---- dict writer (dict.cpp:53-115; species.cpp:32-61, 84-103, 176-188) -------------

def check_dict_closed_form(orc) -> None:
    """`dict` on hand-built FASTA: groups in order of first appearance in the
    manifest, each group's files in lexicographic order (a std::set), ACGT
    counts case-folded over valid bases only, TOT_SEQ_LENGTH their sum,
    sequence names = the first token after '>' in file order (CRLF and leading
    blanks stripped), blank manifest lines skipped."""
    with tempfile.TemporaryDirectory() as d:
        mf, want = write_dict_case(d)
        assert orc.main(["dict", mf]) == 0
        got = open(mf + ".dict").read()
    assert got == want, got


def write_dict_case(d):
    """Hand-built FASTA files and manifest under d; returns the manifest path
    and the .dict text the reference writes for it (closed form)."""
    f0, f1, f2 = (os.path.join(d, x) for x in ("f0.fa", "f1.fa", "f2.fa"))
    with open(f1, "wb") as f:
        f.write(b">s1 first\nACGTacgtNNAC\nGT\n>s2\nnnnn\n> s3 x\nAAAA\n")
    with open(f0, "wb") as f:
        f.write(b">z9\r\nCCCC\r\n")
    with open(f2, "wb") as f:
        f.write(b">only\nggggtttt\n")
    mf = os.path.join(d, "sequences.mf")
    with open(mf, "w") as f:
        f.write(f"grpB\t{f1}\n\ngrpA\t{f2}\ngrpB\t{f0}\n")
    want = (f"NUM_SPECIES\t2\nSPECIES\tgrpB\nNUM_FASTA_FILES\t2\n{f0}\n{f1}\nTOT_SEQ_LENGTH\t20\n"
            f"NUCL_COUNT_ACGT\t7\t7\t3\t3\nNUM_SEQUENCES\t4\nz9\ns1\ns2\ns3\n"
            f"SPECIES\tgrpA\nNUM_FASTA_FILES\t1\n{f2}\nTOT_SEQ_LENGTH\t8\nNUCL_COUNT_ACGT\t0\t0\t4\t4\n"
            f"NUM_SEQUENCES\t1\nonly\n")
    return mf, want


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
