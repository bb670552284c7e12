"""The MI355X engine against the FP64 oracle (needs the GPU).

Every test calls the engine through the C ABI (libblamm_hip.so via ctypes, or
the `blamm` executable linked against it) and compares with the oracle on the
same seeded inputs.  The bar: the (position, column) hit set is identical to
the FP64 left-to-right oracle; reported scores are within the engine's stated
bound (blamm_hip_library_bounds) of the FP64 score.
"""
import os
import sys

import numpy as np
import pytest

import blamm_amd as ba
from blamm_amd import synth
from helpers import (EXAMPLE, assert_same_occurrences, copy_example, hit_set, library, read_occurrences,
                     run_blamm, run_oracle, stream_case)

pytestmark = pytest.mark.gpu

PRECS = [ba.PREC_BF16X2, ba.PREC_BF16X2_LO, ba.PREC_FP32]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def scanner():
    with ba.HipScanner(0) as sc:
        yield sc


def check_against_oracle(sc, orc, text, runs, pwms, thr, prec, lower=None, packed=None):
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    sc.set_library(P, lengths, thr, prec)
    if packed is None:
        packed, lower = ba.pack(text)
    hits = sc.scan(packed, len(text), runs, lower)
    ref = orc.scan_fp64(text, runs, pwms, thr)
    got, want = hit_set(hits), hit_set(ref)
    assert got == want, f"{len(got - want)} extra {sorted(got - want)[:4]}, {len(want - got)} missing {sorted(want - got)[:4]}"
    assert len(hits) == len(got), "duplicate hits"
    _, err = sc.bounds()
    if len(hits):
        a = np.lexsort((hits["col"], hits["pos"]))
        b = np.lexsort((ref["col"], ref["pos"]))
        diff = np.abs(hits["score"][a].astype(np.float64) - ref["score64"][b])
        assert diff.max() <= err, (diff.max(), err)
    return len(want)


@pytest.mark.parametrize("prec", PRECS)
@pytest.mark.parametrize("seed,lower,lmin,lmax,pval", [
    (1, 0, 5, 30, 1e-3), (2, 13, 1, 32, 1e-3), (3, 0, 6, 20, 1e-4), (4, 101, 10, 32, 1e-2), (5, 7, 1, 8, 1e-3),
    (6, 0, 33, 64, 1e-3), (7, 11, 20, 64, 1e-3)])
def test_parity_random(scanner, orc, prec, seed, lower, lmin, lmax, pval):
    say(f"parity seed={seed} prec={prec} L={lmin}..{lmax} lower/{lower} p={pval}")
    text, runs, bg = stream_case(seed, 120_000, lower_every=lower)
    ms = synth.motifs(40, lmin, lmax, 0.5, seed=seed)
    pwms, thr = library(ms, bg, pvalue=pval)
    n = check_against_oracle(scanner, orc, text, runs, pwms, thr, prec)
    assert n > 100


@pytest.mark.parametrize("prec", PRECS)
def test_ties_on_exact_grid(scanner, orc, prec):
    """Weights on a 1/8 grid are exact in bf16 hi+lo and in FP32, so window
    sums are exact; thresholds set to achieved sums must report the ties."""
    rng = np.random.default_rng(31)
    n = 50_000
    codes = rng.integers(0, 4, size=n)
    text = synth.CODE2BASE[codes].tobytes()
    pwms = [(rng.integers(-24, 25, size=(int(rng.integers(2, 33)), 4)) / 8.0).astype(np.float32) for _ in range(40)]
    thr = []
    for w in pwms:
        q = int(rng.integers(0, n - 40))
        thr.append(sum(float(w[p, codes[q + p]]) for p in range(w.shape[0])))
    thr = np.maximum(np.asarray(thr, np.float32), np.float32(0.5))
    say(f"ties prec={prec}")
    check_against_oracle(scanner, orc, text, [0], pwms, thr, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_thresholds_at_window_scores_with_large_lo_parts(scanner, orc, prec):
    """Weights whose bf16 lo parts are large (positive, near half an ulp of
    hi) and thresholds equal to the largest float not above a chosen window's
    FP64 score: the window must be reported.  Its hi-only score falls short of
    the threshold by about the sum of the lo parts, so a lazy-lo bound that
    left the lo terms out would drop it."""
    rng = np.random.default_rng(77)
    n = 40_000
    codes = rng.integers(0, 4, size=n)
    text = synth.CODE2BASE[codes].tobytes()
    pwms, thr = [], []
    for _ in range(48):
        L = int(rng.integers(1, 33))
        hi = rng.uniform(1.0, 2.0, size=(L, 4)).astype(np.float32)
        # bf16-representable hi, lo just under half a bf16 ulp of hi
        hi = (hi.view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)
        ulp = np.float32(2.0 ** -7) * np.float32(2.0) ** np.floor(np.log2(hi))
        w = (hi + np.float32(0.49) * ulp).astype(np.float32)
        pwms.append(w)
        q = int(rng.integers(0, n - 40))
        s64 = sum(float(w[p, codes[q + p]]) for p in range(L))
        t = np.float32(s64)
        if float(t) > s64:
            t = np.nextafter(t, np.float32(-np.inf))
        thr.append(t)
    say(f"lo-heavy ties prec={prec}")
    n_hits = check_against_oracle(scanner, orc, text, [0], pwms, np.asarray(thr, np.float32), prec)
    assert n_hits >= 48


@pytest.mark.parametrize("prec", PRECS)
def test_edge_cases(scanner, orc, prec):
    rng = np.random.default_rng(41)
    w1 = [np.array([[1.0, -1.0, 0.5, -0.5]], np.float32)]
    w32 = [rng.normal(0, 1, size=(32, 4)).astype(np.float32)]
    cases = [
        ("empty stream", b"", [], w1, [0.0]),
        ("one base", b"A", [0], w1, [0.5]),
        ("shorter than motif", b"ACGTACG", [0], w32, [-1e9]),
        ("every base its own run", b"ACGTACGTAC", list(range(10)), w1 + w32, [0.0, -1e9]),
        ("length 31/32/33 around L=32", b"ACGT" * 8 + b"A", [0], w32, [-1e9]),
        ("all lowercase", b"acgtacgtacgtacgtacgtacgt" * 3, [0], w1, [-0.1]),
        ("threshold -inf", synth.CODE2BASE[rng.integers(0, 4, size=3000)].tobytes(), [0, 1000], w1 + w32,
         [-np.inf, -np.inf]),
        ("threshold +inf", synth.CODE2BASE[rng.integers(0, 4, size=3000)].tobytes(), [0], w1 + w32, [np.inf, np.inf]),
    ]
    for name, text, runs, pwms, thr in cases:
        say(f"edge '{name}' prec={prec} n={len(text)}")
        check_against_oracle(scanner, orc, text, runs, pwms, np.asarray(thr, np.float32), prec)


@pytest.mark.parametrize("prec", PRECS)
def test_many_classes_and_partial_blocks(scanner, orc, prec):
    """Columns of every chunk count up to 64 positions, block sizes that are
    not multiples of 32."""
    text, runs, bg = stream_case(51, 60_000, lower_every=0)
    ms = synth.motifs(140, 1, 64, 0.5, seed=51)
    pwms, thr = library(ms, bg, pvalue=1e-3)
    say(f"classes prec={prec} cols={len(pwms)}")
    check_against_oracle(scanner, orc, text, runs, pwms, thr, prec)


@pytest.mark.parametrize("prec", PRECS)
def test_repeated_scans_are_identical(scanner, orc, prec):
    """The same scan repeated under every tile shape gives the oracle's hit
    set each time.  Guards against accumulator reads that race the matrix
    cores (seen once as about one wrong FP32 scan in ten)."""
    text, runs, bg = stream_case(5, 120_000, lower_every=0)
    ms = synth.motifs(40, 1, 8, 0.5, seed=5)
    pwms, thr = library(ms, bg, pvalue=1e-3)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, prec)
    want = hit_set(orc.scan_fp64(text, runs, pwms, thr))
    packed, lower = ba.pack(text)
    nbs = [1, 2] if prec != ba.PREC_FP32 else [1, 2, 4]
    bad = []
    try:
        for rep in range(2):
            for nb in nbs:
                for kp in (4, 8):
                    for ct in (1, 512):
                        scanner.set_tuning(ct, nb, kp)
                        got = hit_set(scanner.scan(packed, len(text), runs, lower))
                        if got != want:
                            bad.append((rep, nb, kp, ct, len(want - got), len(got - want)))
    finally:
        scanner.set_tuning(512, 2, 4)
    say(f"repeat prec={prec}: {len(bad)} wrong")
    assert not bad, bad


def test_pinned_scan_matches_host_scan(scanner, orc):
    text, runs, bg = stream_case(131, 90_000, lower_every=11)
    ms = synth.motifs(30, 4, 28, 0.5, seed=131)
    pwms, thr = library(ms, bg, pvalue=1e-3)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    packed, lower = ba.pack(text)
    a = scanner.scan(packed, len(text), runs, lower)
    for _ in range(2):
        b = scanner.scan_pinned(packed, len(text), runs, lower)
        assert np.array_equal(np.sort(a, order=["pos", "col"]), np.sort(b, order=["pos", "col"]))
    assert hit_set(a) == hit_set(orc.scan_fp64(text, runs, pwms, thr))


def test_hit_capacity_overflow_is_reported(scanner):
    text, runs, bg = stream_case(61, 30_000)
    ms = synth.motifs(8, 5, 12, 0.5, seed=61)
    pwms, thr = library(ms, bg, pvalue=1e-2)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    packed, lower = ba.pack(text)
    full = scanner.scan(packed, len(text), runs, lower)
    assert len(full) > 10
    tiny = scanner.scan(packed, len(text), runs, lower, cap=3)   # grows on OVERFLOW
    assert hit_set(tiny) == hit_set(full)
    # raw ABI: too small a buffer returns OVERFLOW with the required count
    import ctypes
    hits = np.zeros(4, dtype=ba.HIT_DTYPE)
    n = ctypes.c_uint64()
    pw = int(ba.hip_lib().blamm_hip_packed_words(len(text)))
    buf = np.zeros(pw, np.uint64)
    buf[:len(packed)] = packed
    runs64 = np.asarray(runs, np.uint64)
    rc = ba.hip_lib().blamm_hip_scan(scanner.ctx, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), None,
                                     len(text), runs64.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                     len(runs64), hits.ctypes.data, 4, ctypes.byref(n))
    assert rc == ba.BLAMM_ERR_OVERFLOW and n.value == len(full)


def test_invalid_arguments(scanner):
    w = [np.zeros((ba.MAX_MOTIF_LEN + 1, 4), np.float32)]
    P, lengths, _ = ba.build_pattern_matrix(w)
    with pytest.raises(ba.BlammError) as e:
        scanner.set_library(P, lengths, np.zeros(1, np.float32))
    assert e.value.code == ba.BLAMM_ERR_ARG and "length" in str(e.value)
    P, lengths, _ = ba.build_pattern_matrix([np.array([[np.inf, 0, 0, 0]], np.float32)])
    with pytest.raises(ba.BlammError) as e:
        scanner.set_library(P, lengths, np.zeros(1, np.float32))
    assert e.value.code == ba.BLAMM_ERR_ARG
    P, lengths, _ = ba.build_pattern_matrix([np.zeros((3, 4), np.float32)])
    with pytest.raises(ba.BlammError):
        scanner.set_library(P, lengths, np.array([np.nan], np.float32))
    with ba.HipScanner(0) as fresh:
        with pytest.raises(ba.BlammError) as e:
            fresh.scan(np.zeros(4, np.uint64), 10, [0])
        assert e.value.code == ba.BLAMM_ERR_STATE


def test_device_api_matches_host_api(scanner):
    import torch
    text, runs, bg = stream_case(71, 200_000, lower_every=11)
    ms = synth.motifs(30, 5, 30, 0.5, seed=71)
    pwms, thr = library(ms, bg, pvalue=1e-3)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    scanner.set_library(P, lengths, thr, ba.PREC_BF16X2)
    packed, lower = ba.pack(text)
    want = hit_set(scanner.scan(packed, len(text), runs, lower))
    n = len(text)
    dev = torch.device("cuda:0")
    pw = int(ba.hip_lib().blamm_hip_packed_words(n))
    mw = int(ba.hip_lib().blamm_hip_mask_words(n))
    hp = np.zeros(pw, np.uint64)
    hp[:len(packed)] = packed
    hm = np.zeros(mw, np.uint64)
    hm[:len(lower)] = lower
    d_codes = torch.from_numpy(hp.view(np.int64)).to(dev)
    d_low = torch.from_numpy(hm.view(np.int64)).to(dev)
    d_runs = torch.from_numpy(np.asarray(runs, np.uint64).view(np.int64)).to(dev)
    cap = len(want) + 16
    d_hits = torch.zeros(2 * cap, dtype=torch.int64, device=dev)
    d_n = torch.zeros(1, dtype=torch.int64, device=dev)
    torch.cuda.synchronize()
    for _ in range(3):   # the scratch may grow on the first call
        scanner.scan_device(d_codes, d_low, n, d_runs, len(runs), d_hits, cap, d_n)
        try:
            nh = scanner.scan_finish()
            break
        except ba.BlammError as e:
            assert e.code == ba.BLAMM_ERR_OVERFLOW
    assert int(d_n.item()) == nh
    got = d_hits[: 2 * nh].cpu().numpy().view(ba.HIT_DTYPE)
    assert hit_set(got) == want


def test_precisions_agree_mid_size(scanner, orc):
    """4 Mbp x 400 columns: split-bf16 and fp32 give the same hit set, and a
    slice of it equals the oracle's."""
    rng = np.random.default_rng(81)
    n = 4_000_000
    codes = synth.bases(n, 0.44, rng)
    nm = synth.n_runs(n, 0.005, 10, 500, rng)
    sc, runs = synth.filtered([synth.Record("m", codes, nm)])
    ms = synth.motifs(200, 5, 30, 0.4, seed=81)
    pwms, thr = library(ms, synth.counts_of(sc), pvalue=1e-4)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    packed = ba.pack_codes(sc)
    out = {}
    for prec in PRECS:
        say(f"mid-size prec={prec}")
        scanner.set_library(P, lengths, thr, prec)
        out[prec] = scanner.scan(packed, sc.size, runs, None, cap=1 << 22)
    assert hit_set(out[PRECS[0]]) == hit_set(out[PRECS[1]]) == hit_set(out[PRECS[2]])
    a, b = 1_500_000, 1_700_000
    text = synth.CODE2BASE[sc[a:b]].tobytes()
    sub_runs = [0] + [int(r) - a for r in runs if a < r < b]
    ref = orc.scan_fp64(text, sub_runs, pwms, thr)
    lmax = int(lengths.max())
    want = {(int(p) + a, int(c)) for p, c in zip(ref["pos"], ref["col"]) if p + lmax <= b - a}
    got = {(int(p), int(c)) for p, c in zip(out[PRECS[0]]["pos"], out[PRECS[0]]["col"]) if a <= p <= b - lmax}
    assert got == want and len(want) > 1000


def test_cli_example_matches_oracle(tmp_path):
    """`blamm dict/hist/scan -rc -pt 0.0001` on the reference example (its
    documented run, README.md:48-57) against the oracle's command line."""
    a, b = str(tmp_path / "prod"), str(tmp_path / "orc")
    copy_example(a)
    copy_example(b)
    for d, run in ((a, run_blamm), (b, run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "motifs.jaspar", "sequences.mf"], d)
    say("cli example scan")
    run_blamm(["scan", "-rc", "-pt", "0.0001", "motifs.jaspar", "sequences.mf"], a)
    run_oracle(["scan", "-rc", "-pt", "0.0001", "motifs.jaspar", "sequences.mf"], b)
    n = assert_same_occurrences(os.path.join(a, "occurrences.txt"), os.path.join(b, "occurrences.txt"))
    assert n > 50
    assert open(os.path.join(a, "PWMthresholds.txt")).read() == open(os.path.join(b, "PWMthresholds.txt")).read()
    golden = os.path.join(EXAMPLE, "..", "example_rc_pt1e-4.occurrences.txt")
    assert_same_occurrences(os.path.join(a, "occurrences.txt"), golden)


@pytest.mark.parametrize("opts", [["-at", "9.5"], ["-rt", "0.85"], ["-rc", "-rt", "0.9", "-p", "fp32"],
                                  ["-rc", "-pt", "0.001", "-o", "out.gtf"]])
def test_cli_variants_match_oracle(tmp_path, opts):
    from test_host import _write_synthetic
    a, b = tmp_path / "prod", tmp_path / "orc"
    for d in (a, b):
        d.mkdir()
        _write_synthetic(str(d), seed=9, lower=True, crlf=True)
        with open(d / "settings.cnf", "w") as f:
            f.write("PSEUDOCOUNT 0.5\nMATRIX_S_W 17\n")
    for d, run in ((str(a), run_blamm), (str(b), run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "motifs.jaspar", "sequences.mf"], d)
    say(f"cli variant {opts}")
    run_blamm(["scan"] + opts + ["motifs.jaspar", "sequences.mf"], str(a))
    oopts = [o for o in opts if o not in ("-p", "fp32")]
    run_oracle(["scan"] + oopts + ["motifs.jaspar", "sequences.mf"], str(b))
    out = "out.gtf" if "-o" in opts else "occurrences.txt"
    assert_same_occurrences(str(a / out), str(b / out))
    assert open(a / "PWMthresholds.txt").read() == open(b / "PWMthresholds.txt").read()


def test_cli_tuning_cache(tmp_path):
    """`blamm scan -T cache` tunes the tile shapes on a sample of the group,
    records them, reuses them on the next run, and the output still matches
    the oracle."""
    from test_host import _write_synthetic
    a, b = tmp_path / "prod", tmp_path / "orc"
    for d in (a, b):
        d.mkdir()
        _write_synthetic(str(d), seed=11, lower=False, crlf=False)
    for d, run in ((str(a), run_blamm), (str(b), run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "motifs.jaspar", "sequences.mf"], d)
    run_oracle(["scan", "-rc", "-pt", "0.001", "motifs.jaspar", "sequences.mf"], str(b))
    counts = []
    for k in range(2):
        say(f"cli tune run {k}")
        r = run_blamm(["scan", "-rc", "-pt", "0.001", "-T", "tune.txt", "motifs.jaspar", "sequences.mf"], str(a))
        assert "GPU tile shapes" in r.stdout
        assert_same_occurrences(str(a / "occurrences.txt"), str(b / "occurrences.txt"))
        counts.append(len(open(a / "tune.txt").read().splitlines()))
    # one entry per distinct key (device, library, sample size); the second
    # run finds them all in the cache
    assert 1 <= counts[0] <= 2 and counts[1] == counts[0], counts


def test_full_size_plain_gpu_config(scanner, orc):
    """The 100 Mbp x 2,000-column configuration: both precisions give the same
    hit set; three slices equal the oracle's; sampled hits re-score within the
    stated bound."""
    stream, runs = synth.config1_sequence(seed=1)
    ms = synth.config1_motifs(seed=1)
    bg = synth.counts_of(stream)
    pwms, thr = library(ms, bg, pvalue=1e-4)
    P, lengths, _ = ba.build_pattern_matrix(pwms)
    packed = ba.pack_codes(stream)
    out = {}
    for prec in PRECS:
        say(f"full size prec={prec}")
        scanner.set_library(P, lengths, thr, prec)
        out[prec] = scanner.scan(packed, stream.size, runs, None, cap=40_000_000)
    h = out[ba.PREC_BF16X2]
    key = lambda x: np.sort(x["pos"].astype(np.uint64) * np.uint64(4096) + x["col"].astype(np.uint64))
    assert np.array_equal(key(h), key(out[ba.PREC_FP32]))
    assert np.array_equal(key(h), key(out[ba.PREC_BF16X2_LO]))
    assert len(h) > 10_000_000
    lmax = int(lengths.max())
    rng = np.random.default_rng(5)
    for a in rng.integers(0, stream.size - 200_000, size=3):
        a = int(a)
        b = a + 150_000
        text = synth.CODE2BASE[stream[a:b]].tobytes()
        ref = orc.scan_fp64(text, [0] + [int(r) - a for r in runs if a < r < b], pwms, thr)
        want = {(int(p) + a, int(c)) for p, c in zip(ref["pos"], ref["col"]) if p + lmax <= b - a}
        sel = (h["pos"] >= a) & (h["pos"] <= b - lmax)
        got = set(zip(h["pos"][sel].tolist(), h["col"][sel].tolist()))
        assert got == want, (a, len(got ^ want))
    _, err = scanner.bounds()
    idx = rng.choice(len(h), size=5000, replace=False)
    for i in idx[:2000]:
        q, c = int(h["pos"][i]), int(h["col"][i])
        w = pwms[c]
        s64 = 0.0
        for p in range(w.shape[0]):
            s64 += float(w[p, stream[q + p]])
        assert s64 >= float(thr[c]) and abs(s64 - float(h["score"][i])) <= err


def test_cli_cluster_buster_permutations(tmp_path):
    """A cluster-buster motif file (any name not ending in .jaspar,
    motif.cpp:323-368) with permutation motifs (names ending in "__" + digits,
    motif.h:256-281): `hist` skips the permutations, `scan -pt` gives each one
    its base motif's histogram.  Two groups, lowercase, CRLF."""
    from test_host import _write_synthetic
    rng = np.random.default_rng(17)
    names = ["M0", "M0__1", "M0__22", "M1", "M1__3", "X_1", "Y", "Y__", "Z"]
    lines = []
    for nm in names:
        lines.append(">" + nm)
        for _ in range(int(rng.integers(4, 13))):
            lines.append(" ".join(str(int(v)) for v in rng.integers(0, 60, size=4)))
    a, b = tmp_path / "prod", tmp_path / "orc"
    for d in (a, b):
        d.mkdir()
        _write_synthetic(str(d), seed=12, lower=True, crlf=True)
        (d / "motifs.cb").write_text("\n".join(lines) + "\n")
    for d, run in ((str(a), run_blamm), (str(b), run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "motifs.cb", "sequences.mf"], d)
    hists = sorted(f for f in os.listdir(a) if f.startswith("hist_") and f.endswith(".dat"))
    assert hists == sorted(f for f in os.listdir(b) if f.startswith("hist_") and f.endswith(".dat"))
    assert not any("__" in f for f in hists)
    say("cli cluster-buster scan")
    for opts in (["-rc", "-pt", "0.001"], ["-rt", "0.8"]):
        run_blamm(["scan"] + opts + ["motifs.cb", "sequences.mf"], str(a))
        run_oracle(["scan"] + opts + ["motifs.cb", "sequences.mf"], str(b))
        n = assert_same_occurrences(str(a / "occurrences.txt"), str(b / "occurrences.txt"))
        assert n > 0
        assert open(a / "PWMthresholds.txt").read() == open(b / "PWMthresholds.txt").read()
        motifs = {k[1] for k in read_occurrences(str(a / "occurrences.txt"))}
        assert {"M0__1", "M0__22", "Y__"} & motifs


def test_cli_reference_cpu_config(tmp_path):
    """The reference's CPU-runnable configuration end to end: 64 motifs
    (L 6-20) and 4 records of 250 kbp with 1% N-runs, `dict`, theoretical
    `hist`, `scan -rc -pt 1e-4`, against the oracle's command line."""
    ms, recs = synth.config0(seed=0)
    a, b = tmp_path / "prod", tmp_path / "orc"
    for d in (a, b):
        d.mkdir()
        synth.write_fasta(str(d / "c0.fa"), recs)
        ms.write_jaspar(str(d / "motifs.jaspar"))
        (d / "sequences.mf").write_text("c0\tc0.fa\n")
    for d, run in ((str(a), run_blamm), (str(b), run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "motifs.jaspar", "sequences.mf"], d)
    say("cli reference CPU config scan")
    run_blamm(["scan", "-rc", "-pt", "1e-4", "motifs.jaspar", "sequences.mf"], str(a))
    run_oracle(["scan", "-rc", "-pt", "1e-4", "motifs.jaspar", "sequences.mf"], str(b))
    n = assert_same_occurrences(str(a / "occurrences.txt"), str(b / "occurrences.txt"))
    assert n > 1000
    # lines follow the length sort of the motifs, whose tie order is not
    # fixed (std::sort is unstable, motif.cpp:436): compare them as sets
    ta = sorted(open(a / "PWMthresholds.txt").read().splitlines())
    tb = sorted(open(b / "PWMthresholds.txt").read().splitlines())
    assert ta == tb and len(ta) > 50
