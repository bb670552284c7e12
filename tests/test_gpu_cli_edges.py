"""Command-line edge cases on the GPU against the oracle's command line.

- Segments: `blamm scan` and `blamm hist -e` cut each group's stream into
  segments (BLAMM_SEGMENT_BASES, a multiple of 64) with an Lmax - 1 halo;
  hits that start in a halo belong to the next segment and histogram windows
  are counted where they start.  Tiny segments make windows, run starts and
  record boundaries straddle every segment and halo edge.
- NaN thresholds: getScoreCutoff returns 0/0 when the top bin that reaches
  p * total is empty (`-pt 0` on most theoretical histograms, any all-zero
  histogram); the reference then reports no occurrence for that motif
  (R >= NaN is false, kernel.cu:29, pwmscan.cpp:117-119) and prints the NaN
  in PWMthresholds.txt.
"""
import filecmp
import os
import sys

import numpy as np
import pytest

import blamm_amd as ba
from blamm_amd import synth
from helpers import assert_same_occurrences, read_occurrences, run_blamm, run_oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def thresholds(d):
    """PWMthresholds.txt lines as a sorted list: the lines follow the length
    sort of the motifs, whose tie order is not fixed (std::sort is unstable,
    motif.cpp:436)."""
    return sorted(open(os.path.join(d, "PWMthresholds.txt")).read().splitlines())


def write_inputs(d, seed=5):
    """Two groups.  Group g0: records whose valid-base counts put record starts
    on multiples of 64 and 4096 of the stream (4096, then 64 + 4096 ...), an
    N-run ending on a segment boundary, lowercase bases; group g1: one record
    with dense N-runs.  25 jaspar motifs, L 5-30."""
    rng = np.random.default_rng(seed)
    recs = []
    for i, n in enumerate([4096, 4160, 2048 + 17, 8192, 333, 6000]):
        codes = synth.bases(n, 0.42, rng)
        nm = np.zeros(n, bool)
        if i == 3:
            nm[4096 - 100:4096] = True       # an N-run that ends 100 bases later in the stream
        recs.append(synth.Record(f"r{i}", codes, nm))
    synth.write_fasta(os.path.join(d, "a.fa"), recs, width=61)
    data = bytearray(open(os.path.join(d, "a.fa"), "rb").read())
    for k in range(0, len(data), 11):
        if chr(data[k]) in "ACGT":
            data[k] |= 0x20
    open(os.path.join(d, "a.fa"), "wb").write(bytes(data))
    n = 30_000
    synth.write_fasta(os.path.join(d, "b.fa"), [synth.Record("b0", synth.bases(n, 0.38, rng),
                                                             synth.n_runs(n, 0.05, 1, 40, rng))])
    with open(os.path.join(d, "sequences.mf"), "w") as f:
        f.write("g0\ta.fa\ng1\tb.fa\n")
    synth.motifs(25, 5, 30, 0.5, seed=seed).write_jaspar(os.path.join(d, "motifs.jaspar"))


@pytest.fixture
def dirs(tmp_path):
    a, b = tmp_path / "prod", tmp_path / "orc"
    for d in (a, b):
        d.mkdir()
        write_inputs(str(d))
    for d, run in ((str(a), run_blamm), (str(b), run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "motifs.jaspar", "sequences.mf"], d)
    return str(a), str(b)


@pytest.mark.parametrize("seg", [64, 4096, 10_000])
def test_cli_scan_segments_match_oracle(dirs, seg):
    a, b = dirs
    run_oracle(["scan", "-rc", "-pt", "0.01", "motifs.jaspar", "sequences.mf"], b)
    say(f"cli scan with {seg}-base segments")
    r = run_blamm(["scan", "-rc", "-pt", "0.01", "motifs.jaspar", "sequences.mf"], a,
                  env={"BLAMM_SEGMENT_BASES": str(seg)})
    n = assert_same_occurrences(os.path.join(a, "occurrences.txt"), os.path.join(b, "occurrences.txt"))
    assert n > 5000, r.stdout
    # occurrences straddle segment edges (stream offsets of the first three
    # records of group g0, which have no N)
    occ = read_occurrences(os.path.join(a, "occurrences.txt"))
    step = seg // 64 * 64
    start = {"r0": 0, "r1": 4096, "r2": 8256}
    crossing = [k for k in occ if k[0] in start and
                (start[k[0]] + k[2]) // step != (start[k[0]] + k[3] - 1) // step]
    assert crossing, "no occurrence straddles a segment edge"


@pytest.mark.parametrize("seg", [64, 4096])
def test_cli_hist_empirical_segments_match_oracle(dirs, seg):
    a, b = dirs
    for d in (a, b):
        os.makedirs(os.path.join(d, "h"), exist_ok=True)
    run_oracle(["hist", "-e", "-b", "60", "-H", "h", "motifs.jaspar", "sequences.mf"], b)
    say(f"cli hist -e with {seg}-base segments")
    run_blamm(["hist", "-e", "-b", "60", "-H", "h", "motifs.jaspar", "sequences.mf"], a,
              env={"BLAMM_SEGMENT_BASES": str(seg)})
    fa = sorted(f for f in os.listdir(os.path.join(a, "h")) if f.endswith(".dat"))
    fb = sorted(f for f in os.listdir(os.path.join(b, "h")) if f.endswith(".dat"))
    assert fa == fb and len(fa) == 50
    for f in fa:
        assert filecmp.cmp(os.path.join(a, "h", f), os.path.join(b, "h", f), shallow=False), f


def test_cli_pt_zero_matches_oracle(dirs):
    """-pt 0: motifs whose top histogram bin is empty get a NaN threshold and
    no occurrence; the others report only windows at their top bin."""
    a, b = dirs
    run_oracle(["scan", "-rc", "-pt", "0", "motifs.jaspar", "sequences.mf"], b)
    say("cli scan -pt 0")
    run_blamm(["scan", "-rc", "-pt", "0", "motifs.jaspar", "sequences.mf"], a)
    assert_same_occurrences(os.path.join(a, "occurrences.txt"), os.path.join(b, "occurrences.txt"))
    ta = thresholds(a)
    assert ta == thresholds(b)
    assert any("nan" in ln for ln in ta)


def test_cli_all_zero_histogram_matches_oracle(dirs):
    """A histogram file with every count 0 gives that motif a NaN threshold
    at any p-value; the rest of the scan is unaffected."""
    a, b = dirs
    victim = sorted(f for f in os.listdir(a) if f.startswith("hist_g0_") and f.endswith(".dat"))[3]
    for d in (a, b):
        lines = open(os.path.join(d, victim)).read().splitlines()
        out = [lines[0]] + [ln.split("\t")[0] + "\t0" for ln in lines[1:]]
        open(os.path.join(d, victim), "w").write("\n".join(out) + "\n")
    run_oracle(["scan", "-rc", "-pt", "0.01", "motifs.jaspar", "sequences.mf"], b)
    say("cli scan with an all-zero histogram")
    run_blamm(["scan", "-rc", "-pt", "0.01", "motifs.jaspar", "sequences.mf"], a)
    n = assert_same_occurrences(os.path.join(a, "occurrences.txt"), os.path.join(b, "occurrences.txt"))
    assert n > 1000
    assert thresholds(a) == thresholds(b)
    name = victim[len("hist_g0_"):-len(".dat")]
    occ = read_occurrences(os.path.join(a, "occurrences.txt"))
    # no occurrence of that motif in group g0 (records r0..r5); group g1 has
    # its own histogram
    assert not any(k[1] == name and k[0].startswith("r") for k in occ)
    assert any(k[1] == name and k[0] == "b0" for k in occ)


def test_library_and_scratch_changes_rejected_while_scan_pending():
    """set_library, reserve and set_tuning return BLAMM_ERR_STATE between
    scan_device and scan_finish (the kernels still read the library and the
    scratch); after scan_finish they work again."""
    import torch
    rng = np.random.default_rng(3)
    n = 200_000
    text = synth.CODE2BASE[rng.integers(0, 4, n)].tobytes()
    pwms = [rng.normal(0, 1, size=(int(rng.integers(5, 20)), 4)).astype(np.float32) for _ in range(40)]
    thr = np.array([w.max(1).sum() * 0.6 for w in pwms], np.float32)
    P, L, _ = ba.build_pattern_matrix(pwms)
    with ba.HipScanner(0) as sc:
        sc.set_library(P, L, thr, ba.PREC_BF16X2)
        packed, _ = ba.pack(text)
        pw = int(sc.lib.blamm_hip_packed_words(n))
        buf = np.zeros(pw, np.uint64)
        buf[:packed.size] = packed
        d_codes = torch.from_numpy(buf.view(np.int64)).cuda()
        d_runs = torch.zeros(1, dtype=torch.int64, device="cuda:0")
        d_hits = torch.zeros(2 << 20, dtype=torch.int64, device="cuda:0")
        torch.cuda.synchronize()
        sc.scan_device(d_codes, None, n, d_runs, 0, d_hits, 1 << 20, None)
        for call in (lambda: sc.set_library(P, L, thr, ba.PREC_FP32), lambda: sc.reserve(1 << 24),
                     lambda: sc.set_tuning(512, 1, 4)):
            with pytest.raises(ba.BlammError) as e:
                call()
            assert e.value.code == ba.BLAMM_ERR_STATE
        nh = sc.scan_finish()
        assert nh > 0
        sc.reserve(1 << 21)
        sc.set_library(P, L, thr, ba.PREC_FP32)
        sc.scan_device(d_codes, None, n, d_runs, 0, d_hits, 1 << 20, None)
        assert sc.scan_finish() == nh


def test_cli_long_motifs_match_oracle(tmp_path):
    """Motifs of 33-64 positions (the reference has no length cap,
    motif.cpp:549-551): theoretical hist, scan -rc -pt and hist -e against
    the oracle's command line; a 65-position motif is refused with a
    message naming the limit."""
    a, b = tmp_path / "prod", tmp_path / "orc"
    ms = synth.motifs(14, 30, 64, 0.5, seed=8)
    for d in (a, b):
        d.mkdir()
        write_inputs(str(d), seed=8)
        ms.write_jaspar(str(d / "long.jaspar"))
        os.makedirs(d / "h", exist_ok=True)
    for d, run in ((str(a), run_blamm), (str(b), run_oracle)):
        run(["dict", "sequences.mf"], d)
        run(["hist", "long.jaspar", "sequences.mf"], d)
        run(["hist", "-e", "-b", "50", "-H", "h", "long.jaspar", "sequences.mf"], d)
    say("cli scan with 30-64 position motifs")
    run_blamm(["scan", "-rc", "-pt", "0.02", "long.jaspar", "sequences.mf"], str(a))
    run_oracle(["scan", "-rc", "-pt", "0.02", "long.jaspar", "sequences.mf"], str(b))
    n = assert_same_occurrences(str(a / "occurrences.txt"), str(b / "occurrences.txt"))
    assert n > 100
    occ = read_occurrences(str(a / "occurrences.txt"))
    assert max(k[3] - k[2] for k in occ) > 40
    dats = sorted(f for f in os.listdir(a / "h") if f.endswith(".dat"))
    assert dats == sorted(f for f in os.listdir(b / "h") if f.endswith(".dat")) and len(dats) == 28
    for f in dats:
        assert filecmp.cmp(str(a / "h" / f), str(b / "h" / f), shallow=False), f
    too_long = synth.MotifSet(["TOOLONG.1"], [np.full((65, 4), 25, np.uint64)])
    too_long.write_jaspar(str(a / "toolong.jaspar"))
    r = run_blamm(["scan", "-at", "0", "toolong.jaspar", "sequences.mf"], str(a), check=False)
    assert r.returncode != 0 and "64" in (r.stdout + r.stderr)
