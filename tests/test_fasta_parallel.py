"""The parallel FASTA reader against the oracle's serial filter (CPU).

The reader is built into a small driver with 5-byte minimum chunks and
23-byte file segments (shorter than many lines), so chunk and segment
boundaries fall on headers, CRLF pairs, N-runs, lowercase runs and the maxLen
stop; every output (stream bytes, run table, record names, record count) must
equal the oracle's FASTA filter (sequence.cpp:142-293 restated); record
counts are compared for untruncated streams.
"""
import os
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fasta") / "fasta_harness")
    subprocess.run(["g++", "-O1", "-std=c++17", "-pthread", "-DBLAMM_FASTA_MIN_CHUNK=5", "-DBLAMM_FASTA_SEGMENT=23", "-o", out,
                    os.path.join(REPO, "tests", "fasta_harness.cpp"),
                    os.path.join(REPO, "blamm_amd", "csrc", "host", "fasta.cpp")], check=True)
    return out


def run_harness(harness, files, max_len):
    r = subprocess.run([harness, str(max_len)] + files, capture_output=True, text=True, check=True)
    lines = r.stdout.splitlines()
    if lines[0].startswith("error"):
        return ("error",)
    head = lines[0].split()
    n, nseq = int(head[1]), int(head[3])
    stream = lines[1][len("stream "):].encode() if len(lines) > 1 else b""
    runs = [tuple(int(x) for x in ln.split()[1:]) for ln in lines if ln.startswith("run ")]
    names = [ln[len("name "):] for ln in lines if ln.startswith("name ")]
    return n, nseq, stream, runs, names


def oracle_fasta(orc, files, max_len):
    try:
        stream, rs, rq, rp, nseq = orc.filter_fasta(files, max_len)
    except RuntimeError:
        return ("error",)
    runs = list(zip(rs.tolist(), rq.tolist(), rp.tolist()))
    return len(stream), nseq, stream, runs


def random_fasta(rng, path, first_header=True):
    parts = []
    nrec = int(rng.integers(0, 6))
    if not first_header:
        parts.append(b"ACGTNacgt\n")
    for r in range(nrec):
        head = rng.choice([b">", b"> ", b">\t"]) + f"rec{r}".encode() + rng.choice([b"", b" desc x", b"\t\tz"])
        parts.append(head + rng.choice([b"\n", b"\r\n"]))
        for _ in range(int(rng.integers(0, 5))):
            L = int(rng.integers(0, 70))
            line = bytearray(rng.choice(list(b"ACGTacgtNNnRY"), size=L).astype(np.uint8).tobytes())
            parts.append(bytes(line) + rng.choice([b"\n", b"\r\n", b"\n\n"]))
    data = b"".join(parts)
    if data and rng.random() < 0.3:
        data = data.rstrip(b"\n")      # no trailing newline
    with open(path, "wb") as f:
        f.write(data)


@pytest.mark.parametrize("seed", range(40))
def test_parallel_reader_matches_oracle(harness, orc, tmp_path, seed):
    rng = np.random.default_rng(seed)
    files = []
    for k in range(int(rng.integers(1, 4))):
        fn = str(tmp_path / f"f{k}.fa")
        random_fasta(rng, fn, first_header=(k > 0 or rng.random() < 0.9))
        files.append(fn)
    want = oracle_fasta(orc, files, (1 << 64) - 1)
    total = 0 if want[0] == "error" else want[0]
    for max_len in sorted({(1 << 64) - 1, 0, 1, total, max(total - 1, 0), int(rng.integers(0, total + 2))}):
        g = run_harness(harness, files, max_len)
        if max_len == 0:
            # an empty stream: nothing is read (the serial reader did not open the files either)
            assert g == (0, 0, b"", [], []), (seed, g)
            continue
        w = oracle_fasta(orc, files, max_len)
        if w[0] == "error" or g[0] == "error":
            assert w[0] == g[0], (seed, max_len, w[:1], g[:1])
            continue
        assert g[0] == w[0] and g[2] == w[2], (seed, max_len, "stream")
        assert g[3] == w[3], (seed, max_len, "runs")
        assert len(g[4]) == g[1], (seed, max_len, "names")
        if max_len > total:
            # past a truncation point the reader stops (the serial reader did
            # too); records there reach no output, so counts are compared
            # only for untruncated streams
            assert g[1] == w[1], (seed, max_len, "record count", g[1], w[1])


def test_reader_stops_when_full(harness, orc, tmp_path):
    """Once maxLen bases are packed the reader stops: a later file is not
    opened (a missing one is no error), as in the serial reader."""
    fn = str(tmp_path / "a.fa")
    with open(fn, "wb") as f:
        f.write(b">r1\n" + b"ACGT" * 40 + b"\n>r2\n" + b"GGCC" * 10 + b"\n")
    missing = str(tmp_path / "missing.fa")
    g = run_harness(harness, [fn, missing], 100)
    assert g[0] == 100 and g[2] == (b"ACGT" * 25)
    assert run_harness(harness, [fn, missing], 1000) == ("error",)
