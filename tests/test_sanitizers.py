"""Host code under sanitizers (CPU only).

The concurrent host pieces are built with ThreadSanitizer and with
AddressSanitizer + UndefinedBehaviorSanitizer and run on the inputs of their
functional tests:
- the parallel FASTA reader (fasta.cpp; 5-byte chunks and 23-byte segments so
  that many workers cut every line), whose output must still equal the
  oracle's serial filter;
- the occurrence writer (occurrences.cpp: worker-local formatting, pwrite at
  final offsets by one thread per part), whose file must equal the
  single-threaded one;
- the serving loop (serve.hpp: admission, the launcher / harvester /
  completer threads, the slot state machine, result hand-out, shutdown with a
  request still queued) over a CPU stand-in for the batch API
  (tests/server_harness.cpp), with 8 client threads and 1, 3 and 8 slots;
  every result must equal the stand-in's hits for that request alone;
- the `blamm` host code (settings, dict, motif loaders, theoretical
  histograms, .dat writer and reader, thresholds, error paths) under
  ASan + UBSan, with the commands that need no GPU.
Any sanitizer report fails the test (halt_on_error, nonzero exit code).
GPU code is not built with sanitizers (device sanitizers are not available).
"""
import os
import subprocess

import numpy as np
import pytest

import blamm_amd as ba
from helpers import EXAMPLE, copy_example, run_oracle

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(REPO, "blamm_amd", "csrc", "host")

SAN_ENV = {"ASAN_OPTIONS": "halt_on_error=1:abort_on_error=0:exitcode=66:detect_leaks=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1:exitcode=67",
           "TSAN_OPTIONS": "halt_on_error=1:exitcode=68:second_deadlock_stack=1"}
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
TSAN = ["-fsanitize=thread"]


def _build(out, sources, san, defines=()):
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-ffp-contract=off"] + list(defines) + san + \
          ["-o", out] + sources
    subprocess.run(cmd, check=True)
    return out


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    small = ["-DBLAMM_FASTA_MIN_CHUNK=5", "-DBLAMM_FASTA_SEGMENT=23"]
    fasta = [os.path.join(REPO, "tests", "fasta_harness.cpp"), os.path.join(HOST, "fasta.cpp")]
    writer = [os.path.join(REPO, "tests", "writer_harness.cpp"), os.path.join(HOST, "occurrences.cpp"),
              os.path.join(HOST, "fasta.cpp"), os.path.join(HOST, "inputs.cpp")]
    server = [os.path.join(REPO, "tests", "server_harness.cpp")]
    return {
        "fasta_asan": _build(str(d / "fasta_asan"), fasta, ASAN, small),
        "fasta_tsan": _build(str(d / "fasta_tsan"), fasta, TSAN, small),
        "writer_asan": _build(str(d / "writer_asan"), writer, ASAN),
        "writer_tsan": _build(str(d / "writer_tsan"), writer, TSAN),
        "server_asan": _build(str(d / "server_asan"), server, ASAN, ("-Wno-subobject-linkage",)),
        "server_tsan": _build(str(d / "server_tsan"), server, TSAN, ("-Wno-subobject-linkage",)),
    }


def _run(cmd, cwd=None):
    env = dict(os.environ)
    env.update(SAN_ENV)
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd, env=env)
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.returncode in (0, 1), (r.returncode, r.stderr[-4000:])
    return r


@pytest.mark.parametrize("kind", ["fasta_asan", "fasta_tsan"])
def test_fasta_reader_under_sanitizers(bins, orc, tmp_path, kind):
    from test_fasta_parallel import oracle_fasta, random_fasta
    for seed in range(12):
        rng = np.random.default_rng(1000 + seed)
        files = []
        for k in range(int(rng.integers(1, 4))):
            fn = str(tmp_path / f"s{seed}_{k}.fa")
            random_fasta(rng, fn, first_header=True)
            files.append(fn)
        want = oracle_fasta(orc, files, (1 << 64) - 1)
        for max_len in ((1 << 64) - 1, 37):
            r = _run([bins[kind], str(max_len)] + files)
            lines = r.stdout.splitlines()
            w = want if max_len > 1 << 62 else oracle_fasta(orc, files, max_len)
            if w[0] == "error":
                assert lines[0].startswith("error"), (seed, lines[:1])
                continue
            head = lines[0].split()
            stream = lines[1][len("stream "):].encode() if len(lines) > 1 else b""
            assert int(head[1]) == w[0] and stream == w[2], (seed, max_len)


def _random_writer_case(tmp_path, seed):
    rng = np.random.default_rng(seed)
    recs = []
    for r in range(int(rng.integers(1, 6))):
        n = int(rng.integers(50, 3000))
        body = bytes(rng.choice(list(b"ACGTacgtNN"), size=n).astype(np.uint8))
        recs.append(b">rec%d desc\n" % r + b"\n".join(body[i:i + 61] for i in range(0, n, 61)) + b"\n")
    fa = str(tmp_path / f"w{seed}.fa")
    with open(fa, "wb") as f:
        f.write(b"".join(recs))
    stream, *_ = __import__("oracle").default().filter_fasta([fa])
    n = len(stream)
    k = int(rng.integers(0, 40_000))
    hits = np.zeros(k, ba.HIT_DTYPE)
    if n:
        hits["pos"] = rng.integers(0, n, size=k)
    hits["col"] = rng.integers(0, 64, size=k)
    hits["score"] = rng.normal(0, 10, size=k).astype(np.float32)
    hb = str(tmp_path / f"w{seed}.bin")
    hits.tofile(hb)
    payload = int(rng.integers(n // 2, n + 1)) if n else 0
    return fa, hb, payload, int((hits["pos"] < payload).sum())


def _expected_lines(fa, hb, payload):
    """The occurrence lines of the writer harness's hits, built from the
    oracle's FASTA filter (run starts -> record and position) and the
    harness's column names, lengths and strands; scores as the reference's
    ostream prints a float (%g of the value as a double)."""
    _, rs, rq, rp, _ = __import__("oracle").default().filter_fasta([fa])
    hits = np.fromfile(hb, ba.HIT_DTYPE)
    out = []
    for pos, col, score in zip(hits["pos"].tolist(), hits["col"].tolist(), hits["score"].tolist()):
        if pos >= payload:
            continue
        r = int(np.searchsorted(rs, pos, side="right")) - 1
        p = int(rp[r]) + pos - int(rs[r])
        ln = 5 + col % 7
        out.append(f"rec{int(rq[r])}\tblamm\tM{col}\t{p}\t{p + ln}\t{'%g' % score}\t{'-' if col % 2 else '+'}\t.\t.\n")
    return "".join(out).encode()


@pytest.mark.parametrize("kind", ["writer_asan", "writer_tsan"])
def test_occurrence_writer_under_sanitizers(bins, tmp_path, kind):
    """8 writer threads produce the same file as 1, and that file is the
    lines built independently from the oracle's FASTA filter."""
    for seed in range(6):
        fa, hb, payload, expect = _random_writer_case(tmp_path, seed)
        outs = []
        for threads in (1, 8):
            out = str(tmp_path / f"o{seed}_{threads}.txt")
            r = _run([bins[kind], fa, hb, str(payload), str(threads), out])
            assert r.stdout.strip() == str(expect), r.stdout
            outs.append(open(out, "rb").read())
        assert outs[0] == outs[1]
        assert outs[0].count(b"\n") == expect
        assert outs[0] == _expected_lines(fa, hb, payload), seed


@pytest.mark.parametrize("kind", ["server_asan", "server_tsan"])
@pytest.mark.parametrize("slots", [1, 3, 8])
def test_server_threads_under_sanitizers(bins, kind, slots):
    """8 client threads x 40 requests through the server's three threads;
    each result equals the stand-in batch's hits for that request alone."""
    r = _run([bins[kind], "8", "40", str(slots)])
    assert r.returncode == 0 and r.stdout.startswith("ok 320 "), (r.stdout, r.stderr[-3000:])


@pytest.fixture(scope="module")
def blamm_asan(tmp_path_factory):
    """The blamm executable with its host code under ASan + UBSan (linked to
    the unsanitized libblamm_hip.so, which dict and theoretical hist do not
    call into)."""
    out = str(tmp_path_factory.mktemp("san_cli") / "blamm_asan")
    srcs = [os.path.join(HOST, f) for f in ("inputs.cpp", "fasta.cpp", "commands.cpp", "occurrences.cpp", "main.cpp")]
    lib = os.path.join(REPO, "blamm_amd", "lib")
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-pthread", "-ffp-contract=off"] + ASAN + ["-o", out] + srcs +
                   ["-L" + lib, "-lblamm_hip", "-Wl,-rpath," + lib, "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return out


def test_cli_host_code_under_asan(blamm_asan, tmp_path):
    """dict, theoretical hist (jaspar and cluster-buster), hist with options
    and the scan's input checks run clean under ASan + UBSan and write what
    the oracle writes."""
    import pinning
    a, b = tmp_path / "prod", tmp_path / "orc"
    for d in (a, b):
        d.mkdir()
        copy_example(str(d))
        (d / "motifs.cb").write_text(pinning.CLUSTER_BUSTER)
        (d / "settings.cnf").write_text("PSEUDOCOUNT 0.5\nMATRIX_S_W 17\n")
    env_leaks = {"ASAN_OPTIONS": SAN_ENV["ASAN_OPTIONS"].replace("detect_leaks=1", "detect_leaks=0")}
    for args in (["dict", "sequences.mf"], ["hist", "motifs.jaspar", "sequences.mf"],
                 ["hist", "-b", "97", "-l", "12345", "motifs.cb", "sequences.mf"]):
        env = dict(os.environ)
        env.update(SAN_ENV)
        env.update(env_leaks)
        r = subprocess.run([blamm_asan] + args, cwd=str(a), capture_output=True, text=True, env=env)
        assert r.returncode == 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
        run_oracle(args, str(b))
    files = sorted(p.name for p in a.iterdir() if p.name.endswith((".dat", ".dict")))
    assert len(files) > 10 and files == sorted(p.name for p in b.iterdir() if p.name.endswith((".dat", ".dict")))
    for f in files:
        assert (a / f).read_bytes() == (b / f).read_bytes(), f
    # scan input errors are reported before the device is opened
    for args, msg in ((["scan", "-pt", "1e-4", "missing.jaspar", "sequences.mf"], "open"),
                      (["scan", "-at", "1", "-rt", "0.5", "motifs.jaspar", "sequences.mf"], "threshold"),
                      (["scan", "-pt", "1e-4", "motifs.jaspar", "nothing.mf"], "open")):
        env = dict(os.environ)
        env.update(SAN_ENV)
        env.update(env_leaks)
        r = subprocess.run([blamm_asan] + args, cwd=str(a), capture_output=True, text=True, env=env)
        assert r.returncode != 0 and "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
        assert msg in (r.stdout + r.stderr).lower(), (args, r.stdout, r.stderr)
