"""Shared helpers for the tests (workload builders, output normalisation)."""
from __future__ import annotations

import math
import os
import shutil
import subprocess

import numpy as np

import blamm_amd as ba
from blamm_amd import synth

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXAMPLE = os.path.join(REPO, "tests", "golden", "example")


def run_blamm(args, cwd, check=True, env=None):
    e = dict(os.environ)
    if env:
        e.update(env)
    r = subprocess.run([ba.BLAMM_BIN] + list(args), cwd=cwd, capture_output=True, text=True, env=e)
    if check and r.returncode != 0:
        raise AssertionError(f"blamm {' '.join(args)} failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
    return r


def run_oracle(args, cwd, check=True):
    import oracle
    return oracle.run_cli(args, cwd, check=check)


def copy_example(dst: str) -> str:
    shutil.copytree(EXAMPLE, dst, dirs_exist_ok=True)
    return dst


def read_occurrences(path: str):
    """Occurrence lines -> {(seq, motif, start, end, strand): score}."""
    out = {}
    with open(path) as f:
        for line in f:
            t = line.rstrip("\n").split("\t")
            assert len(t) == 9 and t[1] == "blamm" and t[7] == "." and t[8] == ".", line
            key = (t[0], t[2], int(t[3]), int(t[4]), t[6])
            assert key not in out, f"duplicate occurrence {key}"
            out[key] = float(t[5])
    return out


def print_unit(v: float) -> float:
    """One unit in the 6th significant digit (ostream's default float print)."""
    return 10.0 ** (math.floor(math.log10(abs(v))) - 5) if v != 0 else 1e-30


def assert_same_occurrences(a: str, b: str, score_tol: float = 2.5e-4):
    """Same occurrence set; printed scores within score_tol (the engine's
    stated bound for these libraries) plus one print unit."""
    x, y = read_occurrences(a), read_occurrences(b)
    missing = sorted(set(y) - set(x))[:5]
    extra = sorted(set(x) - set(y))[:5]
    assert set(x) == set(y), f"{len(set(y) - set(x))} missing {missing}, {len(set(x) - set(y))} extra {extra}"
    worst = max((abs(x[k] - y[k]) - score_tol - print_unit(max(abs(x[k]), abs(y[k]))) for k in x), default=-1.0)
    assert worst <= 0, f"score differs beyond tolerance by {worst}"
    return len(x)


def library(motifset, bg, pvalue=None, rc=True, pseudocount=0.25, thr_scale=None):
    """PWM columns (fwd, rc interleaved) and thresholds as `blamm scan` builds them."""
    pwms, thr = [], []
    for pfm in motifset.pfms:
        fwd = ba.pwm(pfm, bg, pseudocount, False)
        cols = [fwd]
        if rc:
            cols.append(ba.pwm(synth.reverse_complement_pfm(pfm), bg, pseudocount, True))
        if pvalue is not None:
            t = ba.pvalue_threshold(fwd, bg, pvalue, pseudocount)
            ts = [t] * len(cols)
        else:
            ts = [np.float32(thr_scale * float(c.max(1).sum())) for c in cols]
        pwms += cols
        thr += ts
    return pwms, np.asarray(thr, np.float32)


def stream_case(seed, n, lower_every=0, nfrac=0.01, gc=0.41):
    rng = np.random.default_rng(seed)
    codes = synth.bases(n, gc, rng)
    nm = synth.n_runs(n, nfrac, 10, 200, rng)
    sc, runs = synth.filtered([synth.Record("s", codes, nm)])
    text = bytearray(synth.CODE2BASE[sc].tobytes())
    if lower_every:
        for i in range(0, len(text), lower_every):
            text[i] |= 0x20
    return bytes(text), runs, synth.counts_of(sc)


def hit_set(h):
    return set(zip(np.asarray(h["pos"]).tolist(), np.asarray(h["col"]).tolist()))
