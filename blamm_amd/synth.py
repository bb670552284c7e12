"""Seeded synthetic workloads for `blamm scan` benchmarks and tests.

Motifs: each position holds counts = round(100 * Dirichlet(alpha * 1_4)) over
ACGT, lengths uniform in [lmin, lmax], written in the JASPAR layout of the
reference's example/motifs.jaspar.  Sequences: i.i.d. bases at a GC content,
optionally with the CG dinucleotide depleted, plus runs of 'N'.
All generators use numpy.random.default_rng(seed).
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

CODE2BASE = np.frombuffer(b"ACGT", dtype=np.uint8)


@dataclass
class MotifSet:
    names: List[str]
    pfms: List[np.ndarray]        # L x 4 uint64 counts (ACGT)

    def write_jaspar(self, path: str) -> None:
        with open(path, "w") as f:
            for name, pfm in zip(self.names, self.pfms):
                f.write(f">{name}\tSYN\n")
                for b, row in zip("ACGT", pfm.T):
                    f.write(b + "  [" + "".join(f"{int(v):>7d}" for v in row) + " ]\n")


def motifs(n: int, lmin: int, lmax: int, alpha: float = 0.5, seed: int = 0, prefix: str = "SYN") -> MotifSet:
    rng = np.random.default_rng(seed)
    names, pfms = [], []
    for i in range(n):
        L = int(rng.integers(lmin, lmax + 1))
        p = rng.dirichlet(np.full(4, alpha), size=L)
        pfms.append(np.rint(100.0 * p).astype(np.uint64))
        names.append(f"{prefix}{i:05d}.1")
    return MotifSet(names, pfms)


def bases(n: int, gc: float, rng: np.random.Generator, cpg_factor: float = 1.0) -> np.ndarray:
    """n base codes (0..3 = ACGT); CG dinucleotides kept with prob cpg_factor."""
    at = (1.0 - gc) / 2.0
    s = rng.choice(4, size=n, p=[at, gc / 2, gc / 2, at]).astype(np.uint8)
    if cpg_factor < 1.0:
        # replace the G of a CG by A/C/T (composition-proportional) with
        # probability 1 - cpg_factor; repeat so that created CGs are thinned too
        for _ in range(4):
            idx = np.nonzero((s[:-1] == 1) & (s[1:] == 2))[0] + 1
            if idx.size == 0:
                break
            drop = idx[rng.random(idx.size) >= cpg_factor]
            s[drop] = rng.choice(np.array([0, 1, 3], np.uint8), size=drop.size, p=np.array([at, gc / 2, at]) / (1 - gc / 2))
    return s


def n_runs(n: int, frac: float, lo: int, hi: int, rng: np.random.Generator) -> np.ndarray:
    """Boolean mask of positions covered by N-runs of length U{lo..hi}."""
    mask = np.zeros(n, dtype=bool)
    if frac <= 0 or n == 0:
        return mask
    mean = (lo + hi) / 2.0
    k = max(1, int(round(frac * n / mean)))
    starts = rng.integers(0, n, size=k)
    lens = rng.integers(lo, hi + 1, size=k)
    for s, L in zip(starts, lens):
        mask[s:s + L] = True
    return mask


@dataclass
class Record:
    name: str
    codes: np.ndarray             # uint8 0..3
    nmask: np.ndarray             # True where the byte is 'N'

    def text(self) -> bytes:
        b = CODE2BASE[self.codes].copy()
        b[self.nmask] = ord("N")
        return b.tobytes()


def write_fasta(path: str, records: Sequence[Record], width: int = 60) -> None:
    with open(path, "wb") as f:
        for r in records:
            f.write(b">" + r.name.encode() + b"\n")
            t = r.text()
            for i in range(0, len(t), width):
                f.write(t[i:i + width] + b"\n")


def filtered(records: Sequence[Record]) -> Tuple[np.ndarray, np.ndarray]:
    """Filtered stream of records (N removed) and its run starts.

    A run starts at stream position 0, after every N-run and at every record
    (the reference's SeqBlock markers, sequence.cpp:81-88)."""
    parts, starts, off = [], [], 0
    for r in records:
        keep = ~r.nmask
        codes = r.codes[keep]
        if codes.size == 0:
            continue
        # a new run wherever the kept index jumps
        idx = np.nonzero(keep)[0]
        brk = np.nonzero(np.diff(idx) != 1)[0] + 1
        starts.append(off)
        starts.extend((off + brk).tolist())
        parts.append(codes)
        off += codes.size
    s = np.concatenate(parts) if parts else np.zeros(0, np.uint8)
    return s, np.asarray(starts, dtype=np.uint64)


def write_manifest(dirpath: str, groups: Sequence[Tuple[str, Sequence[str]]]) -> str:
    mf = os.path.join(dirpath, "sequences.mf")
    with open(mf, "w") as f:
        for g, files in groups:
            for fn in files:
                f.write(f"{g}\t{fn}\n")
    return mf


# ---- named workloads ---------------------------------------------------------

def config1_record(seed: int = 1, n: int = 100_000_000) -> Record:
    """One 100 Mbp record, GC 0.41, CG at 0.2x, 0.5% N-runs of 10..500."""
    rng = np.random.default_rng(seed)
    codes = bases(n, 0.41, rng, cpg_factor=0.2)
    nm = n_runs(n, 0.005, 10, 500, rng)
    return Record("chrSyn", codes, nm)


def config1_sequence(seed: int = 1, n: int = 100_000_000) -> Tuple[np.ndarray, np.ndarray]:
    """Filtered stream and run starts of config1_record."""
    return filtered([config1_record(seed, n)])


def config1_motifs(seed: int = 1) -> MotifSet:
    return motifs(1000, 5, 30, 0.5, seed=seed)


def config0(seed: int = 0, scale: float = 1.0) -> Tuple[MotifSet, List[Record]]:
    """64 motifs (L 6..20), 4 records of 250 kbp at GC 0.41 with 1% N-runs."""
    rng = np.random.default_rng(seed)
    m = motifs(64, 6, 20, 0.5, seed=seed)
    recs = []
    for i in range(4):
        n = int(250_000 * scale)
        recs.append(Record(f"syn{i}", bases(n, 0.41, rng), n_runs(n, 0.01, 10, 500, rng)))
    return m, recs


def counts_of(codes: np.ndarray) -> np.ndarray:
    return np.bincount(codes, minlength=4).astype(np.uint64)


def reverse_complement_pfm(pfm: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(pfm[::-1, ::-1])


def config2(seed: int = 2, n_each: int = 62_500_000, groups: int = 4):
    """Precision-form workload: 2,000 motifs (L 5..30, Dirichlet 0.3) and
    `groups` records of n_each bases, each its own manifest group with GC drawn
    uniformly from 0.35..0.50.  Returns (motifs, [(name, codes, gc)])."""
    rng = np.random.default_rng(seed)
    m = motifs(2000, 5, 30, 0.3, seed=seed)
    recs = []
    for i in range(groups):
        gc = float(rng.uniform(0.35, 0.50))
        recs.append((f"grp{i}", bases(n_each, gc, rng), gc))
    return m, recs


def config3(seed: int = 3, nseq: int = 20_000):
    """Batched uneven workload: nseq records with lengths lognormal (median
    2 kbp, sigma 1.0, clipped to 200 bp..200 kbp) at GC 0.45 and 1,500 motifs
    (L 6..25).  Returns (motifs, codes, run_starts, lengths): the records as
    one filtered stream whose runs are the records."""
    rng = np.random.default_rng(seed)
    m = motifs(1500, 6, 25, 0.5, seed=seed)
    lengths = np.clip(np.rint(rng.lognormal(np.log(2000.0), 1.0, size=nseq)), 200, 200_000).astype(np.int64)
    codes = bases(int(lengths.sum()), 0.45, rng)
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
    return m, codes, starts, lengths


def config4_library(seed: int = 4) -> MotifSet:
    """Serving library: 2,000 motifs drawn like the precision workload's."""
    return motifs(2000, 5, 30, 0.3, seed=seed)


def config4_request(rng: np.random.Generator, nmotifs: int, gc: float = 0.41, max_seqs: int = 5000):
    """One serving request: 1..max_seqs sequences with lengths lognormal
    (median 1 kbp, sigma 1.0, at most 1 Mbp), a motif subset of 10..2,000 drawn
    uniformly from the library, and a p-value from {1e-3, 1e-4, 1e-5, 1e-6}.
    Returns (codes, run_starts, motif_subset, pvalue)."""
    nseq = int(rng.integers(1, max_seqs + 1))
    lengths = np.clip(np.rint(rng.lognormal(np.log(1000.0), 1.0, size=nseq)), 1, 1_000_000).astype(np.int64)
    codes = bases(int(lengths.sum()), gc, rng)
    starts = np.concatenate([[0], np.cumsum(lengths)[:-1]]).astype(np.uint64)
    k = int(rng.integers(10, min(2000, nmotifs) + 1))
    subset = np.sort(rng.choice(nmotifs, size=k, replace=False))
    pvalue = float(rng.choice([1e-3, 1e-4, 1e-5, 1e-6]))
    return codes, starts, subset, pvalue
