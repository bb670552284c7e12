"""The occurrence writer's score formatting (CPU).

`put_g6` (blamm_amd/csrc/host/format_g6.hpp) prints a float as ostream's
default output does, i.e. printf's %g with precision 6 of the value as a
double, which is what the reference's occurrence lines carry
(pwmscan.cpp:88-95).  It computes the six digits with exact integer
arithmetic for 1e-4 <= |v| < 1e6 and hands every other value to
std::to_chars.  The harness compares it with std::to_chars(general, 6) on:
random bit patterns, every m / 2^s with m < 2^16 (exact decimal ties for the
round-half-even rule), and every float within 2^16 bit patterns of each decade
boundary and of each 9.999995 x 10^k rounding edge.  Setting
BLAMM_FORMAT_EXHAUSTIVE=1 also checks every float of the integer path
(671,088,640 values, both signs; a few minutes on 8 cores).
"""
import os
import struct
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("fmt") / "format_harness")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", out, os.path.join(REPO, "tests", "format_harness.cpp")],
                   check=True)
    return out


def run(harness, *args):
    r = subprocess.run([harness] + [str(a) for a in args], capture_output=True, text=True, check=True)
    last = r.stdout.strip().splitlines()[-1].split()
    assert last[0] == "mismatches", r.stdout
    return int(last[1]), int(last[3]), r.stdout


def bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def test_random_bit_patterns(harness):
    bad, checked, out = run(harness, "random", 4_000_000, 12345)
    assert bad == 0, out
    assert checked > 3_900_000


def test_exact_ties(harness):
    bad, checked, out = run(harness, "ties")
    assert bad == 0, out
    assert checked == 2 * 40 * ((1 << 16) - 1)


def test_decade_and_rounding_edges(harness):
    edges = []
    for k in range(-6, 8):
        edges.append(bits(10.0 ** k))
        edges.append(bits(9.999995 * 10.0 ** (k - 1)))
    total = 0
    for e in edges:
        for sign in (0, 0x80000000):
            lo, hi = sign | max(0, e - (1 << 16)), sign | min(0x7F7FFFFF, e + (1 << 16))
            bad, checked, out = run(harness, "range", lo, hi)
            assert bad == 0, out
            total += checked
    assert total > 3_000_000


@pytest.mark.skipif(os.environ.get("BLAMM_FORMAT_EXHAUSTIVE") != "1", reason="exhaustive run only on request")
def test_every_float_of_the_integer_path(harness):
    lo, hi = 110 << 23, 150 << 23
    parts = [(lo + i * (hi - lo) // 8, lo + (i + 1) * (hi - lo) // 8) for i in range(8)]
    jobs = [(a | s, b | s) for a, b in parts for s in (0, 0x80000000)]
    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(lambda ab: run(harness, "range", ab[0], ab[1]), jobs))
    assert sum(r[0] for r in res) == 0
    assert sum(r[1] for r in res) == 2 * (hi - lo)
