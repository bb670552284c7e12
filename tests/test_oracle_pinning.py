"""The oracle against closed forms, brute force and hand-derived cases (CPU)."""
import pytest

import pinning

CHECKS = [name for name in dir(pinning) if name.startswith("check_")]


@pytest.mark.parametrize("name", CHECKS)
def test_pinning(orc, name):
    getattr(pinning, name)(orc)


def test_sgemm_path_matches_fp64_outside_the_band(orc):
    """The restated sgemm scan (the CPU baseline) reports the FP64 hit set
    except for windows whose FP64 score lies within FP32 rounding of the
    threshold."""
    import numpy as np
    rng = np.random.default_rng(21)
    n = 40_000
    codes = rng.integers(0, 4, size=n)
    text = np.frombuffer(b"ACGT", np.uint8)[codes].tobytes()
    runs = np.unique(np.concatenate([[0], rng.choice(n, 30, replace=False)])).astype(np.uint64)
    pwms = sorted([rng.normal(0, 1.5, size=(int(rng.integers(3, 25)), 4)).astype(np.float32) for _ in range(24)],
                  key=lambda w: w.shape[0])
    thr = np.array([w.max(1).sum() * 0.5 for w in pwms], np.float32)
    a = orc.scan_fp64(text, runs, pwms, thr)
    b = orc.scan_sgemm(text, runs, pwms, thr, h=37, w=11, min_zero_area=64, threads=4)
    sa = set(zip(a["pos"].tolist(), a["col"].tolist()))
    sb = set(zip(b["pos"].tolist(), b["col"].tolist()))
    assert len(sa) > 200
    band = {c: 4 * 2.0 ** -24 * 4 * w.shape[0] * np.abs(w).max(1).sum() for c, w in enumerate(pwms)}
    for (q, c) in sa ^ sb:
        s64 = sum(float(pwms[c][p, codes[q + p]]) for p in range(pwms[c].shape[0]))
        assert abs(s64 - float(thr[c])) <= band[c], (q, c, s64, thr[c])


def test_sgemm_independent_of_strip_shape(orc):
    """On a 1/8 weight grid every sum is exact, so any h, w and tiling give
    the same hits and scores."""
    import numpy as np
    rng = np.random.default_rng(8)
    n = 20_000
    codes = rng.integers(0, 4, size=n)
    text = np.frombuffer(b"ACGT", np.uint8)[codes].tobytes()
    runs = np.array([0, 5000, 5003, 17000], np.uint64)
    pwms = sorted([(rng.integers(-16, 17, size=(int(rng.integers(2, 20)), 4)) / 8.0).astype(np.float32)
                   for _ in range(12)], key=lambda w: w.shape[0])
    thr = np.array([w.max(1).sum() * 0.6 for w in pwms], np.float32)
    ref = orc.scan_fp64(text, runs, pwms, thr)
    want = sorted(zip(ref["pos"].tolist(), ref["col"].tolist(), ref["score32"].tolist()))
    for h, w, mza in ((1000, 250, 4096), (7, 3, 1), (50, 400, 10**9)):
        got = orc.scan_sgemm(text, runs, pwms, thr, h=h, w=w, min_zero_area=mza, threads=3)
        assert sorted(zip(got["pos"].tolist(), got["col"].tolist(), got["score32"].tolist())) == want
