"""The finalize scan as a difference of one exclusive prefix.

The reference's finalize (shaders/analyze_rate_control_finalize.comp:14-38)
runs a Hillis-Steele scan over 512 threads holding four savings totals each,
with steps 1 .. 128 only (quirk Q1); the kernels compute the same result as
the thread's in-thread prefix plus P[t] - P[max(0, t - 255)], P the exclusive
prefix of the thread totals (P[t] alone in the exact-prefix mode), modulo
2^32.  Checked here against the oracle's literal simulation of the loop
(itself pinned against the shader loop) on totals that are sparse, dense and
large enough to wrap.
"""
import numpy as np
import pytest

from oracle import oracle as O

M = (1 << 32) - 1


def window_scan(totals, exact):
    v = np.cumsum(totals.astype(np.int64).reshape(512, 4), axis=1) & M
    w = v[:, 3]
    p = np.concatenate([[0], np.cumsum(w)[:-1]]) & M
    t = np.arange(512)
    add = p if exact else (p - np.where(t >= 255, p[np.maximum(t - 255, 0)], 0)) & M
    return ((v + add[:, None]) & M).reshape(-1).astype(np.uint32)


@pytest.mark.parametrize("kind", ["sparse", "dense", "wrapping"])
@pytest.mark.parametrize("exact", [False, True])
def test_window_scan_equals_oracle_finalize(kind, exact):
    rng = np.random.default_rng({"sparse": 1, "dense": 2, "wrapping": 3}[kind])
    for _ in range(20):
        if kind == "wrapping":
            totals = rng.integers(0, 1 << 31, size=2048)
        else:
            totals = rng.integers(0, 6000, size=2048)
            if kind == "sparse":
                totals[rng.random(2048) < 0.8] = 0
        totals = totals.astype(np.uint32)
        assert np.array_equal(window_scan(totals, exact), O.finalize(totals, q1=not exact))
