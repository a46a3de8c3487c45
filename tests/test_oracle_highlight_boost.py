"""The oracle's camera-carrier highlight boost (oracle/fj_oracle.c:
fjo_highlight_boost, after src/Cuda/Film/JuicerCudaHighlightBoost.cu).

Pinned by what the curve fixes: the carrier maximum m maps to m * gain,
samples at or below the protection p are unchanged, the curve is monotone
and leaves p with slope 1, a zero sharpness gives the plain quadratic
s + m (g - 1) ((s - p) / (m - p))^2, a carrier whose maximum is at most p is
bypassed, and invalid or overflowing carriers are reported.
"""
import numpy as np
import pytest

from fj import abi
from oracle import oracle


def desc(p, k, g):
    d = abi.FjHighlightBoost()
    d.recipe_hash, d.protection, d.sharpness, d.peak_gain = 1, p, k, g
    return d


def planes(values):
    v = np.asarray(values, np.float64)
    return [v, v * 0.5, v * 0.25]


@pytest.mark.parametrize("k", [0.0, 0.3, 2.0, 9.0, 28.0])  # 28: the recipe's largest sharpness
def test_endpoints_monotone_and_slope(k):
    p, g = 0.8, 3.0
    s = np.linspace(0.0, 4.0, 4001)
    rc, out = oracle.highlight_boost(desc(p, k, g), planes(s))
    assert rc == 0
    b = out[0]
    assert b[-1] == pytest.approx(4.0 * g, rel=1e-12)
    np.testing.assert_array_equal(b[s <= p], s[s <= p])
    assert np.all(np.diff(b) > 0)
    h = 1e-6
    _, near = oracle.highlight_boost(desc(p, k, g), [np.array([p + h, 4.0]), np.zeros(2), np.zeros(2)])
    assert (near[0][0] - p) / h == pytest.approx(1.0, abs=1e-4)


def test_zero_sharpness_is_quadratic():
    p, g, m = 0.5, 2.5, 3.0
    s = np.linspace(0.0, m, 301)
    _, out = oracle.highlight_boost(desc(p, 0.0, g), planes(s))
    f = np.clip((s - p) / (m - p), 0, None)
    want = np.where(s > p, s + m * (g - 1) * f * f, s)
    np.testing.assert_allclose(out[0], want, rtol=1e-13, atol=1e-15)


def test_bypass_invalid_overflow():
    s = np.linspace(0.0, 0.7, 50)
    rc, out = oracle.highlight_boost(desc(0.8, 2.0, 3.0), planes(s))
    assert rc == 0
    np.testing.assert_array_equal(out[0], s)
    bad = s.copy()
    bad[3] = np.nan
    assert oracle.highlight_boost(desc(0.1, 2.0, 3.0), planes(bad))[0] == 6
    neg = s.copy()
    neg[5] = -1e-3
    assert oracle.highlight_boost(desc(0.1, 2.0, 3.0), planes(neg))[0] == 6
    assert oracle.highlight_boost(desc(0.1, 2.0, 1e39), planes(s))[0] == 7
