"""The oracle's gate weave (JuicerCudaScanPipeline.cu:1514-1558, :2122-2146)
against what the Mitchell-Netravali filter fixes in closed form.

The weave is linear in the planes it resamples, so each check renders a
record with and without the weave and compares:

* no motion: the B = C = 1/3 cubic is not interpolating; at integer
  positions its taps are (1, 16, 1) / 18 in each direction, with replicated
  edges;
* a 90 degree rotation of a square window (cos 0, sin 1): output (x, y) reads
  source (W-1-y, x) before the same (1, 16, 1) / 18 filter (numpy's rot90,
  which fixes the rotation's sign);
* a half-pixel shift in x: taps mitchell(1.5), mitchell(0.5), mitchell(0.5),
  mitchell(1.5) = (-5, 77, 77, -5) / 144 along x, the (1, 16, 1) / 18 taps
  along y;
* a shift far outside the window: every output pixel reads the clamped
  border column.
"""
import numpy as np
import pytest

from fj import prepared, synth
from oracle import oracle


def taps_181(a, axis):
    lo = np.take(a, np.r_[0, np.arange(a.shape[axis] - 1)], axis=axis)
    hi = np.take(a, np.r_[np.arange(1, a.shape[axis]), a.shape[axis] - 1], axis=axis)
    return (lo + 16.0 * a + hi) / 18.0


def pair(w, h, dx, dy, deg, cos_sin=None):
    src = synth.frame(w, h, seed=w * 7 + h)
    plain = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    woven = prepared.with_gate_weave(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), dx, dy, deg)
    if cos_sin is not None:
        woven.record.effects.weave_cos_rot, woven.record.effects.weave_sin_rot = cos_sin
    _, lin0, _ = oracle.Plan(plain).render(src, threads=4, want_dst=False)
    _, lin1, _ = oracle.Plan(woven).render(src, threads=4, want_dst=False)
    return lin0, lin1


def test_no_motion_is_the_181_filter():
    lin0, lin1 = pair(37, 23, 0.0, 0.0, 0.0)
    np.testing.assert_allclose(lin1, taps_181(taps_181(lin0, 0), 1), rtol=0, atol=1e-12)


def test_quarter_turn_of_a_square_window():
    lin0, lin1 = pair(29, 29, 0.0, 0.0, 0.0, cos_sin=(0.0, 1.0))
    rot = np.rot90(lin0, k=1, axes=(0, 1))
    np.testing.assert_allclose(lin1, taps_181(taps_181(rot, 0), 1), rtol=0, atol=1e-12)


def test_half_pixel_shift():
    lin0, lin1 = pair(31, 19, 0.5, 0.0, 0.0)
    a = taps_181(lin0, 0)
    W = a.shape[1]
    idx = lambda k: np.clip(np.arange(W) + k, 0, W - 1)  # noqa: E731
    want = (-5 * a[:, idx(-1)] + 77 * a[:, idx(0)] + 77 * a[:, idx(1)] - 5 * a[:, idx(2)]) / 144.0
    # the position is clamped to W-1 before the taps: the last column reads
    # integer position W-1
    want[:, -1] = taps_181(a, 1)[:, -1]
    np.testing.assert_allclose(lin1, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dx", [1e4, -1e4])
def test_shift_beyond_the_window_reads_the_border(dx):
    lin0, lin1 = pair(17, 13, dx, 0.0, 0.0)
    a = taps_181(lin0, 0)
    col = taps_181(a, 1)[:, -1 if dx > 0 else 0, :]
    np.testing.assert_allclose(lin1, np.repeat(col[:, None, :], a.shape[1], axis=1), rtol=0, atol=1e-12)
