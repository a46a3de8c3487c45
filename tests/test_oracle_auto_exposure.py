"""The oracle's camera auto-exposure meter (oracle/fj_oracle.c:
fjo_auto_exposure, after src/Cuda/JuicerCudaAutoExposure.cu).

The reference holds no metering golden vectors, so the meter is pinned by
laws a wrong implementation breaks: the raw exposure is linear in input
intensity, so a power-of-two input scale divides the metered scale exactly
(a meter that returned Y instead of 1/Y, or the wrong log base, fails); a
uniform frame meters the same under every weighted method; on a frame with a
bright centre, centre, spot and highlight weighting meter brighter than the
plain average; and the median of a two-level frame follows from the
2048-bin histogram formula.
"""
import numpy as np
import pytest

from fj import abi, prepared
from oracle import oracle

W, H = 96, 64
WEIGHTED = [abi.FJ_METER_CENTER_WEIGHTED, abi.FJ_METER_AVERAGE, abi.FJ_METER_PARTIAL, abi.FJ_METER_MATRIX,
            abi.FJ_METER_MULTI_ZONE, abi.FJ_METER_HIGHLIGHT_WEIGHTED]


def plan(method, case="PrintPreflash", meter=None):
    p = prepared.simple_view(prepared.load(case), encode=False)
    prepared.with_auto_exposure(p, W, H, method, meter)
    return p, oracle.Plan(p)


def uniform(rgb):
    f = np.empty((H, W, 4), np.float32)
    f[..., :3] = rgb
    f[..., 3] = 1.0
    return f


def test_off_is_one():
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    assert oracle.Plan(p).auto_exposure(uniform((0.2, 0.3, 0.4))) == 1.0


@pytest.mark.parametrize("method", WEIGHTED + [abi.FJ_METER_MEDIAN])
def test_power_of_two_input_scale(method):
    _, pl = plan(method)
    rng = np.random.default_rng(method)
    f = uniform((0.0, 0.0, 0.0))
    f[..., :3] = rng.uniform(0.02, 0.9, size=(H, W, 3))
    s1 = pl.auto_exposure(f)
    g = f.copy()
    g[..., :3] *= 4.0
    s4 = pl.auto_exposure(g)
    assert s1 != 1.0
    assert s4 == pytest.approx(s1 / 4.0, rel=3e-7)


def test_uniform_frame_same_for_weighted_methods():
    f = uniform((0.18, 0.22, 0.15))
    scales = [plan(m)[1].auto_exposure(f) for m in WEIGHTED]
    assert max(scales) == pytest.approx(min(scales), rel=1e-6)
    assert 0.1 < scales[0] < 1e3


def test_bright_centre_orders_the_methods():
    f = uniform((0.05, 0.05, 0.05))
    f[H // 2 - 6:H // 2 + 6, W // 2 - 8:W // 2 + 8, :3] = 0.8
    s = {m: plan(m)[1].auto_exposure(f) for m in WEIGHTED}
    avg = s[abi.FJ_METER_AVERAGE]
    # brighter metering -> smaller scale
    assert s[abi.FJ_METER_CENTER_WEIGHTED] < avg
    assert s[abi.FJ_METER_PARTIAL] < s[abi.FJ_METER_CENTER_WEIGHTED]
    assert s[abi.FJ_METER_HIGHLIGHT_WEIGHTED] < avg
    assert s[abi.FJ_METER_MULTI_ZONE] < avg


def test_median_two_level_frame():
    dark, bright = (0.06, 0.07, 0.05), (0.5, 0.55, 0.45)
    m_dark = 1.0 / plan(abi.FJ_METER_AVERAGE)[1].auto_exposure(uniform(dark))
    m_bright = 1.0 / plan(abi.FJ_METER_AVERAGE)[1].auto_exposure(uniform(bright))
    f = uniform(dark)
    f[:, int(W * 0.6):, :3] = bright  # 40 % bright columns
    _, pl = plan(abi.FJ_METER_MEDIAN)
    n = W * H
    n_dark = H * int(W * 0.6)
    bin_dark = int(np.float32(np.float32(m_dark) / np.float32(m_bright)) * np.float32(2047))
    frac = (n // 2) / n_dark
    expected = 1.0 / ((bin_dark + frac) * m_bright / 2048)
    assert pl.auto_exposure(f) == pytest.approx(expected, rel=2e-4)


def test_meter_window_and_source_mismatch():
    f = uniform((0.05, 0.05, 0.05))
    f[:, W // 2:, :3] = 0.6
    _, left = plan(abi.FJ_METER_AVERAGE, meter=(0, 0, W // 2, H))
    _, right = plan(abi.FJ_METER_AVERAGE, meter=(W // 2, 0, W, H))
    assert left.auto_exposure(f) > right.auto_exposure(f)
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    p.record.auto_exposure = prepared.auto_exposure_descriptor((0, 0, W + 1, H), (0, 0, W, H), 1)
    p.record.film_exposure.flags |= abi.FJ_FILM_AUTO_EXPOSURE
    assert oracle.Plan(p).auto_exposure(f) == -1.0


def test_preview_descriptor_dimensions():
    d = prepared.auto_exposure_descriptor((0, 0, 3840, 2160), (0, 0, 3840, 2160), 0)
    assert (d.preview_width, d.preview_height) == (256, 144)
    d = prepared.auto_exposure_descriptor((0, 0, 100, 700), (10, 20, 110, 720), 2)
    assert (d.preview_width, d.preview_height) == (37, 256)
    d = prepared.auto_exposure_descriptor((0, 0, 64, 48), (0, 0, 64, 48), 1)
    assert (d.preview_width, d.preview_height) == (64, 48) and d.hash != 0
