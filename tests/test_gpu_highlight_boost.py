"""GPU parity of the camera-carrier highlight boost (fj_highlight.hip, after
src/Cuda/Film/JuicerCudaHighlightBoost.cu) against the FP64 oracle
(oracle/fj_oracle.c: fjo_highlight_boost): print and direct routes, with
scatter/halation after the boost, a bypassed boost (protection above the
carrier maximum) and the peak-overflow refusal. Gate: linear output max
|err| <= 2e-4, alpha bit-exact.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402

W, H = 128, 80


def render_gpu(p, src):
    h, w, c = src.shape
    s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream)
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy(), out, text


def check(p, src):
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
    assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    return got


def boosted(case, protection, sharpness=3.0, gain=2.0, halation=False):
    p = prepared.simple_view(prepared.load(case), encode=False)
    if halation:
        prepared.with_halation(p, W, H, 2.0, primary_amount=(0.3, 0.1, 0.015))
    return prepared.with_highlight_boost(p, W, H, protection, sharpness, gain)


def bright_frame(seed):
    src = synth.frame(W, H, seed=seed)
    src[20:30, 40:60, :3] *= 8.0  # a highlight well above the rest
    return src


def test_print_route_boost_changes_highlights():
    src = bright_frame(31)
    got = check(boosted("PrintPreflash", 0.05), src)
    plain, _, _ = render_gpu(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), src)
    assert np.abs(got[20:30, 40:60, :3] - plain[20:30, 40:60, :3]).max() > 1e-3


def test_direct_route_with_halation():
    check(boosted("NegativeDirectHanatos", 0.05, sharpness=6.0, gain=3.0, halation=True), bright_frame(32))


def test_bypass_when_protection_exceeds_the_maximum():
    check(boosted("PrintPreflash", 1e30), bright_frame(33))


def test_peak_overflow_refused():
    _, out, text = render_gpu(boosted("PrintPreflash", 0.0, gain=1e38), bright_frame(34))
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "HighlightBoostPeakOverflow" in text


def test_sharpness_at_the_recipe_bound():
    """28 = 28^(1 - range) at range 0, the recipe's largest sharpness
    (rust optics/highlight_boost.rs:84)."""
    check(boosted("PrintPreflash", 0.05, sharpness=28.0, gain=4.0), bright_frame(35))


def test_sharpness_above_the_bound_refused():
    _, out, text = render_gpu(boosted("PrintPreflash", 0.05, sharpness=28.5), bright_frame(36))
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "highlight_boost" in text
