"""GPU parity of the scanner lens blur (blur_plane_in_place on the three
scanner-linear planes before the unsharp mask, src/Cuda/Scan/
JuicerCudaScanPipeline.cu:2686-2809) against the FP64 oracle, alone and
with the unsharp mask, on print and direct routes. Gate: linear output max
|err| <= 2e-4, alpha bit-exact.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402


def render_gpu(p, src):
    h, w, c = src.shape
    s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream)
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy(), out, text


@pytest.mark.parametrize("case,unsharp,sigma,shape", [("PrintPreflash", False, 1.5, (120, 72)),
                                                      ("PrintPreflash", True, 3.0, (97, 61)),
                                                      ("NegativeDirectHanatos", True, 0.8, (64, 40)),
                                                      ("NegativeDirectHanatos", False, 12.0, (40, 30))])
def test_lens_blur(case, unsharp, sigma, shape):
    w, h = shape
    p = prepared.simple_view(prepared.load(case), encode=False, unsharp=unsharp)
    prepared.with_lens_blur(p, sigma)
    src = synth.frame(w, h, seed=w + h)
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
    assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    plain = prepared.simple_view(prepared.load(case), encode=False, unsharp=unsharp)
    ref, _, _ = render_gpu(plain, src)
    assert np.abs(got - ref).max() > 1e-4  # the blur took effect


@pytest.mark.parametrize("sigma,radius", [(25.0, 75), (9.0, 27)])
def test_lens_blur_large_radius(sigma, radius):
    """The largest admitted radius (75) runs the vertical pass with a 166-row
    LDS tile; FP32 accumulation of 151 normalized taps stays within the gate."""
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False, unsharp=True)
    prepared.with_lens_blur(p, sigma, radius=radius)
    src = synth.frame(211, 167, seed=radius)
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
    assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"


@pytest.mark.parametrize("radius", [17, 40, 75])
def test_wide_unsharp(radius):
    """Unsharp radii above the fused post kernel's tile (16) run as a separate
    blur + combine, to the same gate."""
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False, unsharp=True)
    sp = p.record.scanner_post
    sp.unsharp_sigma_px = radius / 3.0
    sp.unsharp_radius = radius
    sp.unsharp_amount = 0.8
    src = synth.frame(150, 97, seed=radius)
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
    assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"


@pytest.mark.parametrize("radius", [0, 76, 120])
def test_lens_blur_radius_limit(radius):
    """A positive sigma with a radius outside 1..75 is a preparation failure
    (ensure_gaussian_kernel, src/Cuda/JuicerCudaResources.cpp:6436-6448)."""
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_lens_blur(p, 40.0, radius=radius)
    _, out, text = render_gpu(p, synth.frame(32, 24, seed=1))
    assert out.status.category == abi.FJ_STATUS_PREPARATION_FAILURE, text
    assert "gaussian descriptor is invalid" in text


def test_lens_blur_with_unused_wide_unsharp():
    """An unsharp kernel with amount 0 is prepared but never launched: lens
    blur renders whatever its (valid) radius."""
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False, unsharp=True)
    prepared.with_lens_blur(p, 2.0)
    sp = p.record.scanner_post
    sp.unsharp_sigma_px, sp.unsharp_radius, sp.unsharp_amount = 10.0, 30, 0.0
    src = synth.frame(64, 48, seed=3)
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    assert np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1)).max() <= 2e-4
