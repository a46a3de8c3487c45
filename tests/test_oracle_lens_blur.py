"""The oracle's scanner lens blur (oracle/fj_oracle.c: lens_plane, after
blur_plane_in_place, src/Cuda/Scan/JuicerCudaScanPipeline.cu:2686-2731):
normalized Gaussian taps (float32 taps: their sum is 1 to ~1e-7) with
half-sample reflection leave a uniform frame unchanged at any radius (also
radii beyond the frame), and a non-uniform frame is smoothed (each
channel's spatial variance drops)."""
import numpy as np
import pytest

from fj import prepared
from oracle import oracle


@pytest.mark.parametrize("sigma", [0.7, 3.0, 15.0])
def test_uniform_frame_unchanged(sigma):
    src = np.empty((20, 30, 4), np.float32)
    src[..., :3] = (0.21, 0.17, 0.12)
    src[..., 3] = 1.0
    plain = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    _, a, _ = oracle.Plan(plain).render(src)
    blurred = prepared.with_lens_blur(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), sigma)
    _, b, _ = oracle.Plan(blurred).render(src)
    np.testing.assert_allclose(b, a, rtol=1e-6, atol=0)


def test_smooths_a_checkerboard():
    y, x = np.mgrid[0:24, 0:32]
    src = np.empty((24, 32, 4), np.float32)
    src[..., :3] = np.where(((x + y) % 2)[..., None] == 0, 0.05, 0.4)
    src[..., 3] = 1.0
    plain = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    _, a, _ = oracle.Plan(plain).render(src)
    blurred = prepared.with_lens_blur(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), 1.5)
    _, b, _ = oracle.Plan(blurred).render(src)
    for c in range(3):
        assert b[..., c].var() < 0.05 * a[..., c].var()


@pytest.mark.parametrize("radius", [0, 76])
def test_invalid_radius_refused(radius):
    """A positive sigma with a radius outside 1..75 fails preparation
    (ensure_gaussian_kernel, src/Cuda/JuicerCudaResources.cpp:6436-6448)."""
    p = prepared.with_lens_blur(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), 4.0,
                                radius=radius)
    with pytest.raises(RuntimeError, match="rc=4"):
        oracle.Plan(p).render(np.ones((8, 8, 4), np.float32))


def test_unsharp_without_sigma_is_skipped():
    """The unsharp kernel is prepared only for a positive sigma
    (ProcessRoot.cpp:3138-3146); without one the mask is not applied, whatever
    its amount and radius."""
    src = np.random.default_rng(3).uniform(0.01, 0.5, (12, 16, 4)).astype(np.float32)
    plain = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    _, a, _ = oracle.Plan(plain).render(src)
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False, unsharp=True)
    sp = p.record.scanner_post
    sp.unsharp_sigma_px, sp.unsharp_radius, sp.unsharp_amount = 0.0, 5, 0.9
    _, b, _ = oracle.Plan(p).render(src)
    np.testing.assert_array_equal(a, b)
