"""GPU parity of the camera and enlarger diffusion stages against the FP64 oracle.

Stage descriptors come from fj.descriptors (the resolved controls, sample and
stage identities, pinned to the reference's diffusion cohort); the pixel pitch
sets the PSF radius (ceil(20 lambda_bloom / pitch), capped at half the short
side), so the cases cover small supports, supports at the cap, odd and
degenerate frame sides, and the full bench stage set. Gate: linear output
max |err| <= 2e-4 (encoded: 2e-4 + 3e-4 |ref|), alpha bit-exact.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402


def render_gpu(p, src, window=None):
    h, w, c = src.shape
    s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream,
                          window=window)
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy(), out, text


def configure(case, w, h, px, *, camera=None, enlarger=None, encode=False, unsharp=False, dir_spatial=False,
              halation=False):
    p = prepared.simple_view(prepared.load(case), encode=encode, unsharp=unsharp)
    if dir_spatial:
        descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, descriptors.pixel_size_for(w, h), w, h)
    if halation:
        prepared.with_halation(p, w, h, px, primary_amount=(0.3, 0.1, 0.015))
    return prepared.with_diffusion(p, w, h, px, camera=camera, enlarger=enlarger)


def check(got, out, text, src, dst, lin, linear):
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    if linear:
        err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
        assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"
    else:
        ref = dst[..., :3].astype(np.float64)
        err = np.abs(got[..., :3] - ref)
        assert np.all(err <= 2e-4 + 3e-4 * np.abs(ref)), f"max {err.max():.3e}"
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    return err


MIST = descriptors.diffusion_recipe(1, 0.5)                       # black pro-mist
GLIMMER = descriptors.diffusion_recipe(2, 1.0, halo_warmth=0.6)   # warm halo
CINEBLOOM = descriptors.diffusion_recipe(3, 2.0, spatial_scale=0.7)


@pytest.mark.parametrize("shape,px", [((160, 90), 600.0), ((97, 61), 250.0), ((64, 48), 30.0), ((300, 40), 400.0),
                                      ((7, 5), 100.0), ((2, 9), 50.0)])
def test_camera_diffusion_shapes(shape, px):
    w, h = shape
    src = synth.frame(w, h, seed=w * 5 + h)
    p = configure("NegativePrintHanatos", w, h, px, camera=MIST)
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, out, text, src, None, lin, True)
    print(f"camera diffusion {shape} r={p.record.optics.camera.radius_px}: max {err.max():.3e} mean {err.mean():.3e}")


@pytest.mark.parametrize("recipe", [MIST, GLIMMER, CINEBLOOM], ids=["mist", "glimmer", "cinebloom"])
def test_enlarger_diffusion_families(recipe):
    w, h = 150, 100
    src = synth.frame(w, h, seed=31)
    p = configure("PrintPreflash", w, h, 350.0, enlarger=recipe)
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, out, text, src, None, lin, True)
    print(f"enlarger diffusion r={p.record.optics.enlarger.radius_px}: max {err.max():.3e}")


def test_camera_diffusion_direct_route():
    w, h = 128, 72
    src = synth.frame(w, h, seed=5)
    p = configure("NegativeDirectHanatos", w, h, 300.0, camera=GLIMMER)
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    check(got, out, text, src, None, lin, True)


@pytest.mark.parametrize("linear", [True, False], ids=["linear", "encoded"])
def test_bench_stage_set(linear):
    """Carrier + camera diffusion + halation, spatial DIR, print-linear +
    enlarger diffusion, print, scan, unsharp."""
    w, h = 192, 108
    src = synth.frame(w, h, seed=8)
    px = 400.0
    p = configure("PrintPreflash", w, h, px, camera=MIST, enlarger=MIST, encode=not linear, unsharp=True,
                  dir_spatial=True, halation=True)
    got, out, text = render_gpu(p, src)
    dst, lin, _ = oracle.Plan(p).render(src, threads=4)
    check(got, out, text, src, dst, lin, linear)


def test_diffusion_at_radius_cap():
    w, h = 120, 80
    src = synth.frame(w, h, seed=9)
    p = configure("PrintPreflash", w, h, 5.0, camera=CINEBLOOM, enlarger=GLIMMER)
    assert p.record.optics.camera.radius_px == h // 2 - 1
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    check(got, out, text, src, None, lin, True)


def test_diffusion_moves_the_image():
    w, h = 160, 90
    src = synth.frame(w, h, seed=12)
    a, out, text = render_gpu(configure("PrintPreflash", w, h, 200.0, camera=CINEBLOOM), src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    b, _, _ = render_gpu(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), src)
    assert np.abs(a[..., :3] - b[..., :3]).max() > 1e-2


def _expect(p, src, category, fragment):
    got, out, text = render_gpu(p, src)
    assert out.status.category == category, text
    assert fragment in text
    assert np.all(got == -7.0)


def test_malformed_diffusion_rejected():
    w, h = 64, 48
    src = synth.frame(w, h, seed=2)
    p = configure("PrintPreflash", w, h, 300.0, camera=MIST)
    p.record.optics.camera.warmth += 0.25  # fields no longer match the sample identity
    _expect(p, src, abi.FJ_STATUS_PREPARATION_FAILURE, "InvalidDiffusionHostBehavior field=sample_descriptor_hash")
    p = configure("PrintPreflash", w, h, 300.0, camera=MIST)
    p.record.optics.diffusion_hash = 0
    _expect(p, src, abi.FJ_STATUS_UNSUPPORTED_INPUT, "optics.diffusion_hash")
    p = configure("PrintPreflash", w, h, 300.0, camera=MIST)
    p.record.optics.camera.family = 4
    _expect(p, src, abi.FJ_STATUS_UNSUPPORTED_INPUT, "optics.diffusion")
    p = configure("PrintPreflash", w, h, 300.0)
    p.record.optics.diffusion_hash = 99
    _expect(p, src, abi.FJ_STATUS_UNSUPPORTED_INPUT, "optics.inactive_diffusion")
    p = configure("NegativeDirectHanatos", w, h, 300.0)
    p = prepared.with_diffusion(p, w, h, 300.0, enlarger=MIST)
    _expect(p, src, abi.FJ_STATUS_UNSUPPORTED_INPUT, "optics.flags")
    p = configure("PrintPreflash", w, h, 300.0, camera=MIST)
    p.record.optics.full_frame.width = w - 1
    _expect(p, src, abi.FJ_STATUS_UNSUPPORTED_INPUT, "optics.frame")


def test_uhd_diffusion_constant_frame():
    """A constant frame stays constant through both stages (normalized PSF,
    mirror boundary): the UHD bench stage set with and without diffusion
    renders the same constant image. Exercises the bench radii (834 px)
    and padded sizes (5625 x 3840)."""
    w, h = 3840, 2160
    src = np.empty((h, w, 4), np.float32)
    src[..., :3] = np.array([0.18, 0.12, 0.3], np.float32)
    src[..., 3] = 1.0
    px = descriptors.pixel_size_for(w, h)
    p = configure("PrintPreflash", w, h, px, camera=MIST, enlarger=MIST)
    a, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    b, out, text = render_gpu(configure("PrintPreflash", w, h, px), src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    err = np.abs(a[..., :3].astype(np.float64) - b[..., :3])
    assert err.max() <= 2e-4, f"max {err.max():.3e}"
    assert p.record.optics.camera.radius_px == 834
