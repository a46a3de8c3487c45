"""GPU parity of the staged schedules against the FP64 oracle.

* scanner post stage: unsharp mask (sigma 0.7 px, radius 2, amount 0.7) on
  the linear planes, then output encoding;
* spatial DIR: raw donor corrections, the DCT-domain core+tail filter
  (20 um core, 200 um tail, weight 0.03), final development.
Gate: linear output max |err| <= 2e-4; encoded output within the reference's
pixel gate 2e-4 + 3e-4*|ref|.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402


def render_gpu(p, src):
    h, w, c = src.shape
    s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream)
    out = runtime.shared_renderer().render(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy(), out


def configure(case, w, h, *, dir_spatial, unsharp, encode, tail_um=200.0, tail_weight=0.03):
    p = prepared.simple_view(prepared.load(case), encode=encode, unsharp=unsharp)
    if dir_spatial:
        descriptors.spatial_dir(p.record, 20.0, tail_um, tail_weight, descriptors.pixel_size_for(w, h), w, h)
    return p


def compare(got, src, dst, lin, linear):
    if linear:
        err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
        assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"
    else:
        ref = dst[..., :3].astype(np.float64)
        err = np.abs(got[..., :3] - ref)
        assert np.all(err <= 2e-4 + 3e-4 * np.abs(ref)), f"max {err.max():.3e}"
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    return err


@pytest.mark.parametrize("linear", [True, False], ids=["linear", "encoded"])
@pytest.mark.parametrize("case", ["NegativePrintHanatos", "PrintPreflash", "NegativeDirectHanatos"])
def test_unsharp_post(case, linear):
    src = synth.frame(200, 120, seed=4)
    p = configure(case, 200, 120, dir_spatial=False, unsharp=True, encode=not linear)
    assert p.record.scanner_post.unsharp_radius == 2
    got, _ = render_gpu(p, src)
    dst, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = compare(got, src, dst, lin, linear)
    print(f"unsharp {case}: max {err.max():.3e}")


@pytest.mark.parametrize("shape", [(257, 131), (240, 135), (64, 48), (1, 7), (5, 1)])
@pytest.mark.parametrize("case", ["NegativePrintHanatos", "NegativeDirectHanatos"])
def test_spatial_dir(case, shape):
    w, h = shape
    src = synth.frame(w, h, seed=w * 31 + h)
    p = configure(case, w, h, dir_spatial=True, unsharp=False, encode=False)
    got, out = render_gpu(p, src)
    _, lin, flags = oracle.Plan(p).render(src, threads=4)
    err = compare(got, src, None, lin, True)
    print(f"spatial DIR {case} {shape}: max {err.max():.3e} mean {err.mean():.3e}")


@pytest.mark.parametrize("linear", [True, False], ids=["linear", "encoded"])
def test_spatial_dir_with_unsharp_long_tail(linear):
    w, h = 384, 216
    src = synth.frame(w, h, seed=8)
    # tail 400 um / weight 0.06: the long-tail form of the DIR stage
    p = configure("PrintPreflash", w, h, dir_spatial=True, unsharp=True, encode=not linear, tail_um=400.0,
                  tail_weight=0.06)
    got, _ = render_gpu(p, src)
    dst, lin, _ = oracle.Plan(p).render(src, threads=4)
    compare(got, src, dst, lin, linear)


@pytest.mark.parametrize("shape,pitch_width,tail_um,tail_weight,max_sigma", [
    ((640, 360), 3840, 200.0, 0.03, 60.0),     # the UHD frame's pitch: tails sigma 11.8 / 33.4 / 60.8 px
    ((1152, 648), 8192, 400.0, 0.06, 259.0)])  # the 8192-wide frame's pitch, 400 um tail: sigma up to 259 px
def test_spatial_dir_at_benchmark_pitch(shape, pitch_width, tail_um, tail_weight, max_sigma):
    """Spatial DIR at the pixel pitch of the benchmarked frames, on frames
    large enough to hold the tails, against the FP64 oracle."""
    w, h = shape
    pitch = 35000.0 / pitch_width
    src = synth.frame(w, h, seed=w + 7 * h)
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    descriptors.spatial_dir(p.record, 20.0, tail_um, tail_weight, pitch, w, h)
    sd = p.record.spatial_dir
    sigmas = [sd.components[i].sigma_px for i in range(sd.component_count)]
    assert max(sigmas) >= max_sigma, sigmas
    got, _ = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=8)
    err = compare(got, src, None, lin, True)
    print(f"spatial DIR {shape} pitch {pitch:.4f} um sigmas {sigmas}: max {err.max():.3e} mean {err.mean():.3e}")


def test_spatial_dir_moves_the_image():
    """DIR spatial tails must change the result vs point DIR (guards against a no-op filter)."""
    w, h = 160, 90
    src = synth.frame(w, h, seed=12)
    a, _ = render_gpu(configure("NegativePrintHanatos", w, h, dir_spatial=True, unsharp=False, encode=False), src)
    b, _ = render_gpu(configure("NegativePrintHanatos", w, h, dir_spatial=False, unsharp=False, encode=False), src)
    assert np.abs(a[..., :3] - b[..., :3]).max() > 1e-3


def test_uhd_staged_properties():
    """Full UHD frame through DIR + unsharp: finite, in range, alpha exact, repeatable."""
    w, h = 3840, 2160
    src = synth.frame(w, h, seed=20261203)
    p = configure("PrintPreflash", w, h, dir_spatial=True, unsharp=True, encode=True)
    a, _ = render_gpu(p, src)
    b, _ = render_gpu(p, src)
    assert np.isfinite(a).all() and a[..., :3].min() >= 0.0 and a[..., :3].max() <= 1.0
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a[..., 3], src[..., 3])
