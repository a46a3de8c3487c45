"""GPU parity of the camera-carrier scatter/halation stages against the FP64 oracle.

The frame builder's descriptors (fj.descriptors.halation_frame, pinned to the
reference's frame-builder words) are composed at chosen pixel pitches so the
schedule exercises every Gaussian kind: FIR taps (sigma < 3 px), legacy
replicate-seeded YVV recursions and the factored form. Gate: linear output
max |err| <= 2e-4 (encoded: 2e-4 + 3e-4 |ref|), alpha bit-exact.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402

STRONG = (0.3, 0.1, 0.015)  # still/no antihalation amounts


def render_gpu(p, src, window=None):
    h, w, c = src.shape
    s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream,
                          window=window)
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy(), out, text


def configure(case, w, h, px, *, encode=False, unsharp=False, dir_spatial=False, legacy=True, **recipe):
    p = prepared.simple_view(prepared.load(case), encode=encode, unsharp=unsharp)
    if dir_spatial:
        descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, descriptors.pixel_size_for(w, h), w, h)
    recipe.setdefault("primary_amount", STRONG)
    return prepared.with_halation(p, w, h, px, legacy=legacy, **recipe)


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


def kinds(p):
    d = p.record.optics.scatter_halation
    c = d.channels[0]
    return (c.core.kind, tuple(c.tail[i].kind for i in range(3)), tuple(c.bounce[i].kind for i in range(d.bounce_count)))


# Column kernels: heights up to 64 x 36 rows keep each chunk in registers
# (2160: 34-row chunks and a short top chunk; 2304: the largest such height;
# 131 and 7: short top chunks and empty chunks); 2305 takes the streaming
# column kernel. Widths that are not a multiple of 16 leave idle columns.
@pytest.mark.parametrize("shape", [(257, 131), (300, 70), (64, 48), (17, 65), (1, 7), (5, 1), (40, 2160),
                                   (96, 2304), (33, 2305)])
def test_halation_shapes(shape):
    w, h = shape
    src = synth.frame(w, h, seed=w * 7 + h)
    p = configure("NegativePrintHanatos", w, h, 2.0)
    assert abi.FJ_HALATION_YVV_REPLICATE in kinds(p)[1] and abi.FJ_HALATION_FIR_REFLECT in kinds(p)[1]
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, out, text, src, None, lin, True)
    print(f"halation {shape} {kinds(p)}: max {err.max():.3e} mean {err.mean():.3e}")


@pytest.mark.parametrize("px", [0.6, 9.1146, 40.0])
@pytest.mark.parametrize("legacy", [True, False], ids=["legacy", "factored"])
def test_halation_pitches_and_forms(px, legacy):
    w, h = 240, 135
    src = synth.frame(w, h, seed=17)
    p = configure("PositivePrintHanatos", w, h, px, legacy=legacy, bounce_decay=0.5 if legacy else 0.6)
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = check(got, out, text, src, None, lin, True)
    print(f"halation px={px} {kinds(p)}: max {err.max():.3e}")


@pytest.mark.parametrize("recipe", [dict(halation_amount=0.0), dict(scatter_amount=0.0),
                                    dict(scatter_amount=2.0, halation_amount=2.0)],
                         ids=["scatter-only", "reflection-only", "maximum-amounts"])
def test_halation_stage_selection(recipe):
    w, h = 160, 90
    src = synth.frame(w, h, seed=23)
    p = configure("NegativeDirectHanatos", w, h, 3.0, **recipe)
    got, out, text = render_gpu(p, src)
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    check(got, out, text, src, None, lin, True)


@pytest.mark.parametrize("linear", [True, False], ids=["linear", "encoded"])
def test_halation_with_dir_and_unsharp(linear):
    """The bench stage set: carrier + halation, spatial DIR, print, unsharp."""
    w, h = 384, 216
    src = synth.frame(w, h, seed=8)
    p = configure("PrintPreflash", w, h, descriptors.pixel_size_for(w, h) / 10.0, encode=not linear, unsharp=True,
                  dir_spatial=True)
    got, out, text = render_gpu(p, src)
    dst, lin, _ = oracle.Plan(p).render(src, threads=4)
    check(got, out, text, src, dst, lin, linear)


def test_halation_moves_the_image():
    w, h = 160, 90
    src = synth.frame(w, h, seed=12)
    a, out, text = render_gpu(configure("NegativePrintHanatos", w, h, 3.0), src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    b, _, _ = render_gpu(prepared.simple_view(prepared.load("NegativePrintHanatos"), encode=False), src)
    assert np.abs(a[..., :3] - b[..., :3]).max() > 1e-2


def test_malformed_halation_rejected():
    w, h = 64, 48
    src = synth.frame(w, h, seed=2)
    p = configure("NegativePrintHanatos", w, h, 3.0)
    p.record.optics.scatter_halation.channels[1].tail[0].radius = 12
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert "optics.scatter_halation" in text
    assert np.all(got == -7.0)


def test_partial_window_rejected_with_optics():
    w, h = 64, 48
    src = synth.frame(w, h, seed=2)
    p = configure("NegativePrintHanatos", w, h, 3.0)
    got, out, text = render_gpu(p, src, window=(0, 0, 32, 48))
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert "complete frame" in text


def test_uhd_halation_properties():
    """UHD print route with halation + DIR + unsharp: finite, in range,
    repeatable, alpha exact, and the halation stages take effect."""
    w, h = 3840, 2160
    src = synth.frame(w, h, seed=20261203)
    px = descriptors.pixel_size_for(w, h)
    p = configure("PrintPreflash", w, h, px, encode=True, unsharp=True, dir_spatial=True)
    a, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    b, _, _ = render_gpu(p, src)
    assert np.isfinite(a).all() and a[..., :3].min() >= 0.0 and a[..., :3].max() <= 1.0
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a[..., 3], src[..., 3])
