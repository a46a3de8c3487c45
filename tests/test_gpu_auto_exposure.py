"""GPU parity of camera auto-exposure metering (fj_autoexposure.hip, after
src/Cuda/JuicerCudaAutoExposure.cu) against the FP64 oracle's meter
(oracle/fj_oracle.c: fjo_auto_exposure): every method on the print route,
the direct route, the camera-carrier path (scatter/halation on, where the
metered scale multiplies the carrier without the route correction), a batch
whose frames meter concurrently, and the boundary's refusals (median on the
direct route, source bounds other than the source image's, an empty preview).
Gate: linear output max |err| <= 2e-4, alpha bit-exact.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402

W, H = 160, 96
METHODS = [abi.FJ_METER_CENTER_WEIGHTED, abi.FJ_METER_AVERAGE, abi.FJ_METER_MEDIAN, abi.FJ_METER_PARTIAL,
           abi.FJ_METER_MATRIX, abi.FJ_METER_MULTI_ZONE, abi.FJ_METER_HIGHLIGHT_WEIGHTED]


def render_gpu(p, src):
    h, w, c = src.shape
    s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream)
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy(), out, text


def frame(w=W, h=H, seed=11, gain=1.0):
    src = synth.frame(w, h, seed=seed)
    src[..., :3] *= gain
    return src


def check(p, src):
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
    assert err.max() <= 2e-4, f"max {err.max():.3e} mean {err.mean():.3e}"
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    return got


@pytest.mark.parametrize("method", METHODS)
def test_print_route_methods(method):
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_auto_exposure(p, W, H, method)
    src = frame(gain=0.3)
    scale = oracle.Plan(p).auto_exposure(src)
    assert np.isfinite(scale) and scale > 0.0 and abs(scale - 1.0) > 1e-3, scale  # metering took effect
    check(p, src)


@pytest.mark.parametrize("method", [abi.FJ_METER_AVERAGE, abi.FJ_METER_MATRIX])
def test_direct_route_and_meter_window(method):
    p = prepared.simple_view(prepared.load("NegativeDirectHanatos"), encode=False)
    prepared.with_auto_exposure(p, W, H, method, meter=(20, 10, 130, 80))
    check(p, frame(seed=12, gain=2.5))


def test_metering_cancels_an_input_gain():
    """A power-of-two input gain changes the metered scale by its inverse:
    the rendered frames agree."""
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_auto_exposure(p, W, H, abi.FJ_METER_CENTER_WEIGHTED)
    a, _, _ = render_gpu(p, frame(seed=13))
    b, _, _ = render_gpu(p, frame(seed=13, gain=4.0))
    np.testing.assert_allclose(a, b, atol=2e-6)


def test_carrier_path_with_halation():
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_halation(p, W, H, 2.0, primary_amount=(0.3, 0.1, 0.015))
    prepared.with_auto_exposure(p, W, H, abi.FJ_METER_MULTI_ZONE)
    check(p, frame(seed=14, gain=0.5))


def test_batch_frames_meter_independently():
    records, frames, keep = [], [], []
    rng = np.random.default_rng(3)
    for i in range(6):
        p = prepared.simple_view(prepared.load("PrintPreflash"), encode=True)
        prepared.with_auto_exposure(p, W, H, METHODS[i % len(METHODS)])
        src = frame(seed=20 + i, gain=float(rng.uniform(0.2, 3.0)))
        s = torch.from_numpy(np.ascontiguousarray(src)).cuda()
        d = torch.full_like(s, -7.0)
        frames.append(runtime.frame_for(s.data_ptr(), d.data_ptr(), W, H, 4,
                                        stream=torch.cuda.current_stream().cuda_stream))
        records.append(p.record)
        keep.append((p, src, s, d))
    st, text, outs = runtime.shared_renderer().render_batch(frames, records)
    torch.cuda.synchronize()
    assert st.category == abi.FJ_STATUS_SUCCESS, text
    for (p, src, s, d), out in zip(keep, outs):
        assert out.status.category == abi.FJ_STATUS_SUCCESS
        single, _, _ = render_gpu(p, src)
        np.testing.assert_array_equal(d.cpu().numpy(), single)


def test_refusals():
    src = frame()
    p = prepared.simple_view(prepared.load("NegativeDirectHanatos"), encode=False)
    prepared.with_auto_exposure(p, W, H, abi.FJ_METER_MEDIAN)
    _, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "Median" in text
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    p.record.auto_exposure = prepared.auto_exposure_descriptor((0, 0, W + 8, H), (0, 0, W, H), 1)
    p.record.film_exposure.flags |= abi.FJ_FILM_AUTO_EXPOSURE
    _, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT, text
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    p.record.auto_exposure = prepared.auto_exposure_descriptor((0, 0, W, H), (5, 5, 5, 5), 1)
    p.record.film_exposure.flags |= abi.FJ_FILM_AUTO_EXPOSURE
    _, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "auto_exposure_buffers" in text
