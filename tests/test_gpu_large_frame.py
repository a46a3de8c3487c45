"""The largest single frame of the workload set: 8192 x 4320 print route with
1 % speculars at 64.0, scatter/halation at four times the default spatial
scale with five bounces, camera and enlarger diffusion at the radius cap
(2159 px: 13824 x 8640 padded planes), spatial DIR with a 400 um tail at
weight 0.06, density-space grain (seed 7) and unsharp. A whole-frame FP64
oracle is out of reach at this size (the diffusion alone is 1.9e13 taps per
channel); each spatial stage is checked against the oracle's operators on
smaller frames and, for diffusion, at this size at sample pixels
(tests/test_gpu_diffusion_long.py). Here the frame must render, stay finite
and in range, keep alpha bit-exact, repeat bit for bit, and respond to its
inputs; the per-stage times are printed.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402

W, H = 8192, 4320
PX = descriptors.pixel_size_for(W, H)


def workload(optics=True):
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=True, unsharp=True)
    p.record.print.preflash_exposure = 0.01
    descriptors.spatial_dir(p.record, 20.0, 400.0, 0.06, PX, W, H)
    if optics:
        prepared.with_halation(p, W, H, PX, halation_spatial_scale=4.0, bounce_count=5)
        mist = descriptors.diffusion_recipe(2, 1.0)
        prepared.with_diffusion(p, W, H, PX, camera=mist, enlarger=mist)
    prepared.with_grain(p, W, H, PX, session_seed=7, noise_shape=(512, 512, 16))
    return p


def test_large_frame_renders_and_repeats():
    p = workload()
    assert p.record.optics.camera.radius_px == 2159 and p.record.optics.enlarger.radius_px == 2159
    src = synth.frame(W, H, seed=20261203, specular_fraction=0.01)
    assert (src[..., 0] == 64.0).mean() > 0.009
    s = torch.from_numpy(src).cuda()
    outs = []
    r = runtime.shared_renderer()
    for k in range(2):
        d = torch.full_like(s, -5.0)
        f = runtime.frame_for(s.data_ptr(), d.data_ptr(), W, H, 4, stream=torch.cuda.current_stream().cuda_stream,
                              pixel_size_um=PX, session_seed=7)
        if k == 1:
            f.flags |= abi.FJ_FRAME_TRACE_VERBOSE
        out, text = r.render_raw(f, p.record)
        torch.cuda.synchronize()
        assert out.status.category == abi.FJ_STATUS_SUCCESS, text
        outs.append(d.cpu().numpy())
    a, b = outs
    np.testing.assert_array_equal(a, b)
    assert np.isfinite(a).all() and a[..., :3].min() >= 0.0 and a[..., :3].max() <= 1.0
    np.testing.assert_array_equal(a[..., 3], src[..., 3])
    # The highlights bloom: pixels next to a specular are brighter than the
    # same pixels rendered without diffusion and halation.
    q = workload(optics=False)
    d = torch.full_like(s, -5.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), W, H, 4, stream=torch.cuda.current_stream().cuda_stream,
                          pixel_size_um=PX, session_seed=7)
    out, text = r.render_raw(f, q.record)
    torch.cuda.synchronize()
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    plain = d.cpu().numpy()
    lum = lambda x: x[..., :3].mean(axis=-1)  # noqa: E731
    dark = src[..., :3].max(axis=-1) < 0.05
    assert lum(a)[dark].mean() > lum(plain)[dark].mean() + 1e-3
    timings = r.stage_timings(torch.cuda.current_device())
    total = sum(ms for _, ms in timings)
    print(f"8192x4320 print frame: {total:.2f} ms over stages")
    for name, ms in timings:
        print(f"  {name:28s} {ms:8.3f} ms")
    filt = [ms for name, ms in timings if name.endswith("diffusion.filter")]
    if filt:
        pw, ph = 13824, 8640
        gb = 3 * (8 * pw * ph + 4 * pw * H) / (np.mean(filt) * 1e-3) / 1e9
        print(f"  diffusion column filter: {np.mean(filt):.3f} ms, {gb:.0f} GB/s algorithmic")
