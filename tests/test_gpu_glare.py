"""Scanner glare on the GPU against the oracle.

* stage: fjx_glare_field's raw uniforms equal the oracle's Philox-derived
  FP32 uniforms bit for bit (the reference's words, pinned in
  tests/test_oracle_glare.py), at a negative window origin and for both
  media; the unblurred field matches the oracle's FP64 lognormal within FP32
  rounding of its transcendental steps (relative 2e-6), and the blurred field
  matches a numpy half-sample-reflected Gaussian of the oracle's field;
* frames: print routes with the plug-in's default glare (0.03 %, roughness
  0.7, blur 0.5 px) and a strong one, a direct route, a source image whose
  bounds start away from the origin and a non-zero frame token, each within
  2e-4 of the oracle in linear output; the strong glare visibly changes the
  frame.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from fj import abi, prepared, runtime, synth  # noqa: E402
from oracle import oracle  # noqa: E402


def post_of(percent, roughness, blur):
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_glare(p, percent, roughness, blur)
    return p.record.scanner_post


def gpu_field(post, seed, medium, origin, w, h, uniforms=False):
    f = torch.full((h, w), -1.0, dtype=torch.float32, device="cuda")
    u = torch.full((2, h, w), -1.0, dtype=torch.float32, device="cuda") if uniforms else None
    err = runtime.ErrorBuffer()
    st = runtime.lib().fjx_glare_field(runtime.shared_renderer().handle, 0, C.byref(post), seed, medium, origin[0],
                                       origin[1], w, h, f.data_ptr(), u.data_ptr() if u is not None else None,
                                       torch.cuda.current_stream().cuda_stream, err.ptr())
    assert st.category == 0, err.text
    return f.cpu().numpy(), (u.cpu().numpy() if u is not None else None)


@pytest.mark.parametrize("medium", [1, 0])
def test_glare_uniforms_bit_exact_and_field(medium):
    w, h, origin = 96, 48, (-17, -3)
    post = post_of(0.03, 0.7, 0.0)
    seed = oracle.glare_seed(0xC0FFEE1234567, 12, origin[0], origin[1])
    field, uni = gpu_field(post, seed, medium, origin, w, h, uniforms=True)
    want = np.empty((2, h, w), np.float32)
    for y in range(h):
        for x in range(w):
            want[:, y, x] = oracle.glare_uniforms(seed, origin[0] + x, origin[1] + y, medium)
    np.testing.assert_array_equal(uni.view(np.uint32), want.view(np.uint32))
    ref = oracle.glare_plane(post, seed, medium, w, h, origin=origin)
    rel = np.abs(field - ref) / ref
    print(f"medium {medium}: field max rel err {rel.max():.3e}")
    assert rel.max() <= 2e-6


def blur_reflect(plane, weights):
    r = len(weights) // 2
    pad = np.pad(plane, r, mode="symmetric")  # half-sample reflection
    h, w = plane.shape
    tmp = sum(weights[k] * pad[r:r + h, k:k + w] for k in range(2 * r + 1))
    tmp = np.pad(tmp, ((r, r), (0, 0)), mode="symmetric")
    return sum(weights[k] * tmp[k:k + h, :] for k in range(2 * r + 1))


@pytest.mark.parametrize("sigma", [0.5, 3.0])
def test_blurred_glare_field(sigma):
    w, h = 320, 200
    post = post_of(0.03, 0.7, sigma)
    seed = oracle.glare_seed(99, 0, 0, 0)
    field, _ = gpu_field(post, seed, 1, (0, 0), w, h)
    weights = oracle.gaussian_weights(post.glare_blur_radius, post.glare_blur_sigma_px).astype(np.float64)
    ref = blur_reflect(oracle.glare_plane(post, seed, 1, w, h), weights)
    rel = np.abs(field - ref) / ref
    print(f"sigma {sigma} (radius {post.glare_blur_radius}): max rel err {rel.max():.3e}")
    assert rel.max() <= 5e-6


def render_gpu(p, src, time_frames=0.0, origin=(0, 0)):
    h, w, c = src.shape
    s = torch.from_numpy(src).cuda()
    d = torch.full_like(s, -7.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), w, h, c, stream=torch.cuda.current_stream().cuda_stream,
                          time_frames=time_frames)
    if origin != (0, 0):
        x1, y1 = origin
        f.source = runtime.image(s.data_ptr(), w, h, c, x1=x1, y1=y1)
        f.destination = runtime.image(d.data_ptr(), w, h, c, x1=x1, y1=y1)
        f.render_window = abi.FjRect(x1, y1, x1 + w, y1 + h)
        f.full_frame_extent = abi.FjRect(x1, y1, x1 + w, y1 + h)
    out, text = runtime.shared_renderer().render_raw(f, p.record)
    torch.cuda.synchronize()
    assert out.status.category == 0, text
    return d.cpu().numpy()


CASES = [
    ("PrintPreflash", 0.03, 0.7, 0.5, 0.0, (0, 0)),
    ("PrintPreflash", 2.0, 0.7, 1.5, 7.0, (0, 0)),
    ("NegativePrintHanatos", 1.0, 0.5, 0.5, 3.0, (120, 45)),
    ("NegativeDirectHanatos", 1.0, 0.7, 0.0, 0.0, (0, 0)),
]


@pytest.mark.parametrize("case, percent, roughness, sigma, t, origin", CASES)
def test_glare_frame_parity(case, percent, roughness, sigma, t, origin):
    p = prepared.simple_view(prepared.load(case), encode=False)
    prepared.with_glare(p, percent, roughness, sigma)
    src = synth.frame(200, 120, seed=21)
    got = render_gpu(p, src, t, origin)
    _, lin, _ = oracle.Plan(p).set_frame(int(np.floor(t)), origin).render(src, threads=4)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0.0, 1.0))
    print(f"{case} glare {percent}/{roughness}/{sigma} t={t} origin={origin}: max {err.max():.3e}")
    assert err.max() <= 2e-4
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    if percent >= 1.0:
        plain = prepared.simple_view(prepared.load(case), encode=False)
        base = render_gpu(plain, src, t, origin)
        assert np.abs(got[..., :3] - base[..., :3]).max() > 1e-3
