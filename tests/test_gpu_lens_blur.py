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


def test_lens_blur_radius_limit():
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=False)
    prepared.with_lens_blur(p, 40.0, radius=120)
    _, out, text = render_gpu(p, synth.frame(32, 24, seed=1))
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "lens blur" in text
