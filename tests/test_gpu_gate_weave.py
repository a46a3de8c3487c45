"""GPU parity of the gate weave (the scanner output pass resampling the
post-unsharp linear planes through a Mitchell filter at rotated and shifted
positions, JuicerCudaScanPipeline.cu:1514-1558, :2096-2182; bound by the
gate-output effects descriptor, JuicerCudaExecutor.cpp:1769-1773) against the
FP64 oracle, with and without the unsharp mask, on print and direct routes,
including a shift that clamps every position to the border. Gate: linear
output max |err| <= 2e-4, alpha bit-exact. Effects descriptors outside the
weave are refused with the reference's diagnostics.
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


@pytest.mark.parametrize("case,unsharp,motion,shape", [
    ("PrintPreflash", False, (1.7, -2.3, 0.4), (120, 72)),
    ("PrintPreflash", True, (-0.6, 0.9, -0.25), (97, 61)),
    ("NegativeDirectHanatos", True, (3.2, 1.1, 1.0), (64, 40)),
    ("NegativeDirectHanatos", False, (0.0, 0.0, 0.0), (40, 30)),
    ("PrintPreflash", True, (1e4, -37.5, 0.0), (53, 47)),
])
def test_gate_weave(case, unsharp, motion, shape):
    w, h = shape
    print(f"case {case} unsharp {unsharp} motion {motion} shape {shape}", flush=True)
    p = prepared.with_gate_weave(prepared.simple_view(prepared.load(case), encode=False, unsharp=unsharp), *motion)
    src = synth.frame(w, h, seed=w + 3 * h)
    got, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_SUCCESS, text
    _, lin, _ = oracle.Plan(p).render(src, threads=4, want_dst=False)
    err = np.abs(got[..., :3].astype(np.float64) - np.clip(lin, 0, 1))
    print(f"max {err.max():.3e} mean {err.mean():.3e}")
    assert err.max() <= 2e-4
    np.testing.assert_array_equal(got[..., 3], src[..., 3])
    plain, _, _ = render_gpu(prepared.simple_view(prepared.load(case), encode=False, unsharp=unsharp), src)
    assert np.abs(got - plain).max() > 1e-4  # the weave took effect


@pytest.mark.parametrize("flags,hashes,diag", [
    (abi.FJ_EFFECTS_FILM, True, "film/gate dust and scratches"),
    (abi.FJ_EFFECTS_GATE_TRANSMITTANCE, True, "film/gate dust and scratches"),
    (0, False, "field=effects.identity"),
    (32, True, "field=effects.flags"),
])
def test_effects_outside_the_weave_are_refused(flags, hashes, diag):
    p = prepared.with_gate_weave(prepared.simple_view(prepared.load("PrintPreflash"), encode=False), 1.0, 0.0, 0.0,
                                 flags=flags)
    if not hashes:
        p.record.effects.hash = 0
    src = synth.frame(32, 24, seed=1)
    _, out, text = render_gpu(p, src)
    assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert diag in text, text
