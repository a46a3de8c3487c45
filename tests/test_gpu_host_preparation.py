"""The call scope's host preparation on the GPU (FjHostPreparation,
juicer_cuda_call_scope.h): when the host's render passes scanner-LUT and
Gaussian-kernel callbacks, the frame's scanner LUT (on a cold LUT) and its
Gaussian kernels come from them.

* callbacks that return the library's own builders' results render the same
  bits as a render without host preparation, and both callbacks are called;
* a host whose Gaussian weights differ changes the frame (the library uses
  the host's kernels, not its own);
* a host refusal reaches the render's status and text unchanged (the
  reference's JuicerCudaScannerLut / gaussian kernel preparation forward the
  host's typed failure);
* a view that breaks the callback's contract is refused with the reference's
  texts: "Gaussian kernel projection invalid" (preparation failure,
  tests/resource/gaussian_kernel_test.cpp:166-169) and "Scanner LUT
  preparation: invalid transport contract" (internal failure,
  tests/resource/scanner_lut_test.cpp:590-594).

The scanner LUT is cached per prepared scanner identity, so each case that
needs the LUT callback gives the scanner LUT a hash not rendered before.
"""
import itertools

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import call_harness as H  # noqa: E402
from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402

W, HGT = 160, 96
PX = descriptors.pixel_size_for(3840, 2160)
OPS = [H.PRECHECK, H.ADMIT, H.INSPECT, H.RENDER, H.RELEASE, H.COMPLETE]
RENDER_AT = OPS.index(H.RENDER)
_fresh = itertools.count(1)


def scene(fresh_lut):
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=True, unsharp=True)
    prepared.with_lens_blur(p, 1.5)
    if fresh_lut:
        p.record.scanner_lut.hash = 0x5EED_0000_0000 + next(_fresh)
    return p


def render(p, src, host_mode):
    r = runtime.shared_renderer()
    s = torch.from_numpy(src).cuda()
    d = torch.full_like(s, -3.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), W, HGT, 4, stream=torch.cuda.current_stream().cuda_stream,
                          pixel_size_um=PX)
    sub = runtime.submission_for(p.record)
    scoped = abi.FjPreparedHostData.from_buffer_copy(p.record)
    out = H.run(r.handle, OPS, route=int(p.record.route), frame=f, prepared=scoped, submission=sub,
                host_mode=host_mode)
    torch.cuda.synchronize()
    return out, d.cpu().numpy()


def bare(p, src):
    r = runtime.shared_renderer()
    s = torch.from_numpy(src).cuda()
    d = torch.full_like(s, -3.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), W, HGT, 4, stream=torch.cuda.current_stream().cuda_stream,
                          pixel_size_um=PX)
    r.render(f, p.record)
    torch.cuda.synchronize()
    return d.cpu().numpy()


def test_host_builders_render_the_native_frame():
    src = synth.frame(W, HGT, seed=41)
    p = scene(fresh_lut=True)
    out, got = render(p, src, H.HOST_DELEGATE)
    assert out.op_status(RENDER_AT) == abi.FJ_STATUS_SUCCESS, out.op_text(RENDER_AT)
    assert out.host_calls[0] == 1, "cold scanner LUT comes from the host"
    assert out.host_calls[1] >= 2, "lens blur and unsharp kernels come from the host"
    np.testing.assert_array_equal(got, bare(scene(fresh_lut=False), src))
    # Warm LUT: only the Gaussian callback is visited.
    out, again = render(p, src, H.HOST_DELEGATE)
    assert out.op_status(RENDER_AT) == abi.FJ_STATUS_SUCCESS and out.host_calls[0] == 0 and out.host_calls[1] >= 2
    np.testing.assert_array_equal(again, got)


def test_host_kernels_are_the_ones_used():
    src = synth.frame(W, HGT, seed=42)
    out, got = render(scene(fresh_lut=False), src, H.HOST_GAUSSIAN_HALF)
    assert out.op_status(RENDER_AT) == abi.FJ_STATUS_SUCCESS, out.op_text(RENDER_AT)
    native = bare(scene(fresh_lut=False), src)
    assert np.abs(got[..., :3] - native[..., :3]).max() > 1e-2


@pytest.mark.parametrize("mode, fresh, category, text", [
    (H.HOST_GAUSSIAN_FAIL, False, abi.FJ_STATUS_PREPARATION_FAILURE, "host gaussian refused: sigma"),
    (H.HOST_LUT_FAIL, True, abi.FJ_STATUS_PREPARATION_FAILURE, "host scanner LUT refused: non-finite node"),
    (H.HOST_GAUSSIAN_BROKEN, False, abi.FJ_STATUS_PREPARATION_FAILURE, "Gaussian kernel projection invalid"),
    (H.HOST_LUT_BROKEN, True, abi.FJ_STATUS_INTERNAL_FAILURE, "Scanner LUT preparation: invalid transport contract"),
])
def test_host_failures_reach_the_render(mode, fresh, category, text):
    src = synth.frame(W, HGT, seed=43)
    out, got = render(scene(fresh_lut=fresh), src, mode)
    assert out.op_status(RENDER_AT) == category, out.op_text(RENDER_AT)
    assert out.op_text(RENDER_AT) == text
    assert out.completion_diagnostic.decode() == text
    assert np.all(got == -3.0), "a refused frame writes nothing"
