"""Rendering through the call scope on the GPU (juicer_cuda_call_scope.h):
the host's continuation prechecks, admits, inspects the frame, composes the
spatial-DIR and grain descriptors natively from frame math, renders and
completes. The frame must come out bit-identical to the bare fj_cuda_render
of the same records (whose parity with the oracle the other GPU tests check); the scope binds the render to the inspected context
and the descriptors to the inspected frame; a completion that reports
context loss retires the device state, and the next render rebuilds it.
"""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

import call_harness as H  # noqa: E402
from fj import abi, descriptors, prepared, runtime, synth  # noqa: E402

OK = abi.FJ_STATUS_SUCCESS
W, HGT = 160, 96
PX = descriptors.pixel_size_for(3840, 2160)


def dir_math_of(sd):
    m = abi.FjDirFrameMath()
    m.present = 1
    m.component_count = sd.component_count
    m.recipe_hash = sd.recipe_hash
    m.render_extent, m.full_frame_extent, m.filter_domain_extent = (
        sd.render_extent, sd.full_frame_extent, sd.filter_domain_extent)
    for i in range(4):
        m.components[i] = sd.components[i]
    return m


def grain_math_of(g, polarity):
    m, k = abi.FjGrainFrameMath(), abi.FjGrainCoefficients()
    m.present = 1
    m.capture_polarity = polarity
    for name, _ in abi.FjGrainFrameMath._fields_:
        if name in ("present", "capture_polarity", "requires_full_frame"):
            continue
        setattr(m, name, getattr(g, name))
    m.requires_full_frame = 1 if g.flags & abi.FJ_GRAIN_FULL_FRAME else 0
    for name, _ in abi.FjGrainCoefficients._fields_:
        setattr(k, name, getattr(g, name))
    k.flags = g.flags & ~abi.FJ_GRAIN_FULL_FRAME
    return m, k


def scene():
    p = prepared.simple_view(prepared.load("PrintPreflash"), encode=True, unsharp=True)
    descriptors.spatial_dir(p.record, 20.0, 200.0, 0.03, PX, W, HGT)
    prepared.with_grain(p, W, HGT, PX, session_seed=7)
    math = abi.FjFrameMathInput()
    math.dir = dir_math_of(p.record.spatial_dir)
    math.grain, math.grain_coefficients = grain_math_of(p.record.grain, p.record.capture_polarity)
    return p, math


def frame_on_gpu(src, session_seed=7):
    s = torch.from_numpy(src).cuda()
    d = torch.full_like(s, -3.0)
    f = runtime.frame_for(s.data_ptr(), d.data_ptr(), W, HGT, 4, stream=torch.cuda.current_stream().cuda_stream,
                          pixel_size_um=PX, session_seed=session_seed)
    return s, d, f


FULL = [H.PRECHECK, H.ADMIT, H.INSPECT, H.DESCRIPTORS, H.RENDER, H.RELEASE, H.COMPLETE]


def test_render_through_scope_matches_bare_render():
    r = runtime.shared_renderer()
    p, math = scene()
    src = synth.frame(W, HGT, seed=31)
    s, d, f = frame_on_gpu(src)
    sub = runtime.submission_for(p.record)
    scoped = abi.FjPreparedHostData.from_buffer_copy(p.record)
    out = H.run(r.handle, FULL, route=int(p.record.route), frame=f, math=math, prepared=scoped, submission=sub,
                copy_descriptors=1)
    torch.cuda.synchronize()
    for i, op in enumerate(FULL):
        if op != H.RELEASE:
            assert out.op_status(i) == OK, (op, out.op_text(i))
    assert out.completion_calls == 1 and out.call_result.fatal == 0
    assert out.completion_seen.render_outcome_valid == 1 and out.completion_seen.has_inspected_context == 1
    assert out.context.device_id == torch.cuda.current_device()
    # The scope composed real descriptor identities for both stages.
    assert out.descriptors.spatial_dir.hash != 0 and out.descriptors.grain.hash != 0
    assert out.descriptors.spatial_dir.component_count == p.record.spatial_dir.component_count
    got = d.cpu().numpy()
    # The bare entry point with the host-built records renders the same bits.
    d2 = torch.full_like(s, -3.0)
    f2 = runtime.frame_for(s.data_ptr(), d2.data_ptr(), W, HGT, 4, stream=torch.cuda.current_stream().cuda_stream,
                           pixel_size_um=PX, session_seed=7)
    r.render(f2, p.record)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(got, d2.cpu().numpy())
    assert np.isfinite(got).all() and got[..., :3].min() >= 0.0 and got[..., :3].max() <= 1.0
    np.testing.assert_array_equal(got[..., 3], src[..., 3])


def test_scope_binds_render_and_descriptors_to_the_inspection():
    r = runtime.shared_renderer()
    p, math = scene()
    src = synth.frame(W, HGT, seed=32)
    s, d, f = frame_on_gpu(src, session_seed=8)  # grain math carries session seed 7
    sub = runtime.submission_for(p.record)
    scoped = abi.FjPreparedHostData.from_buffer_copy(p.record)
    out = H.run(r.handle, [H.ADMIT, H.INSPECT, H.DESCRIPTORS, H.RENDER_OTHER_CONTEXT, H.RELEASE, H.COMPLETE],
                route=int(p.record.route), frame=f, math=math, prepared=scoped, submission=sub)
    torch.cuda.synchronize()
    assert out.op_status(1) == OK
    assert out.op_status(2) == abi.FJ_STATUS_PREPARATION_FAILURE and "frame_descriptors" in out.op_text(2)
    assert out.op_status(3) == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert out.op_text(3) == "CUDA render context differs from inspected scope"
    assert out.completion_diagnostic.decode() == "CUDA render context differs from inspected scope"
    assert np.all(d.cpu().numpy() == -3.0)


def test_rejected_render_reports_its_diagnostic_to_the_completion():
    r = runtime.shared_renderer()
    p, math = scene()
    src = synth.frame(W, HGT, seed=33)
    s, d, f = frame_on_gpu(src)
    sub = runtime.submission_for(p.record)
    bad = abi.FjPreparedHostData.from_buffer_copy(p.record)
    bad.grain.recipe_hash = 0
    out = H.run(r.handle, [H.ADMIT, H.INSPECT, H.RENDER, H.RELEASE, H.COMPLETE], route=int(p.record.route), frame=f,
                prepared=bad, submission=sub)
    torch.cuda.synchronize()
    assert out.op_status(2) == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert out.completion_seen.outcome.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert "grain" in out.completion_diagnostic.decode()


def test_context_loss_retirement_rebuilds_device_state():
    r = runtime.shared_renderer()
    p, math = scene()
    src = synth.frame(W, HGT, seed=34)
    s, d, f = frame_on_gpu(src)
    sub = runtime.submission_for(p.record)
    scoped = abi.FjPreparedHostData.from_buffer_copy(p.record)
    out = H.run(r.handle, FULL, route=int(p.record.route), frame=f, math=math, prepared=scoped, submission=sub,
                copy_descriptors=1, report_context_loss=1, recovery_scope_entered=1, retire_on_completion=1)
    torch.cuda.synchronize()
    assert out.op_status(4) == OK
    assert out.retirement == 3
    first = d.cpu().numpy()
    # Device state was released; the next render re-creates it and matches.
    d.fill_(-3.0)
    out = H.run(r.handle, FULL, route=int(p.record.route), frame=f, math=math, prepared=scoped, submission=sub,
                copy_descriptors=1)
    torch.cuda.synchronize()
    assert out.op_status(4) == OK, out.op_text(4)
    np.testing.assert_array_equal(d.cpu().numpy(), first)
