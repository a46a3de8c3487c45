"""The call scope (juicer_cuda_call_scope.h) without a GPU: argument
validation of fj_cuda_call, the precheck/admit/release/complete protocol,
the owner lock and re-entry guard held by an admitted scope, completion
forwarding, the auto-exposure preview descriptor and the trace sink
(native/juicer_cuda_call_scope.cpp:275-733). Scripts run inside
fj_cuda_call through the C driver in tests/native (no compute is issued).
"""
import ctypes as C

import pytest

import call_harness as H
from fj import abi, runtime

OK = abi.FJ_STATUS_SUCCESS
BAD = abi.FJ_STATUS_UNSUPPORTED_INPUT


@pytest.fixture()
def owner():
    r = runtime.Renderer.create()
    yield r
    assert r.destroy().category == OK


def test_call_arguments_are_validated(owner):
    lib = H.lib()
    for args in ((None, 1, 1, 1, 1), (owner.handle, 4, 1, 1, 1), (owner.handle, 1, 0, 1, 1),
                 (owner.handle, 1, 1, 0, 1), (owner.handle, 1, 1, 1, 0)):
        r = lib.fjh_call_raw(*args)
        assert r.fatal == 1 and r.outcome.status.category == BAD, args


def test_protocol_in_order(owner):
    s = H.run(owner.handle, [H.PRECHECK, H.ADMIT, H.RELEASE, H.COMPLETE])
    assert [s.op_status(i) for i in (0, 1, 3)] == [OK, OK, OK]
    assert s.completion_calls == 1 and s.call_result.fatal == 0
    assert s.completion_seen.render_outcome_valid == 0 and s.completion_seen.has_inspected_context == 0
    # The owner is free again after the call.
    assert runtime.lib().fj_cuda_check_terminal(owner.handle, None).category == OK


def test_admitted_scope_holds_the_reentry_guard(owner):
    s = H.run(owner.handle, [H.ADMIT, H.PUBLIC_INSPECT, H.RELEASE, H.COMPLETE], frame=abi.FjFrame())
    assert s.op_status(1) == BAD and "re-entered" in s.op_text(1)
    assert s.op_status(3) == OK and s.completion_calls == 1


def test_misuse_is_refused(owner):
    s = H.run(owner.handle, [H.RENDER, H.INSPECT, H.DESCRIPTORS])
    for i in range(3):
        assert s.op_status(i) == BAD and s.op_text(i) == "CUDA call scope is not admitted", i
    s = H.run(owner.handle, [H.ADMIT, H.ADMIT, H.PRECHECK])
    assert s.op_status(0) == OK
    assert s.op_status(1) == BAD and s.op_text(1) == "invalid CUDA call scope"
    assert s.op_status(2) == BAD and s.op_text(2) == "invalid CUDA call scope"
    # Descriptors need an inspected frame.
    s = H.run(owner.handle, [H.ADMIT, H.DESCRIPTORS, H.RELEASE], math=abi.FjFrameMathInput())
    assert s.op_status(1) == BAD and s.op_text(1) == "CUDA call scope is not admitted"


def test_complete_runs_once_and_only_after_release(owner):
    s = H.run(owner.handle, [H.ADMIT, H.COMPLETE, H.RELEASE])
    assert s.op_status(1) == abi.FJ_STATUS_INTERNAL_FAILURE and s.completed.fatal == 1
    assert s.completion_calls == 0
    s = H.run(owner.handle, [H.PRECHECK, H.ADMIT, H.RELEASE, H.COMPLETE, H.COMPLETE], force_fatal=1)
    assert s.completion_calls == 1
    assert s.op_status(4) == abi.FJ_STATUS_INTERNAL_FAILURE
    assert s.call_result.fatal == 1


def test_scope_left_admitted_is_released_on_return(owner):
    H.run(owner.handle, [H.ADMIT])
    assert runtime.lib().fj_cuda_check_terminal(owner.handle, None).category == OK
    s = H.run(owner.handle, [H.PRECHECK, H.ADMIT, H.RELEASE, H.COMPLETE])
    assert s.op_status(1) == OK


def test_shut_down_owner_is_refused(owner):
    assert owner.shutdown().category == OK
    s = H.run(owner.handle, [H.PRECHECK, H.ADMIT])
    closed = abi.FJ_STATUS_PREPARATION_FAILURE
    assert s.op_status(0) == closed and "shut down" in s.op_text(0)
    assert s.op_status(1) == closed and "shut down" in s.op_text(1)


def test_failure_diagnostic_reaches_the_completion(owner):
    s = H.run(owner.handle, [H.ADMIT, H.DESCRIPTORS, H.RELEASE, H.COMPLETE], math=abi.FjFrameMathInput())
    assert s.completion_calls == 1
    assert s.completion_diagnostic.decode() == "CUDA call scope is not admitted"


def preview(owner, source, meter, method=0):
    s = H.run(owner.handle, [H.PREVIEW], preview_source=abi.FjRect(*source), preview_meter=abi.FjRect(*meter),
              preview_method=method)
    return s.op_status(0), s.op_text(0), s.preview


@pytest.mark.parametrize("meter,dims", [
    ((0, 0, 3840, 2160), (256, 144)),     # (2160*256 + 1920) // 3840 = 144
    ((0, 0, 2160, 3840), (144, 256)),
    ((10, 20, 210, 120), (200, 100)),     # within 256: unchanged
    ((0, 0, 1000, 1), (256, 1)),          # never below one sample
    ((0, 0, 0, 0), (0, 0)),
    ((0, 0, 4096, 1716), (256, 107)),     # (1716*256 + 2048) // 4096 = 107
])
def test_preview_dimensions(owner, meter, dims):
    st, text, p = preview(owner, (0, 0, 3840, 2160), meter)
    assert st == OK, text
    assert (p.preview_width, p.preview_height) == dims
    assert (p.meter_bounds.x2, p.source_bounds.x2, p.method) == (meter[2], 3840, 0)
    assert p.hash != 0


def test_preview_identity_and_validation(owner):
    hashes = {preview(owner, (0, 0, 3840, 2160), (0, 0, 3840, 2160), m)[2].hash for m in range(7)}
    assert len(hashes) == 7
    assert preview(owner, (0, 0, 64, 64), (0, 0, 64, 64))[2].hash == preview(owner, (0, 0, 64, 64), (0, 0, 64, 64))[2].hash
    for source, meter, method in (((0, 0, 8, 8), (0, 0, 8, 8), 7), ((8, 0, 0, 8), (0, 0, 8, 8), 0),
                                  ((0, 0, 8, 8), (0, 8, 8, 0), 0)):
        st, text, p = preview(owner, source, meter, method)
        assert st == BAD and text == "invalid auto-exposure preview arguments"
        assert bytes(p) == bytes(abi.FjAutoExposure())


def test_trace_sink():
    lib = runtime.lib()
    flags = lib.fj_cuda_call_trace_flags()
    assert flags in (0, abi.FJ_FRAME_TRACE_INFO, abi.FJ_FRAME_TRACE_INFO | abi.FJ_FRAME_TRACE_VERBOSE)
    lib.fj_cuda_call_trace_cutoff(abi.FjStringView(b"cutoff", 6))
    lib.fj_cuda_call_trace_failure(abi.FjStringView(None, 0))


def test_retire_context_outside_completion_is_a_no_op():
    assert runtime.lib().fj_cuda_call_retire_context(None, None) == 0


def test_context_loss_retirement_during_completion(owner):
    ctx = abi.FjCudaContext(0, 1)
    ops = [H.ADMIT, H.RELEASE, H.COMPLETE]
    s = H.run(owner.handle, ops, context=ctx, report_context_loss=1, recovery_scope_entered=1,
              retire_on_completion=1)
    assert s.retirement == 3  # attempted and accepted (no device state yet)
    assert s.completion_seen.outcome.status.category == abi.FJ_STATUS_CONTEXT_LOSS
    assert runtime.lib().fj_cuda_check_terminal(owner.handle, None).category == OK  # lock dropped
    s = H.run(owner.handle, ops, context=ctx, report_context_loss=1, recovery_scope_entered=0,
              retire_on_completion=1)
    assert s.retirement == 0  # no recovery scope
    s = H.run(owner.handle, ops, context=ctx, report_context_loss=0, recovery_scope_entered=1,
              retire_on_completion=1)
    assert s.retirement == 0  # nothing to recover
