"""The runtime owner's lifecycle without a GPU (native/juicer_cuda_api.cpp:
24-120, :528-644): re-entrant terminal calls from inside an admitted call
scope, the Retained quarantine a re-entrant destroy leaves, and a destroy
racing other threads' calls on the same handle.

Re-entrant destroy retains the owner for the rest of the process (the
reference's quarantine), so those cases run in a child interpreter.
"""
import ctypes as C
import os
import subprocess
import sys
import threading

import pytest

from fj import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r'''
import sys
sys.path[:0] = [{pkg!r}, {tests!r}]
import call_harness as H
from fj import abi, runtime
r = runtime.Renderer.create()
s = H.run(r.handle, [H.ADMIT, H.{op}, H.RELEASE, H.COMPLETE])
print(s.op_status(1), s.op_text(1))
lib = runtime.lib()
err = runtime.ErrorBuffer()
print(lib.fj_cuda_retire_instance(r.handle, 3, err.ptr()).category, err.text)
print(lib.fj_cuda_shutdown(r.handle, err.ptr()).category, err.text)
print(lib.fj_cuda_destroy(r.handle, err.ptr()).category, err.text)
try:
    runtime.Renderer.create()
    print("created")
except runtime.FjError as e:
    print("refused", e.status.category)
'''


def run_child(op):
    code = CHILD.format(pkg=os.path.join(ROOT, "film-juicer_amd"), tests=os.path.join(ROOT, "tests"), op=op)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return [line.split(" ", 1) for line in out.stdout.strip().splitlines()]


def test_reentrant_destroy_retains_the_owner():
    lines = run_child("PUBLIC_DESTROY")
    # The destroy inside the scope is refused as re-entry ...
    assert int(lines[0][0]) == abi.FJ_STATUS_UNSUPPORTED_INPUT and "re-entered" in lines[0][1]
    # ... and the owner stays registered, closed to calls and to terminal retries.
    assert int(lines[1][0]) == abi.FJ_STATUS_PREPARATION_FAILURE and "admission is closed" in lines[1][1]
    assert int(lines[2][0]) == abi.FJ_STATUS_UNSUPPORTED_INPUT  # the published terminal status
    assert int(lines[3][0]) == abi.FJ_STATUS_UNSUPPORTED_INPUT
    assert lines[4] == ["refused", str(abi.FJ_STATUS_PREPARATION_FAILURE)]


def test_reentrant_shutdown_is_refused_and_harmless():
    lines = run_child("PUBLIC_SHUTDOWN")
    assert int(lines[0][0]) == abi.FJ_STATUS_UNSUPPORTED_INPUT and "re-entered" in lines[0][1]
    assert int(lines[1][0]) == abi.FJ_STATUS_SUCCESS  # still accepting after the refused attempt
    assert int(lines[2][0]) == abi.FJ_STATUS_SUCCESS
    assert int(lines[3][0]) == abi.FJ_STATUS_SUCCESS
    assert lines[4] == ["created"]


def test_destroy_while_other_threads_call():
    """Threads keep calling a cheap admitted entry point on the handle while
    the main thread destroys it: every call ends with success (admitted
    before the close) or a refusal, and none touches the freed owner."""
    lib = runtime.lib()
    for _ in range(20):
        r = runtime.Renderer.create()
        handle = r.handle
        stop = threading.Event()
        seen = set()

        def worker():
            while not stop.is_set():
                st = lib.fj_cuda_retire_instance(handle, 9, None)
                seen.add(st.category)

        threads = [threading.Thread(target=worker) for _ in range(4)]
        for t in threads:
            t.start()
        assert r.destroy().category == abi.FJ_STATUS_SUCCESS
        stop.set()
        for t in threads:
            t.join()
        assert seen <= {abi.FJ_STATUS_SUCCESS, abi.FJ_STATUS_PREPARATION_FAILURE}, seen
    assert lib.fj_cuda_retire_instance(handle, 9, None).category == abi.FJ_STATUS_PREPARATION_FAILURE


def test_bad_error_buffer_is_refused_without_writing():
    lib = runtime.lib()
    bad = abi.FjErrorBuffer(None, 16, 99)
    assert lib.fj_cuda_create(None, C.byref(bad)).category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    r = runtime.Renderer.create()
    try:
        st = lib.fj_cuda_check_terminal(r.handle, C.byref(bad))
        assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    finally:
        assert r.destroy().category == abi.FJ_STATUS_SUCCESS


@pytest.mark.parametrize("field,value", [("device_id", -1), ("context", 0)])
def test_render_arguments_checked_before_admission(field, value):
    """invalid_render_arguments (native/juicer_cuda_api.cpp:141-148)."""
    lib = runtime.lib()
    r = runtime.Renderer.create()
    try:
        ctx = abi.FjCudaContext(0, 1)
        setattr(ctx, field, value)
        err = runtime.ErrorBuffer()
        out = lib.fj_cuda_render(r.handle, C.byref(ctx), C.byref(abi.FjFrame()), C.byref(abi.FjSubmission(1, 1)),
                                 C.byref(abi.FjPreparedHostData()), abi.FjAbortCallback(abi.FjAbortQuery(0), None),
                                 err.ptr())
        assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and "invalid CUDA render arguments" in err.text
    finally:
        assert r.destroy().category == abi.FJ_STATUS_SUCCESS
