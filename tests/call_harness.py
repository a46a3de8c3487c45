"""ctypes view of tests/native/fj_call_harness.c: scripts of call-scope
operations played inside fj_cuda_call (test infrastructure)."""
import ctypes as C
import os

from fj import abi, runtime

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "native", "libfj_call_harness.so")

PRECHECK, ADMIT, INSPECT, DESCRIPTORS, PREVIEW, RENDER, RELEASE, COMPLETE = 1, 2, 3, 4, 5, 6, 7, 8
PUBLIC_INSPECT, TRACE, RENDER_OTHER_CONTEXT, PUBLIC_DESTROY, PUBLIC_SHUTDOWN = 9, 10, 11, 12, 13
MAX_OPS, TEXT = 16, 256
# Host preparation behaviours of the driver's callbacks (fj_call_harness.c).
HOST_NONE, HOST_DELEGATE, HOST_GAUSSIAN_FAIL, HOST_GAUSSIAN_HALF = 0, 1, 2, 3
HOST_LUT_FAIL, HOST_GAUSSIAN_BROKEN, HOST_LUT_BROKEN = 4, 5, 6


class Script(C.Structure):
    _fields_ = [
        ("route", C.c_uint32), ("op_count", C.c_uint32), ("ops", C.c_uint32 * MAX_OPS),
        ("frame", C.POINTER(abi.FjFrame)), ("math", C.POINTER(abi.FjFrameMathInput)),
        ("prepared", C.POINTER(abi.FjPreparedHostData)), ("copy_descriptors", C.c_uint32),
        ("submission", C.POINTER(abi.FjSubmission)), ("preview_source", abi.FjRect),
        ("preview_meter", abi.FjRect), ("preview_method", C.c_uint32), ("recovery_scope_entered", C.c_uint32),
        ("force_fatal", C.c_uint32), ("retire_on_completion", C.c_uint32), ("report_context_loss", C.c_uint32),
        ("host_mode", C.c_uint32),
        ("status", abi.FjStatus * MAX_OPS), ("text", (C.c_char * TEXT) * MAX_OPS),
        ("context", abi.FjCudaContext), ("descriptors", abi.FjNativeFrameDescriptors),
        ("preview", abi.FjAutoExposure), ("outcome", abi.FjRenderOutcome),
        ("completed", abi.FjCallCompletionResult), ("completion_calls", C.c_uint32),
        ("completion_seen", abi.FjCallCompletionInput), ("completion_diagnostic", C.c_char * TEXT),
        ("retirement", C.c_uint32), ("call_result", abi.FjCallCompletionResult),
        ("host_calls", C.c_uint32 * 2)]

    def op_status(self, i):
        return self.status[i].category

    def op_text(self, i):
        return self.text[i].value.decode()


_lib = None


def lib():
    global _lib
    if _lib is None:
        runtime.lib()  # the render library first: the driver binds to the same copy
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"call-scope test driver not built: {LIB_PATH} (run __graft_entry__.build())")
        _lib = C.CDLL(LIB_PATH)
        _lib.fjh_run.restype = C.c_int
        _lib.fjh_run.argtypes = [C.c_void_p, C.POINTER(Script)]
        _lib.fjh_call_raw.restype = abi.FjCallCompletionResult
        _lib.fjh_call_raw.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_int]
        _lib.fjh_script_size.restype = C.c_size_t
        assert _lib.fjh_script_size() == C.sizeof(Script), "driver script layout"
    return _lib


def run(handle, ops, route=abi.FJ_ROUTE_NEGATIVE_PRINT, **fields):
    """Plays `ops` in one fj_cuda_call; keyword fields set script inputs
    (frame/math/prepared/submission as ctypes records, passed by reference)."""
    s = Script()
    s.route = route
    s.op_count = len(ops)
    for i, op in enumerate(ops):
        s.ops[i] = op
    keep = []
    for k, v in fields.items():
        if k in ("frame", "math", "prepared", "submission"):
            keep.append(v)
            setattr(s, k, C.pointer(v))
        else:
            setattr(s, k, v)
    lib().fjh_run(handle, C.byref(s))
    s._keep = keep
    return s
