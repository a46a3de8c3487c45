"""The C-ABI library loads, exports every declared entry point and enforces
its owner/argument contract without a GPU (no compute is issued here)."""
import ctypes as C
import os
import subprocess
import re

import pytest

from fj import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# Headers whose functions the host defines and the library resolves at run
# time (it must not define them itself).
IMPORT_HEADERS = ("juicer_diffusion_api.h", "juicer_dir_boundary_api.h")


def declared_functions(headers=None):
    names = []
    for header in headers or [h for h in os.listdir(os.path.join(ROOT, "include", "fj")) if h not in IMPORT_HEADERS]:
        text = open(os.path.join(ROOT, "include", "fj", header)).read()
        names += re.findall(r"^\s*(?:FjStatus|FjRenderOutcome|uint32_t|void|FjCallCompletionResult)\s+(fj_\w+)\s*\(",
                            text, re.M)
    return sorted(set(names))


def test_exports_every_declared_symbol():
    lib = runtime.lib()
    names = declared_functions()
    assert "fj_cuda_render" in names and len(names) >= 7
    for n in names:
        assert hasattr(lib, n), n


def test_single_owner_lifecycle():
    r = runtime.Renderer.create()
    try:
        with pytest.raises(runtime.FjError) as e:
            runtime.Renderer.create()
        # A duplicate owner is a preparation failure (native/juicer_cuda_api.cpp:262-269).
        assert e.value.status.category == abi.FJ_STATUS_PREPARATION_FAILURE
        err = runtime.ErrorBuffer()
        assert runtime.lib().fj_cuda_check_terminal(r.handle, err.ptr()).category == abi.FJ_STATUS_SUCCESS
        st = r.shutdown()
        assert st.category == abi.FJ_STATUS_SUCCESS
        assert r.shutdown().category == abi.FJ_STATUS_SUCCESS  # Closed: shutdown is idempotent
        # The terminal precheck checks registration and re-entry only.
        assert runtime.lib().fj_cuda_check_terminal(r.handle, err.ptr()).category == abi.FJ_STATUS_SUCCESS
        st = runtime.lib().fj_cuda_retire_instance(r.handle, 5, err.ptr())
        assert st.category == abi.FJ_STATUS_PREPARATION_FAILURE and "admission is closed" in err.text
    finally:
        assert r.destroy().category == abi.FJ_STATUS_SUCCESS
    r2 = runtime.Renderer.create()
    assert r2.destroy().category == abi.FJ_STATUS_SUCCESS


def test_null_arguments_are_rejected():
    lib = runtime.lib()
    err = runtime.ErrorBuffer(8)  # truncation keeps the status
    st = lib.fj_cuda_create(None, err.ptr())
    assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT and st.api == abi.FJ_API_NONE and st.native_code == 0
    assert err.rec.length <= 7
    r = runtime.Renderer.create()
    try:
        ctx = abi.FjCudaContext()
        st = lib.fj_cuda_inspect(r.handle, None, C.byref(ctx), None)
        assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
        out = lib.fj_cuda_render(r.handle, None, None, None, None,
                                 abi.FjAbortCallback(abi.FjAbortQuery(0), None), None)
        assert out.status.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
        st = lib.fj_cuda_retire_instance(r.handle, 0, None)
        assert st.category == abi.FJ_STATUS_UNSUPPORTED_INPUT
    finally:
        r.destroy()


def test_unknown_handle_rejected():
    lib = runtime.lib()
    bogus = C.c_void_p(0x1234)
    assert lib.fj_cuda_check_terminal(bogus, None).category == abi.FJ_STATUS_PREPARATION_FAILURE
    assert lib.fj_cuda_check_terminal(None, None).category == abi.FJ_STATUS_UNSUPPORTED_INPUT


def test_host_imports_are_resolved_not_defined():
    """The optional host functions are looked up at run time: the library
    neither defines nor links against them."""
    names = declared_functions(IMPORT_HEADERS)
    assert sorted(names) == ["fj_diffusion_components_prepare", "fj_dir_boundary_acquire", "fj_dir_boundary_describe",
                             "fj_dir_boundary_release", "fj_dir_boundary_view"]
    out = subprocess.run(["nm", "-D", runtime.LIB_PATH], capture_output=True, text=True, check=True).stdout
    for n in names:
        assert n not in out, n
    assert "dlsym" in out
