"""Loads libfj_mi355x.so and drives the fj_cuda_* C-ABI from Python.

This is the same binding a ctypes consumer of the drop-in library would
write. There is no fallback: if the HIP library is missing the import of the
render entry points fails with an error naming the build step.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import subprocess
from dataclasses import dataclass

from . import abi

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libfj_mi355x.so")
_lib = None


def build(jobs: int = 8) -> str:
    """Compiles every HIP/C++ source of the library for gfx950."""
    subprocess.run(["make", "-s", "-C", PKG_DIR, f"-j{jobs}"], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() or make -C {PKG_DIR}")
        L = C.CDLL(LIB_PATH)
        eb = C.POINTER(abi.FjErrorBuffer)
        L.fj_cuda_create.restype = abi.FjStatus
        L.fj_cuda_create.argtypes = [C.POINTER(C.c_void_p), eb]
        L.fj_cuda_check_terminal.restype = abi.FjStatus
        L.fj_cuda_check_terminal.argtypes = [C.c_void_p, eb]
        L.fj_cuda_inspect.restype = abi.FjStatus
        L.fj_cuda_inspect.argtypes = [C.c_void_p, C.POINTER(abi.FjFrame), C.POINTER(abi.FjCudaContext), eb]
        L.fj_cuda_render.restype = abi.FjRenderOutcome
        L.fj_cuda_render.argtypes = [C.c_void_p, C.POINTER(abi.FjCudaContext), C.POINTER(abi.FjFrame),
                                     C.POINTER(abi.FjSubmission), C.POINTER(abi.FjPreparedHostData),
                                     abi.FjAbortCallback, eb]
        L.fj_cuda_retire_instance.restype = abi.FjStatus
        L.fj_cuda_retire_instance.argtypes = [C.c_void_p, C.c_uint64, eb]
        L.fj_cuda_shutdown.restype = abi.FjStatus
        L.fj_cuda_shutdown.argtypes = [C.c_void_p, eb]
        L.fjx_grain_apply.restype = abi.FjStatus
        L.fjx_grain_apply.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjPreparedHostData), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_uint64, eb]
        L.fjx_render_batch.restype = abi.FjStatus
        L.fjx_render_batch.argtypes = [C.c_void_p, C.POINTER(abi.FjCudaContext), C.c_int32, C.POINTER(abi.FjFrame),
                                       C.POINTER(abi.FjSubmission), C.POINTER(C.POINTER(abi.FjPreparedHostData)),
                                       C.c_int32, abi.FjAbortCallback, C.POINTER(abi.FjRenderOutcome),
                                       C.POINTER(abi.FjErrorBuffer)]
        L.fjx_diffusion_apply.restype = abi.FjStatus
        L.fjx_diffusion_apply.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjDiffusion), C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_int32, C.c_int32, C.c_uint64, C.POINTER(abi.FjErrorBuffer)]
        L.fjx_frame_descriptors.restype = abi.FjStatus
        L.fjx_frame_descriptors.argtypes = [C.c_uint32, C.POINTER(abi.FjFrame), C.POINTER(abi.FjFrameMathInput),
                                            C.POINTER(abi.FjNativeFrameDescriptors), C.POINTER(abi.FjErrorBuffer)]
        L.fj_cuda_call_retire_context.restype = C.c_uint32
        L.fj_cuda_call_retire_context.argtypes = [C.c_void_p, C.POINTER(abi.FjCallCompletionInput)]
        L.fj_cuda_call_trace_flags.restype = C.c_uint32
        L.fj_cuda_call_trace_flags.argtypes = []
        L.fj_cuda_call_trace_cutoff.restype = None
        L.fj_cuda_call_trace_cutoff.argtypes = [abi.FjStringView]
        L.fj_cuda_call_trace_failure.restype = None
        L.fj_cuda_call_trace_failure.argtypes = [abi.FjStringView]
        L.fjx_stage_timings.restype = C.c_int32
        L.fjx_stage_timings.argtypes = [C.c_void_p, C.c_int32, C.POINTER(FjxStageTiming), C.c_int32]
        L.fj_cuda_destroy.restype = abi.FjStatus
        L.fj_cuda_destroy.argtypes = [C.c_void_p, eb]
        L.fjx_highlight_boost_apply.restype = abi.FjStatus
        L.fjx_highlight_boost_apply.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjHighlightBoost), C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_int64, C.c_uint64, eb]
        L.fjx_build_tc_lut.restype = abi.FjStatus
        L.fjx_build_tc_lut.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjFilmTcLutInput), C.c_void_p, C.c_size_t,
                                       C.c_uint64, eb]
        L.fjx_glare_field.restype = abi.FjStatus
        L.fjx_glare_field.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjScannerPostEffects), C.c_uint64,
                                      C.c_uint32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_uint64, eb]
        L.fjx_capture_grain_input.restype = abi.FjStatus
        L.fjx_capture_grain_input.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                              eb]
        L.fjx_prepare_scanner_lut.restype = abi.FjStatus
        L.fjx_prepare_scanner_lut.argtypes = [C.c_void_p, C.POINTER(abi.FjScannerLutInput),
                                              C.POINTER(abi.FjScannerLutView), C.POINTER(abi.FjScannerLutFailure), eb]
        L.fjx_prepare_gaussian.restype = abi.FjStatus
        L.fjx_prepare_gaussian.argtypes = [C.c_void_p, C.POINTER(abi.FjGaussianKernelInput),
                                           C.POINTER(abi.FjGaussianKernelView), C.POINTER(abi.FjGaussianKernelFailure), eb]
        L.fjx_prepare_print_illuminant.restype = abi.FjStatus
        L.fjx_prepare_print_illuminant.argtypes = [C.POINTER(abi.FjPrintIlluminantParams), abi.FjFloatSpan,
                                                   C.POINTER(abi.FjIlluminant),
                                                   C.POINTER(abi.FjPrintIlluminantFailure), eb]
        L.fjx_prepare_print_preflash.restype = abi.FjStatus
        L.fjx_prepare_print_preflash.argtypes = [C.POINTER(abi.FjPrintPreflashInput), C.POINTER(abi.FjPrintPreflashRaw),
                                                 C.POINTER(abi.FjPrintPreflashFailure), eb]
        L.fjx_prepare_print_balance.restype = abi.FjStatus
        L.fjx_prepare_print_balance.argtypes = [C.POINTER(abi.FjPrintBalanceInput), C.POINTER(abi.FjPrintBalance),
                                                C.POINTER(abi.FjPrintBalanceFailure), eb]
        L.fjx_prepare_print_density.restype = abi.FjStatus
        L.fjx_prepare_print_density.argtypes = [C.POINTER(abi.FjPrintDensityInput), C.c_void_p, C.c_void_p,
                                                C.c_size_t, C.POINTER(C.c_uint64),
                                                C.POINTER(abi.FjPrintDensityFailure), eb]
        L.fjx_fit_saved_floats.restype = C.c_size_t
        L.fjx_fit_saved_floats.argtypes = [C.c_int32, C.c_int32]
        L.fjx_fit_forward.restype = abi.FjStatus
        L.fjx_fit_forward.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjPreparedHostData),
                                      C.POINTER(abi.FjFitParams), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_uint64, eb]
        L.fjx_fit_backward.restype = abi.FjStatus
        L.fjx_fit_backward.argtypes = [C.c_void_p, C.c_int32, C.POINTER(abi.FjPreparedHostData),
                                       C.POINTER(abi.FjFitParams), C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, eb]
        _lib = L
    return _lib


class FjxStageTiming(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("milliseconds", C.c_float)]


class FjError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"status={status.category} api={status.api} code={status.native_code}: {message}")
        self.status = status
        self.message = message


class ErrorBuffer:
    def __init__(self, capacity: int = 1024):
        self.buf = C.create_string_buffer(capacity)
        self.rec = abi.FjErrorBuffer(C.cast(self.buf, C.c_char_p), capacity, 0)

    def ptr(self):
        return C.byref(self.rec)

    @property
    def text(self) -> str:
        return self.buf.value.decode(errors="replace")


def _norm(v: int) -> int:
    return v if v else 1


def submission_for(prepared_record: abi.FjPreparedHostData, time_frames: float = 0.0,
                   instance_token: int = 1, submission_id: int = 1) -> abi.FjSubmission:
    """Identity digests the boundary checks (native/juicer_cuda_prepared.cpp:160-176)."""
    s = abi.FjSubmission()
    s.instance_token = instance_token
    s.frame_token = int(math.floor(time_frames)) & 0xFFFFFFFFFFFFFFFF
    s.submission_id = submission_id
    s.upload_core_hash = _norm(prepared_record.recipe_hash)
    s.dir_hash = _norm(prepared_record.dir_couplers.hash)
    s.scanner_hash = _norm(prepared_record.scanner_lut.hash)
    s.auto_exposure_hash = _norm(prepared_record.auto_exposure.hash)
    return s


def image(address: int, width: int, height: int, components: int, row_bytes: int | None = None,
          x1: int = 0, y1: int = 0) -> abi.FjImage:
    img = abi.FjImage()
    img.address = address
    img.bounds = abi.FjRect(x1, y1, x1 + width, y1 + height)
    img.row_bytes = row_bytes if row_bytes is not None else width * components * 4
    img.components = components
    img.depth = abi.FJ_DEPTH_FLOAT32
    return img


def frame_for(src_ptr: int, dst_ptr: int, width: int, height: int, components: int = 4,
              stream: int | None = None, pixel_size_um: float | None = None, time_frames: float = 0.0,
              window: tuple | None = None, session_seed: int = 0, clip_token: int = 0) -> abi.FjFrame:
    f = abi.FjFrame()
    f.source = image(src_ptr, width, height, components)
    f.destination = image(dst_ptr, width, height, components)
    wx1, wy1, wx2, wy2 = window if window else (0, 0, width, height)
    f.render_window = abi.FjRect(wx1, wy1, wx2, wy2)
    f.full_frame_extent = abi.FjRect(0, 0, width, height)
    if stream is not None:
        f.stream = stream
        f.flags = abi.FJ_FRAME_STREAM_PRESENT
    # Frame pixel pitch on a 35 mm long edge (rust film-juicer-core frame.rs:45-58).
    f.pixel_size_um = pixel_size_um if pixel_size_um is not None else 35000.0 / max(width, height)
    f.time_frames = time_frames
    f.frame_rate = 24.0
    f.session_seed = session_seed
    f.clip_token = clip_token
    return f


@dataclass
class Renderer:
    """One registered runtime owner (the library admits a single owner)."""

    handle: C.c_void_p

    @classmethod
    def create(cls) -> "Renderer":
        h = C.c_void_p()
        err = ErrorBuffer()
        st = lib().fj_cuda_create(C.byref(h), err.ptr())
        if st.category != abi.FJ_STATUS_SUCCESS:
            raise FjError(st, err.text)
        return cls(h)

    def inspect(self, frame: abi.FjFrame) -> abi.FjCudaContext:
        ctx = abi.FjCudaContext()
        err = ErrorBuffer()
        st = lib().fj_cuda_inspect(self.handle, C.byref(frame), C.byref(ctx), err.ptr())
        if st.category != abi.FJ_STATUS_SUCCESS:
            raise FjError(st, err.text)
        return ctx

    def render_raw(self, frame, prepared_record, submission=None, context=None, abort=None):
        """Returns (FjRenderOutcome, diagnostic text) without raising."""
        if context is None:
            context = self.inspect(frame)
        if submission is None:
            submission = submission_for(prepared_record, frame.time_frames)
        cb = abort if abort is not None else abi.FjAbortCallback(abi.FjAbortQuery(0), None)
        err = ErrorBuffer()
        out = lib().fj_cuda_render(self.handle, C.byref(context), C.byref(frame), C.byref(submission),
                                   C.byref(prepared_record), cb, err.ptr())
        return out, err.text

    def render(self, frame, prepared_record, **kw):
        out, text = self.render_raw(frame, prepared_record, **kw)
        if out.status.category != abi.FJ_STATUS_SUCCESS:
            raise FjError(out.status, text)
        return out

    def grain_apply(self, prepared_record, planes, width: int, height: int, stream: int = 0,
                    device_id: int = 0):
        """Runs the visual-grain stage on three device density planes
        (addresses C, M, Y); returns (FjStatus, diagnostic text)."""
        err = ErrorBuffer()
        st = lib().fjx_grain_apply(self.handle, device_id, C.byref(prepared_record), planes[0], planes[1], planes[2],
                                   width, height, stream, err.ptr())
        return st, err.text

    def render_batch(self, frames, records, streams: int = 4, context=None, submissions=None):
        """Renders frames[i] with records[i] in one call; returns (FjStatus,
        text, [FjRenderOutcome])."""
        n = len(frames)
        fr = (abi.FjFrame * n)(*frames)
        subs = submissions or [submission_for(r, f.time_frames) for f, r in zip(frames, records)]
        sb = (abi.FjSubmission * n)(*subs)
        ptrs = (C.POINTER(abi.FjPreparedHostData) * n)(*[C.pointer(r) for r in records])
        outs = (abi.FjRenderOutcome * n)()
        ctx = context if context is not None else (self.inspect(frames[0]) if frames else abi.FjCudaContext())
        err = ErrorBuffer()
        st = lib().fjx_render_batch(self.handle, C.byref(ctx), n, fr, sb, ptrs, streams,
                                    abi.FjAbortCallback(abi.FjAbortQuery(0), None), outs, err.ptr())
        return st, err.text, list(outs)

    def diffusion_apply(self, stage, planes, width: int, height: int, stream: int = 0, device_id: int = 0):
        """Runs one diffusion stage on three device planes (addresses R, G, B);
        returns (FjStatus, diagnostic text)."""
        err = ErrorBuffer()
        st = lib().fjx_diffusion_apply(self.handle, device_id, C.byref(stage), planes[0], planes[1], planes[2], width,
                                       height, stream, err.ptr())
        return st, err.text

    def stage_timings(self, device_id: int = 0) -> list:
        """(name, ms) of each stage of the last frame rendered with FJ_FRAME_TRACE_VERBOSE."""
        cap = 64
        buf = (FjxStageTiming * cap)()
        n = lib().fjx_stage_timings(self.handle, device_id, buf, cap)
        if n < 0:
            raise RuntimeError("stage timings unavailable")
        return [(buf[i].name.decode(), float(buf[i].milliseconds)) for i in range(min(n, cap))]

    def shutdown(self):
        err = ErrorBuffer()
        return lib().fj_cuda_shutdown(self.handle, err.ptr())

    def destroy(self):
        if self.handle:
            err = ErrorBuffer()
            st = lib().fj_cuda_destroy(self.handle, err.ptr())
            self.handle = None
            return st
        return None


_shared = None


def shared_renderer() -> Renderer:
    """Process-wide owner for tests and benchmarks."""
    global _shared
    if _shared is None:
        _shared = Renderer.create()
    return _shared
