// fj_api.cpp — the drop-in C-ABI (include/fj/juicer_cuda_api.h) over HIP.
//
// Boundary semantics follow the reference entry points:
//   owner lifecycle         native/juicer_cuda_api.cpp:245, :563-644
//   span validation         native/juicer_cuda_prepared.cpp:58-158
//   frame binding checks    native/juicer_cuda_prepared.cpp:160-222
//   descriptor tag checks   native/juicer_cuda_descriptors.cpp:522-600
//   render + completion     native/juicer_cuda_api.cpp:403-526
// Device work is issued on the caller's stream; host tables are packed into
// GPU-friendly layouts once and cached per device by content identity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fj/fj_mi355x_ext.h"
#include "fj_host.h"
#include "fj_owner.h"
#include "fj_plan.h"

using namespace fjx;

#define FJ_EXPORT __attribute__((visibility("default")))

namespace fjx {

std::atomic<FjCuda*> g_owner{nullptr};
thread_local bool t_in_call = false;

FjStatus make_status(uint32_t category, uint32_t api, int32_t code) {
    FjStatus s;
    s.category = category;
    s.api = api;
    s.native_code = api == FJ_API_NONE ? 0 : code;
    return s;
}

void write_error(FjErrorBuffer* error, const std::string& text) {
    if (!error) return;
    if (error->capacity == 0 || error->data == nullptr) {
        error->length = 0;
        return;
    }
    const size_t n = std::min(text.size(), error->capacity - 1);
    std::memcpy(error->data, text.data(), n);
    error->data[n] = '\0';
    error->length = n;
}

FjStatus fail(FjErrorBuffer* error, uint32_t category, const std::string& text, uint32_t api, int32_t code) {
    write_error(error, text);
    return make_status(category, api, code);
}

FjStatus hip_fail(FjErrorBuffer* error, hipError_t e, const char* where) {
    (void)hipGetLastError();  // do not leave a sticky error for the caller's next HIP call
    std::string text = std::string("HIP failure in ") + where + ": " + hipGetErrorString(e);
    const uint32_t category = (e == hipErrorOutOfMemory) ? FJ_STATUS_ALLOCATION_FAILURE : FJ_STATUS_CUDA_FAILURE;
    return fail(error, category, text, FJ_API_CUDA_RUNTIME, static_cast<int32_t>(e));
}

namespace {

bool image_device(const FjImage& img, int& device) {
    if (img.address == 0) return false;
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, reinterpret_cast<const void*>(img.address)) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) return false;
    device = attr.device;
    return true;
}

}  // namespace

FjStatus inspect_frame(const FjFrame* frame, FjCudaContext* out, FjErrorBuffer* error) {
    if (!frame) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "frame is null");
    int src_dev = -1, dst_dev = -1;
    if (!image_device(frame->source, src_dev) || !image_device(frame->destination, dst_dev))
        return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "frame images are not HIP device memory");
    if (src_dev != dst_dev)
        return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "source and destination live on different devices");
    if (out) {
        out->device_id = src_dev;
        out->context = static_cast<uintptr_t>(src_dev) + 1u;
    }
    return make_status(FJ_STATUS_SUCCESS);
}

FjRenderOutcome render_locked(FjCuda* cuda, const FjCudaContext* context, const FjFrame* frame,
                              const FjSubmission* submission, const FjPreparedHostData* prepared,
                              FjAbortCallback abort_callback, FjErrorBuffer* error) {
    FjRenderOutcome outcome{};
    if (!cuda->open) {
        outcome.status = fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime is shut down");
        return outcome;
    }
    if (!context || !frame || !submission || !prepared) {
        outcome.status = fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null render argument");
        return outcome;
    }
    FjCudaContext current{};
    FjStatus s = inspect_frame(frame, &current, error);
    if (s.category != FJ_STATUS_SUCCESS) {
        outcome.status = s;
        return outcome;
    }
    if (current.device_id != context->device_id || current.context != context->context) {
        outcome.status = fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "render context differs from the frame's device");
        return outcome;
    }
    hipError_t he = hipSetDevice(current.device_id);
    if (he != hipSuccess) {
        outcome.status = hip_fail(error, he, "hipSetDevice");
        return outcome;
    }
    auto& slot = cuda->devices[current.device_id];
    if (!slot) {
        slot = std::make_unique<DeviceState>();
        he = slot->init(current.device_id);
        if (he != hipSuccess) {
            slot.reset();
            outcome.status = hip_fail(error, he, "device state init");
            return outcome;
        }
    }
    outcome = render_frame(*slot, *frame, *submission, *prepared, abort_callback, error);
    if (outcome.status.category == FJ_STATUS_SUCCESS) write_error(error, "");
    return outcome;
}

bool retire_device_locked(FjCuda* cuda, int32_t device_id) {
    auto it = cuda->devices.find(device_id);
    if (it == cuda->devices.end()) return true;
    bool ok = true;
    if (it->second) ok = it->second->release() == hipSuccess;
    (void)hipGetLastError();
    cuda->devices.erase(it);
    return ok;
}

}  // namespace fjx

extern "C" {

FJ_EXPORT FjStatus fj_cuda_create(FjCuda** out_cuda, FjErrorBuffer* error) {
    try {
        if (!out_cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "out_cuda is null");
        *out_cuda = nullptr;
        auto owner = std::make_unique<FjCuda>();
        FjCuda* expected = nullptr;
        if (!g_owner.compare_exchange_strong(expected, owner.get()))
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "a film render runtime is already active");
        *out_cuda = owner.release();
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fj_cuda_create");
    }
}

FJ_EXPORT FjStatus fj_cuda_check_terminal(FjCuda* cuda, FjErrorBuffer* error) {
    if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
    if (t_in_call) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
    std::lock_guard<std::mutex> lock(cuda->mu);
    if (!cuda->open) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime is shut down");
    write_error(error, "");
    return make_status(FJ_STATUS_SUCCESS);
}

FJ_EXPORT FjStatus fj_cuda_inspect(FjCuda* cuda, const FjFrame* frame, FjCudaContext* out_context, FjErrorBuffer* error) {
    try {
        if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
        CallGuard guard;
        if (!guard.ok) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
        std::lock_guard<std::mutex> lock(cuda->mu);
        if (!cuda->open) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime is shut down");
        if (!out_context) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "out_context is null");
        FjCudaContext ctx{};
        const FjStatus s = inspect_frame(frame, &ctx, error);
        if (s.category == FJ_STATUS_SUCCESS) {
            *out_context = ctx;
            write_error(error, "");
        }
        return s;
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fj_cuda_inspect");
    }
}

FJ_EXPORT FjRenderOutcome fj_cuda_render(FjCuda* cuda, const FjCudaContext* context, const FjFrame* frame,
                               const FjSubmission* submission, const FjPreparedHostData* prepared,
                               FjAbortCallback abort_callback, FjErrorBuffer* error) {
    FjRenderOutcome outcome{};
    try {
        if (!cuda || g_owner.load() != cuda) {
            outcome.status = fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
            return outcome;
        }
        CallGuard guard;
        if (!guard.ok) {
            outcome.status = fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
            return outcome;
        }
        std::lock_guard<std::mutex> lock(cuda->mu);
        return render_locked(cuda, context, frame, submission, prepared, abort_callback, error);
    } catch (const std::bad_alloc&) {
        outcome.status = fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
        return outcome;
    } catch (...) {
        outcome.status = fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fj_cuda_render");
        return outcome;
    }
}

FJ_EXPORT FjStatus fj_cuda_retire_instance(FjCuda* cuda, uint64_t instance_token, FjErrorBuffer* error) {
    if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
    if (instance_token == 0) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "instance token is zero");
    CallGuard guard;
    if (!guard.ok) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
    std::lock_guard<std::mutex> lock(cuda->mu);
    // Static grain state is keyed by table content, not by instance: nothing
    // instance-owned remains after a render returns.
    write_error(error, "");
    return make_status(FJ_STATUS_SUCCESS);
}

FJ_EXPORT FjStatus fj_cuda_shutdown(FjCuda* cuda, FjErrorBuffer* error) {
    if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
    CallGuard guard;
    if (!guard.ok) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
    std::lock_guard<std::mutex> lock(cuda->mu);
    FjStatus result = make_status(FJ_STATUS_SUCCESS);
    for (auto& kv : cuda->devices) {
        if (!kv.second) continue;
        const hipError_t e = kv.second->release();
        if (e != hipSuccess && result.category == FJ_STATUS_SUCCESS) result = hip_fail(error, e, "shutdown");
    }
    cuda->devices.clear();
    cuda->open = false;
    if (result.category == FJ_STATUS_SUCCESS) write_error(error, "");
    return result;
}

FJ_EXPORT FjStatus fj_cuda_destroy(FjCuda* cuda, FjErrorBuffer* error) {
    if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
    if (t_in_call) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
    FjStatus s = make_status(FJ_STATUS_SUCCESS);
    {
        std::lock_guard<std::mutex> lock(cuda->mu);
        for (auto& kv : cuda->devices) {
            if (kv.second) {
                const hipError_t e = kv.second->release();
                if (e != hipSuccess && s.category == FJ_STATUS_SUCCESS) s = hip_fail(error, e, "destroy");
            }
        }
        cuda->devices.clear();
        cuda->open = false;
    }
    FjCuda* expected = cuda;
    g_owner.compare_exchange_strong(expected, nullptr);
    delete cuda;
    if (s.category == FJ_STATUS_SUCCESS) write_error(error, "");
    return s;
}

FJ_EXPORT FjStatus fjx_grain_apply(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* prepared,
                                   float* density_c, float* density_m, float* density_y, int32_t width, int32_t height,
                                   uint64_t stream, FjErrorBuffer* error) {
    try {
        if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
        CallGuard guard;
        if (!guard.ok) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
        std::lock_guard<std::mutex> lock(cuda->mu);
        if (!cuda->open) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime is shut down");
        if (!prepared) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null prepared record");
        int count = 0;
        hipError_t he = hipGetDeviceCount(&count);
        if (he != hipSuccess) return hip_fail(error, he, "hipGetDeviceCount");
        if (device_id < 0 || device_id >= count) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "device id out of range");
        he = hipSetDevice(device_id);
        if (he != hipSuccess) return hip_fail(error, he, "hipSetDevice");
        auto& slot = cuda->devices[device_id];
        if (!slot) {
            slot = std::make_unique<DeviceState>();
            he = slot->init(device_id);
            if (he != hipSuccess) {
                slot.reset();
                return hip_fail(error, he, "device state init");
            }
        }
        float* const planes[3] = {density_c, density_m, density_y};
        return grain_apply_planes(*slot, *prepared, planes, width, height, reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_grain_apply");
    }
}

FJ_EXPORT FjStatus fjx_diffusion_apply(FjCuda* cuda, int32_t device_id, const FjDiffusion* stage, float* red,
                                       float* green, float* blue, int32_t width, int32_t height, uint64_t stream,
                                       FjErrorBuffer* error) {
    try {
        if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
        CallGuard guard;
        if (!guard.ok) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
        std::lock_guard<std::mutex> lock(cuda->mu);
        if (!cuda->open) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime is shut down");
        if (!stage) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null diffusion stage");
        int count = 0;
        hipError_t he = hipGetDeviceCount(&count);
        if (he != hipSuccess) return hip_fail(error, he, "hipGetDeviceCount");
        if (device_id < 0 || device_id >= count) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "device id out of range");
        he = hipSetDevice(device_id);
        if (he != hipSuccess) return hip_fail(error, he, "hipSetDevice");
        auto& slot = cuda->devices[device_id];
        if (!slot) {
            slot = std::make_unique<DeviceState>();
            he = slot->init(device_id);
            if (he != hipSuccess) {
                slot.reset();
                return hip_fail(error, he, "device state init");
            }
        }
        float* const planes[3] = {red, green, blue};
        return diffusion_apply_planes(*slot, *stage, planes, width, height, reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_diffusion_apply");
    }
}

FJ_EXPORT FjStatus fjx_render_batch(FjCuda* cuda, const FjCudaContext* context, int32_t count, const FjFrame* frames,
                                    const FjSubmission* submissions, const FjPreparedHostData* const* prepared,
                                    int32_t streams, FjAbortCallback abort_callback, FjRenderOutcome* outcomes,
                                    FjErrorBuffer* error) {
    try {
        if (!cuda || g_owner.load() != cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "unknown runtime handle");
        CallGuard guard;
        if (!guard.ok) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
        std::lock_guard<std::mutex> lock(cuda->mu);
        if (!cuda->open) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "runtime is shut down");
        if (!context || count < 0 || (count > 0 && (!frames || !submissions || !prepared || !outcomes)))
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null batch argument");
        if (count == 0) {
            write_error(error, "");
            return make_status(FJ_STATUS_SUCCESS);
        }
        // One device and one caller stream for the whole batch.
        const bool has_stream = (frames[0].flags & FJ_FRAME_STREAM_PRESENT) != 0;
        for (int32_t i = 0; i < count; ++i) {
            FjCudaContext c{};
            const FjStatus st = inspect_frame(&frames[i], &c, error);
            if (st.category != FJ_STATUS_SUCCESS) return st;
            if (c.device_id != context->device_id || c.context != context->context)
                return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "batch frame on another device");
            if (!prepared[i]) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null prepared record in batch");
            const bool hs = (frames[i].flags & FJ_FRAME_STREAM_PRESENT) != 0;
            if (hs != has_stream || (hs && frames[i].stream != frames[0].stream))
                return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "batch frames name different streams");
        }
        hipError_t he = hipSetDevice(context->device_id);
        if (he != hipSuccess) return hip_fail(error, he, "hipSetDevice");
        auto& slot = cuda->devices[context->device_id];
        if (!slot) {
            slot = std::make_unique<DeviceState>();
            he = slot->init(context->device_id);
            if (he != hipSuccess) {
                slot.reset();
                return hip_fail(error, he, "device state init");
            }
        }
        hipStream_t caller = has_stream ? reinterpret_cast<hipStream_t>(frames[0].stream) : nullptr;
        const FjStatus st =
            render_batch(*slot, count, frames, submissions, prepared, streams, caller, abort_callback, outcomes, error);
        if (!has_stream) {
            he = hipStreamSynchronize(caller);
            if (he != hipSuccess && st.category == FJ_STATUS_SUCCESS) return hip_fail(error, he, "batch synchronize");
        }
        return st;
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_render_batch");
    }
}

FJ_EXPORT int32_t fjx_stage_timings(FjCuda* cuda, int32_t device_id, FjxStageTiming* out, int32_t capacity) {
    try {
        if (!cuda || g_owner.load() != cuda || capacity < 0 || (capacity > 0 && !out)) return -1;
        std::lock_guard<std::mutex> lock(cuda->mu);
        auto it = cuda->devices.find(device_id);
        if (it == cuda->devices.end() || !it->second) return -1;
        DeviceState& dev = *it->second;
        const int n = dev.stage_count;
        if (n <= 0) return 0;
        if (hipEventSynchronize(dev.stage_events[n]) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
        for (int i = 0; i < n && i < capacity; ++i) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, dev.stage_events[i], dev.stage_events[i + 1]) != hipSuccess) {
                (void)hipGetLastError();
                return -1;
            }
            std::snprintf(out[i].name, sizeof(out[i].name), "%s", dev.stage_names[i] ? dev.stage_names[i] : "");
            out[i].milliseconds = ms;
        }
        return n;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
