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
#include <condition_variable>
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

namespace {
// Registration: publishing and clearing the owner, and counting entrants.
std::mutex g_registry;
std::condition_variable g_registry_cv;
}  // namespace

OwnerEntry::OwnerEntry(FjCuda* cuda) {
    if (!cuda) return;
    std::lock_guard<std::mutex> lock(g_registry);
    if (g_owner.load(std::memory_order_acquire) != cuda) return;
    ++cuda->entrants;
    cuda_ = cuda;
}

OwnerEntry::~OwnerEntry() {
    if (!cuda_) return;
    std::lock_guard<std::mutex> lock(g_registry);
    if (--cuda_->entrants == 0) g_registry_cv.notify_all();
}

NativeCall::NativeCall(FjCuda* cuda) : entry(cuda) {
    if (!cuda || !guard.ok) {
        status = make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        text = cuda ? "runtime re-entered from its own call" : "invalid CUDA call";
        return;
    }
    if (!entry.ok()) {
        status = make_status(FJ_STATUS_PREPARATION_FAILURE);
        text = "CUDA call requires the registered owner (unknown runtime handle)";
        return;
    }
    lock = std::unique_lock<std::mutex>(cuda->mu);
    if (!cuda->accepting()) {
        lock.unlock();
        status = make_status(FJ_STATUS_PREPARATION_FAILURE);
        text = "CUDA owner admission is closed (runtime is shut down)";
        return;
    }
    status = make_status(FJ_STATUS_SUCCESS);
}

FjStatus owner_shutdown(FjCuda* cuda) {
    Lifecycle state = cuda->lifecycle.load(std::memory_order_acquire);
    for (;;) {
        if (state == Lifecycle::Closed) return make_status(FJ_STATUS_SUCCESS);
        if (state == Lifecycle::Blocked || state == Lifecycle::Retained) return cuda->terminal;
        if (state == Lifecycle::Closing) {
            cuda->lifecycle.wait(state, std::memory_order_acquire);
        } else if (cuda->lifecycle.compare_exchange_weak(state, Lifecycle::Closing, std::memory_order_acq_rel)) {
            break;
        }
        state = cuda->lifecycle.load(std::memory_order_acquire);
    }
    auto publish = [&](Lifecycle next, FjStatus st, const std::string& text) {
        if (next != Lifecycle::Closed) {
            cuda->terminal = st;
            cuda->terminal_text = text;
        }
        cuda->lifecycle.store(next, std::memory_order_release);
        cuda->lifecycle.notify_all();
        return st;
    };
    try {
        // Closing refuses new admissions; the call in flight (if any) ends first.
        std::lock_guard<std::mutex> lock(cuda->mu);
        hipError_t first = hipSuccess;
        for (auto& kv : cuda->devices) {
            if (!kv.second) continue;
            const hipError_t e = kv.second->release();
            if (e != hipSuccess && first == hipSuccess) first = e;
        }
        cuda->devices.clear();
        if (first != hipSuccess) {
            (void)hipGetLastError();
            return publish(Lifecycle::Blocked, make_status(FJ_STATUS_CUDA_FAILURE, FJ_API_CUDA_RUNTIME, first),
                           std::string("HIP failure in shutdown: ") + hipGetErrorString(first) +
                               "; device state released with errors");
        }
        return publish(Lifecycle::Closed, make_status(FJ_STATUS_SUCCESS), "");
    } catch (const std::bad_alloc&) {
        return publish(Lifecycle::Blocked, make_status(FJ_STATUS_ALLOCATION_FAILURE), "CUDA shutdown allocation failed");
    } catch (...) {
        return publish(Lifecycle::Blocked, make_status(FJ_STATUS_INTERNAL_FAILURE),
                       "CUDA shutdown failed with unknown exception");
    }
}

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
                              FjAbortCallback abort_callback, FjErrorBuffer* error, const FjHostPreparation* host) {
    FjRenderOutcome outcome{};
    if (!cuda->accepting()) {
        outcome.status = fail(error, FJ_STATUS_PREPARATION_FAILURE, "CUDA owner admission is closed (runtime is shut down)");
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
    outcome = render_frame(*slot, *frame, *submission, *prepared, abort_callback, error, host);
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

namespace {

// invalid_render_arguments (native/juicer_cuda_api.cpp:141-148).
bool invalid_render_arguments(const FjCudaContext* context, const FjFrame* frame, const FjSubmission* submission,
                              const FjPreparedHostData* prepared) {
    return !context || !frame || !submission || !prepared || context->device_id < 0 || context->context == 0 ||
           submission->instance_token == 0 || submission->submission_id == 0 || !std::isfinite(frame->time_frames) ||
           frame->time_frames < -0x1p63 || frame->time_frames >= 0x1p63 || !std::isfinite(frame->frame_rate) ||
           frame->frame_rate < 0 || !std::isfinite(frame->pixel_size_um) || frame->pixel_size_um < 0;
}

// A terminal call that cannot complete keeps the owner registered and
// blocked (terminal_admission_failure, native/juicer_cuda_api.cpp:163-177).
FjStatus terminal_failure(FjCuda* cuda, FjStatus result, const std::string& text, bool consuming, FjErrorBuffer* error) {
    if (cuda != g_owner.load(std::memory_order_acquire)) return fail(error, result.category, text);
    Lifecycle state = cuda->lifecycle.load(std::memory_order_acquire);
    for (;;) {
        if (state == Lifecycle::Blocked || state == Lifecycle::Retained) break;
        if (state == Lifecycle::Closing) {
            cuda->lifecycle.wait(state, std::memory_order_acquire);
            state = cuda->lifecycle.load(std::memory_order_acquire);
            continue;
        }
        if (cuda->lifecycle.compare_exchange_weak(state, Lifecycle::Closing, std::memory_order_acq_rel)) {
            cuda->terminal = result;
            cuda->terminal_text = text;
            cuda->lifecycle.store(Lifecycle::Blocked, std::memory_order_release);
            cuda->lifecycle.notify_all();
            break;
        }
    }
    if (consuming) cuda->lifecycle.store(Lifecycle::Retained, std::memory_order_release);
    write_error(error, cuda->terminal_text);
    return cuda->terminal;
}

DeviceState* device_state(FjCuda* cuda, int device_id, FjStatus& st, FjErrorBuffer* error) {
    int count = 0;
    hipError_t he = hipGetDeviceCount(&count);
    if (he != hipSuccess) {
        st = hip_fail(error, he, "hipGetDeviceCount");
        return nullptr;
    }
    if (device_id < 0 || device_id >= count) {
        st = fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "device id out of range");
        return nullptr;
    }
    he = hipSetDevice(device_id);
    if (he != hipSuccess) {
        st = hip_fail(error, he, "hipSetDevice");
        return nullptr;
    }
    auto& slot = cuda->devices[device_id];
    if (!slot) {
        slot = std::make_unique<DeviceState>();
        he = slot->init(device_id);
        if (he != hipSuccess) {
            slot.reset();
            st = hip_fail(error, he, "device state init");
            return nullptr;
        }
    }
    st = make_status(FJ_STATUS_SUCCESS);
    return slot.get();
}

}  // namespace

extern "C" {

FJ_EXPORT FjStatus fj_cuda_create(FjCuda** out_cuda, FjErrorBuffer* error) {
    if (out_cuda) *out_cuda = nullptr;
    if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
    try {
        if (t_in_call || !out_cuda) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "invalid CUDA owner creation input");
        auto candidate = std::make_unique<FjCuda>();
        bool published = false;
        {
            std::lock_guard<std::mutex> lock(g_registry);
            if (!g_owner.load(std::memory_order_acquire)) {
                *out_cuda = candidate.release();
                g_owner.store(*out_cuda, std::memory_order_release);
                published = true;
            }
        }
        if (!published) return fail(error, FJ_STATUS_PREPARATION_FAILURE, "a film render runtime is already active");
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "CUDA owner allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fj_cuda_create");
    }
}

FJ_EXPORT FjStatus fj_cuda_check_terminal(FjCuda* cuda, FjErrorBuffer* error) {
    // Registration and re-entry precheck only: no drain, no lifetime exclusion.
    if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
    try {
        if (!cuda || t_in_call)
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "invalid or reentrant terminal precheck (runtime re-entered from its own call)");
        if (cuda != g_owner.load(std::memory_order_acquire))
            return fail(error, FJ_STATUS_PREPARATION_FAILURE, "terminal precheck requires the registered owner (unknown runtime handle)");
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (...) {
        return make_status(FJ_STATUS_INTERNAL_FAILURE);
    }
}

FJ_EXPORT FjStatus fj_cuda_inspect(FjCuda* cuda, const FjFrame* frame, FjCudaContext* out_context, FjErrorBuffer* error) {
    try {
        if (!cuda || !frame || !out_context || bad_error_buffer(error))
            return fail(bad_error_buffer(error) ? nullptr : error, FJ_STATUS_UNSUPPORTED_INPUT,
                        "invalid CUDA inspection arguments");
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        FjCudaContext ctx{};
        const FjStatus s = inspect_frame(frame, &ctx, error);
        if (s.category == FJ_STATUS_SUCCESS) {
            *out_context = ctx;
            write_error(error, "");
        }
        return s;
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "CUDA inspection allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fj_cuda_inspect");
    }
}

FJ_EXPORT FjRenderOutcome fj_cuda_render(FjCuda* cuda, const FjCudaContext* context, const FjFrame* frame,
                               const FjSubmission* submission, const FjPreparedHostData* prepared,
                               FjAbortCallback abort_callback, FjErrorBuffer* error) {
    FjRenderOutcome outcome{};
    try {
        if (!cuda || invalid_render_arguments(context, frame, submission, prepared) || bad_error_buffer(error)) {
            outcome.status = fail(bad_error_buffer(error) ? nullptr : error, FJ_STATUS_UNSUPPORTED_INPUT,
                                  "invalid CUDA render arguments (null render argument)");
            return outcome;
        }
        NativeCall call(cuda);
        if (!call.ok()) {
            outcome.status = call.refuse(error);
            return outcome;
        }
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
    try {
        if (!cuda || instance_token == 0 || bad_error_buffer(error))
            return fail(bad_error_buffer(error) ? nullptr : error, FJ_STATUS_UNSUPPORTED_INPUT,
                        "invalid CUDA instance retirement arguments (instance token is zero)");
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        // Static grain state is keyed by table content, not by instance: nothing
        // instance-owned remains after a render returns.
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "CUDA instance retirement allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "CUDA instance retirement failed with unknown exception");
    }
}

FJ_EXPORT FjStatus fj_cuda_shutdown(FjCuda* cuda, FjErrorBuffer* error) {
    if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
    if (!cuda || t_in_call)
        return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "invalid or reentrant CUDA shutdown (runtime re-entered from its own call)");
    try {
        OwnerEntry entry(cuda);
        if (!entry.ok()) return fail(error, FJ_STATUS_PREPARATION_FAILURE, "CUDA shutdown requires the registered owner");
        const FjStatus result = owner_shutdown(cuda);
        write_error(error, result.category == FJ_STATUS_SUCCESS ? std::string() : cuda->terminal_text);
        return result;
    } catch (const std::bad_alloc&) {
        return terminal_failure(cuda, make_status(FJ_STATUS_ALLOCATION_FAILURE), "CUDA shutdown admission allocation failed",
                                false, error);
    } catch (...) {
        return terminal_failure(cuda, make_status(FJ_STATUS_INTERNAL_FAILURE), "CUDA shutdown admission failed", false,
                                error);
    }
}

FJ_EXPORT FjStatus fj_cuda_destroy(FjCuda* cuda, FjErrorBuffer* error) {
    const bool diagnostic_valid = !bad_error_buffer(error);
    FjErrorBuffer* out = diagnostic_valid ? error : nullptr;
    if (!cuda) return fail(out, FJ_STATUS_UNSUPPORTED_INPUT, "CUDA destroy requires an owning handle");
    if (cuda != g_owner.load(std::memory_order_acquire))
        return fail(out, FJ_STATUS_PREPARATION_FAILURE, "CUDA destroy requires the registered owner (unknown runtime handle)");
    if (t_in_call)
        return terminal_failure(cuda, make_status(FJ_STATUS_UNSUPPORTED_INPUT),
                                "reentrant CUDA destroy retained the owner (runtime re-entered from its own call)", true, out);
    FjStatus result = make_status(FJ_STATUS_SUCCESS);
    bool closed = false;
    try {
        std::unique_lock<std::mutex> lock(g_registry);
        if (cuda != g_owner.load(std::memory_order_acquire))
            return fail(out, FJ_STATUS_PREPARATION_FAILURE, "CUDA destroy requires the registered owner (unknown runtime handle)");
        ++cuda->entrants;  // this call, until it decides the owner's fate
        lock.unlock();
        // Shutdown takes the call mutex; entrants not yet past admission then
        // see a closed owner.
        result = owner_shutdown(cuda);
        lock.lock();
        if (result.category == FJ_STATUS_SUCCESS && g_owner.load(std::memory_order_acquire) == cuda) {
            g_owner.store(nullptr, std::memory_order_release);
            // Every other entrant that passed the registration check leaves
            // before the owner's memory goes.
            g_registry_cv.wait(lock, [&] { return cuda->entrants == 1; });
            cuda->entrants = 0;
            closed = true;
        } else {
            if (--cuda->entrants == 0) g_registry_cv.notify_all();
            if (result.category == FJ_STATUS_SUCCESS) {
                // A concurrent destroy consumed the owner first.
                return fail(out, FJ_STATUS_PREPARATION_FAILURE, "CUDA destroy requires the registered owner");
            }
            cuda->lifecycle.store(Lifecycle::Retained, std::memory_order_release);
            write_error(out, cuda->terminal_text);
        }
    } catch (const std::bad_alloc&) {
        result = terminal_failure(cuda, make_status(FJ_STATUS_ALLOCATION_FAILURE), "CUDA destroy admission allocation failed",
                                  true, out);
    } catch (...) {
        result = terminal_failure(cuda, make_status(FJ_STATUS_INTERNAL_FAILURE), "CUDA destroy admission failed", true, out);
    }
    if (closed) {
        delete cuda;
        if (!diagnostic_valid) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        write_error(out, "");
        return make_status(FJ_STATUS_SUCCESS);
    }
    return result;
}

FJ_EXPORT FjStatus fjx_grain_apply(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* prepared,
                                   float* density_c, float* density_m, float* density_y, int32_t width, int32_t height,
                                   uint64_t stream, FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!prepared) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null prepared record");
        FjStatus st{};
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        float* const planes[3] = {density_c, density_m, density_y};
        return grain_apply_planes(*dev, *prepared, planes, width, height, reinterpret_cast<hipStream_t>(stream), error);
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
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!stage) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null diffusion stage");
        FjStatus st{};
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        float* const planes[3] = {red, green, blue};
        return diffusion_apply_planes(*dev, *stage, planes, width, height, reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_diffusion_apply");
    }
}

FJ_EXPORT FjStatus fjx_highlight_boost_apply(FjCuda* cuda, int32_t device_id, const FjHighlightBoost* boost, float* red,
                                             float* green, float* blue, int64_t samples, uint64_t stream,
                                             FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!boost) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null highlight boost descriptor");
        FjStatus st{};
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        float* const planes[3] = {red, green, blue};
        return highlight_boost_planes(*dev, *boost, planes, samples, reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_highlight_boost_apply");
    }
}

FJ_EXPORT size_t fjx_fit_saved_floats(int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return 0;
    return static_cast<size_t>(9) * static_cast<size_t>(width) * static_cast<size_t>(height);
}

FJ_EXPORT FjStatus fjx_fit_forward(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* nominal,
                                   const FjFitParams* params, const float* source, int32_t width, int32_t height,
                                   const float* grain_delta_cmy, float* destination, float* saved, uint64_t stream,
                                   FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!nominal || !params) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null fit record or parameters");
        FjStatus st = fit_precheck(*nominal, *params, width, height, error);
        if (st.category != FJ_STATUS_SUCCESS) return st;
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        return fit_forward(*dev, *nominal, *params, source, width, height, grain_delta_cmy, destination, saved,
                           reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_fit_forward");
    }
}

FJ_EXPORT FjStatus fjx_fit_backward(FjCuda* cuda, int32_t device_id, const FjPreparedHostData* nominal,
                                    const FjFitParams* params, const float* source, int32_t width, int32_t height,
                                    const float* saved, const float* grad_destination, float* grad_source,
                                    float* grad_params, uint64_t stream, FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!nominal || !params) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null fit record or parameters");
        FjStatus st = fit_precheck(*nominal, *params, width, height, error);
        if (st.category != FJ_STATUS_SUCCESS) return st;
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        return fit_backward(*dev, *nominal, *params, source, width, height, saved, grad_destination, grad_source,
                            grad_params, reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_fit_backward");
    }
}

FJ_EXPORT FjStatus fjx_glare_field(FjCuda* cuda, int32_t device_id, const FjScannerPostEffects* post, uint64_t seed,
                                   uint32_t medium, int32_t origin_x, int32_t origin_y, int32_t width, int32_t height,
                                   float* field, float* uniforms, uint64_t stream, FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!post) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null scanner post descriptor");
        FjStatus st{};
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        const hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        st = glare_field(*dev, *post, seed, medium, origin_x, origin_y, width, height, field, uniforms, s, error);
        if (st.category != FJ_STATUS_SUCCESS) return st;
        const hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail(error, e, "scanner glare");
        return st;
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "host allocation failed");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_glare_field");
    }
}

FJ_EXPORT FjStatus fjx_build_tc_lut(FjCuda* cuda, int32_t device_id, const FjFilmTcLutInput* input, float* samples,
                                    size_t count, uint64_t stream, FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!input) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null TC LUT input");
        FjStatus st{};
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        return build_tc_lut(*dev, *input, samples, count, reinterpret_cast<hipStream_t>(stream), error);
    } catch (const std::bad_alloc&) {
        return fail(error, FJ_STATUS_ALLOCATION_FAILURE, "film_tc_lut allocation failure");
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_build_tc_lut");
    }
}

FJ_EXPORT FjStatus fjx_capture_grain_input(FjCuda* cuda, int32_t device_id, float* density_c, float* density_m,
                                           float* density_y, int64_t samples, FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        const bool arm = density_c || density_m || density_y;
        if (arm && (!density_c || !density_m || !density_y || samples <= 0))
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "grain capture needs three planes and a sample count");
        FjStatus st{};
        DeviceState* dev = device_state(cuda, device_id, st, error);
        if (!dev) return st;
        dev->grain_capture[0] = density_c;
        dev->grain_capture[1] = density_m;
        dev->grain_capture[2] = density_y;
        dev->grain_capture_samples = arm ? samples : 0;
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_capture_grain_input");
    }
}

FJ_EXPORT FjStatus fjx_render_batch(FjCuda* cuda, const FjCudaContext* context, int32_t count, const FjFrame* frames,
                                    const FjSubmission* submissions, const FjPreparedHostData* const* prepared,
                                    int32_t streams, FjAbortCallback abort_callback, FjRenderOutcome* outcomes,
                                    FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
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
        FjStatus dst{};
        DeviceState* dev = device_state(cuda, context->device_id, dst, error);
        if (!dev) return dst;
        hipStream_t caller = has_stream ? reinterpret_cast<hipStream_t>(frames[0].stream) : nullptr;
        const FjStatus st =
            render_batch(*dev, count, frames, submissions, prepared, streams, caller, abort_callback, outcomes, error);
        if (!has_stream) {
            const hipError_t he = hipStreamSynchronize(caller);
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
        if (capacity < 0 || (capacity > 0 && !out)) return -1;
        NativeCall call(cuda);
        if (!call.ok()) return -1;
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
