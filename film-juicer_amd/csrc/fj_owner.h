// fj_owner.h — the registered runtime owner and the locked operations that
// the C-ABI entry points and the call scope share.
//
// One owner per process (native/juicer_cuda_api.cpp:245); every call holds
// the owner's mutex and a thread-local re-entry guard
// (NativeCall, native/juicer_cuda_api.cpp:674-684).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_plan.h"

struct FjCuda {
    std::mutex mu;
    bool open = true;
    std::map<int, std::unique_ptr<fjx::DeviceState>> devices;
};

namespace fjx {

extern std::atomic<FjCuda*> g_owner;
extern thread_local bool t_in_call;

FjStatus make_status(uint32_t category, uint32_t api = FJ_API_NONE, int32_t code = 0);
// FjErrorBuffer rules: capacity 0 permits a null buffer; otherwise a
// NUL-terminated text, length without the NUL, truncated to fit.
void write_error(FjErrorBuffer* error, const std::string& text);
FjStatus fail(FjErrorBuffer* error, uint32_t category, const std::string& text, uint32_t api = FJ_API_NONE,
              int32_t code = 0);
FjStatus hip_fail(FjErrorBuffer* error, hipError_t e, const char* where);

// Marks the calling thread as inside a native call for the guard's lifetime;
// ok is false when the thread already is.
struct CallGuard {
    bool ok;
    CallGuard() : ok(!t_in_call) {
        if (ok) t_in_call = true;
    }
    ~CallGuard() {
        if (ok) t_in_call = false;
    }
    CallGuard(const CallGuard&) = delete;
    CallGuard& operator=(const CallGuard&) = delete;
};

// Device and context token of a frame whose images are HIP device memory.
FjStatus inspect_frame(const FjFrame* frame, FjCudaContext* out, FjErrorBuffer* error);
// fj_cuda_render's work with the owner's lock already held.
FjRenderOutcome render_locked(FjCuda* cuda, const FjCudaContext* context, const FjFrame* frame,
                              const FjSubmission* submission, const FjPreparedHostData* prepared,
                              FjAbortCallback abort_callback, FjErrorBuffer* error);
// Releases the native state of one device (context-loss retirement) with the
// owner's lock held; false when releasing it reported an error.
bool retire_device_locked(FjCuda* cuda, int32_t device_id);

}  // namespace fjx
