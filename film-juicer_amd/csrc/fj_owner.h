// fj_owner.h — the registered runtime owner, its lifecycle, and the locked
// operations that the C-ABI entry points and the call scope share.
//
// One owner per process (native/juicer_cuda_api.cpp:245-280); every call
// holds the owner's call mutex and a thread-local re-entry guard (NativeCall,
// native/juicer_cuda_api.cpp:648-684); the owner's lifecycle is the
// reference's Accepting -> Closing -> Closed | Blocked | Retained machine
// (native/juicer_cuda_api.cpp:24-120).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>

#include "../../include/fj/juicer_cuda_api.h"
#include "fj_plan.h"

namespace fjx {

enum class Lifecycle : uint8_t {
    Accepting,  // calls are admitted
    Closing,    // one thread is shutting the owner down
    Closed,     // native state released; calls are refused
    Blocked,    // a terminal operation failed; the state is kept, calls are refused
    Retained,   // a destroy could not close the owner; it stays registered
};

}  // namespace fjx

struct FjCuda {
    std::mutex mu;  // the call mutex: one native call at a time
    std::atomic<fjx::Lifecycle> lifecycle{fjx::Lifecycle::Accepting};
    // Threads inside an entry point that passed the registration check; a
    // destroy frees the owner only after they have left. Guarded by the
    // registry mutex.
    int entrants = 0;
    // Published before the lifecycle turns Blocked or Retained.
    FjStatus terminal{FJ_STATUS_PREPARATION_FAILURE, FJ_API_NONE, 0};
    std::string terminal_text;
    std::map<int, std::unique_ptr<fjx::DeviceState>> devices;

    bool accepting() const { return lifecycle.load(std::memory_order_acquire) == fjx::Lifecycle::Accepting; }
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
// A buffer with capacity but no storage (valid_diagnostic,
// native/juicer_cuda_api.cpp:150-156): the call is refused without writing.
inline bool bad_error_buffer(const FjErrorBuffer* error) { return error && error->capacity != 0 && !error->data; }

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

// Counts the calling thread as an entrant of `cuda` when it is the
// registered owner (checked under the registry mutex); ok() is false
// otherwise. The owner is not freed while an entry exists.
class OwnerEntry {
public:
    explicit OwnerEntry(FjCuda* cuda);
    ~OwnerEntry();
    OwnerEntry(const OwnerEntry&) = delete;
    OwnerEntry& operator=(const OwnerEntry&) = delete;
    bool ok() const { return cuda_ != nullptr; }

private:
    FjCuda* cuda_ = nullptr;
};

// Admission of one native call: the re-entry guard, the registration check,
// the call mutex and an accepting lifecycle (check_native_call_admission and
// NativeCall, native/juicer_cuda_api.cpp:648-684). On refusal `status` holds
// the reference's category (re-entry: unsupported input; unregistered or
// closed owner: preparation failure) and `text` the diagnostic.
struct NativeCall {
    OwnerEntry entry;
    CallGuard guard;
    std::unique_lock<std::mutex> lock;
    FjStatus status{};
    std::string text;
    explicit NativeCall(FjCuda* cuda);
    bool ok() const { return status.category == FJ_STATUS_SUCCESS; }
    FjStatus refuse(FjErrorBuffer* error) const { return fail(error, status.category, text); }
};

// The owner's shutdown (FjCuda::shutdown, native/juicer_cuda_api.cpp:63-105):
// Closed returns success, Blocked/Retained the published terminal status,
// Closing waits for the closing thread; Accepting closes: waits for the call
// mutex, releases every device's state and publishes Closed (or Blocked with
// the failure).
FjStatus owner_shutdown(FjCuda* cuda);

// Device and context token of a frame whose images are HIP device memory.
FjStatus inspect_frame(const FjFrame* frame, FjCudaContext* out, FjErrorBuffer* error);
// fj_cuda_render's work with the owner's lock already held.
FjRenderOutcome render_locked(FjCuda* cuda, const FjCudaContext* context, const FjFrame* frame,
                              const FjSubmission* submission, const FjPreparedHostData* prepared,
                              FjAbortCallback abort_callback, FjErrorBuffer* error,
                              const FjHostPreparation* host = nullptr);
// Releases the native state of one device (context-loss retirement) with the
// owner's lock held; false when releasing it reported an error.
bool retire_device_locked(FjCuda* cuda, int32_t device_id);

}  // namespace fjx
