// fj_call_scope.cpp — the synchronous host call scope
// (include/fj/juicer_cuda_call_scope.h).
//
// Behaviour follows the reference's CallScope (native/juicer_cuda_call_scope.cpp):
//   fj_cuda_call / completion forwarding          :648-660, :621-646
//   precheck / admit / release                    :275-323
//   inspect (binds frame, context, pixel size)    :325-377
//   preview (auto-exposure preview descriptor)    :379-423
//   frame_descriptors (extent binding, DIR, grain,
//     effects composition, seed binding)          :425-551
//   render (admitted, inspected context only)     :553-590
//   complete (once, after release)                :592-611
//   retire_context (context loss)                 :662-693
//   trace flags and sinks                         :700-733
// The scope lives on fj_cuda_call's stack; admission holds the owner's mutex
// and the thread's re-entry guard until release.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/fj/juicer_cuda_call_scope.h"
#include "fj_frame_descriptors.h"
#include "fj_owner.h"

using namespace fjx;

#define FJ_EXPORT __attribute__((visibility("default")))

namespace {

const char* const kInvalidScope = "invalid CUDA call scope";
const char* const kMissingAdmission = "CUDA call scope is not admitted";

// Trace level from FJ_MI355X_TRACE (0 off, 1 info, 3 verbose), read once.
int trace_level() {
    static const int level = [] {
        const char* v = std::getenv("FJ_MI355X_TRACE");
        return v ? std::atoi(v) : 0;
    }();
    return level;
}

void trace_line(const char* tag, const char* data, size_t count) {
    std::fprintf(stderr, "[fj %s] %.*s\n", tag, static_cast<int>(count), data ? data : "");
}

FjCallCompletionResult fatal_result(uint32_t category) {
    FjCallCompletionResult r{};
    r.outcome.status = make_status(category);
    r.fatal = 1;
    return r;
}

class CallScope {
public:
    CallScope(FjCuda* owner, uint32_t route, FjCudaCompletionContinuation completion, void* user)
        : owner_(owner), route_(route), completion_(completion), user_(user) {}
    ~CallScope() { drop_admission(); }
    CallScope(const CallScope&) = delete;
    CallScope& operator=(const CallScope&) = delete;

    FjCudaCallAccess access() {
        FjCudaCallAccess a{};
        a.user = this;
        a.precheck = &CallScope::precheck;
        a.admit = &CallScope::admit;
        a.release = &CallScope::release;
        a.inspect = &CallScope::inspect;
        a.preview = &CallScope::preview;
        a.frame_descriptors = &CallScope::frame_descriptors;
        a.render = &CallScope::render;
        a.complete = &CallScope::complete;
        a.trace_submission = &CallScope::trace_submission;
        return a;
    }

    uint32_t retire_context(const FjCallCompletionInput* in) {
        if (!in || !completed_ || admitted_ || !owner_ || in->recovery_scope_entered == 0 ||
            in->has_inspected_context == 0)
            return 0;
        if (!recovery_pending_ && in->outcome.status.category == FJ_STATUS_CONTEXT_LOSS) recovery_pending_ = true;
        if (!recovery_pending_) return 0;
        // Held through the caller's latch invalidation; dropped when the
        // completion continuation returns.
        if (!take_admission()) return 1u;
        const bool accepted = retire_device_locked(owner_, in->inspected_context.device_id);
        recovery_pending_ = false;
        return 1u | (accepted ? 2u : 0u);
    }

private:
    static CallScope* from(void* user) { return static_cast<CallScope*>(user); }

    FjStatus error(FjErrorBuffer* err, uint32_t category, const std::string& text) {
        diagnostic_ = text;
        return fail(err, category, text);
    }

    // Admission holds a NativeCall (re-entry guard, owner entry, call mutex)
    // until release or the scope's end.
    bool take_admission(FjStatus* refused = nullptr, std::string* text = nullptr) {
        if (admitted_) return true;
        call_ = std::make_unique<NativeCall>(owner_);
        if (!call_->ok()) {
            if (refused) *refused = call_->status;
            if (text) *text = call_->text;
            call_.reset();
            return false;
        }
        admitted_ = true;
        return true;
    }

    void drop_admission() {
        call_.reset();
        admitted_ = false;
    }

    static FjStatus precheck(void* user, FjErrorBuffer* err) {
        CallScope* s = from(user);
        if (!s || !s->owner_ || s->completed_ || s->admission_attempted_)
            return fail(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
        try {
            // check_native_call_admission (native/juicer_cuda_api.cpp:648-655),
            // plus the lifecycle the admission itself checks.
            if (t_in_call) return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, "runtime re-entered from its own call");
            OwnerEntry entry(s->owner_);
            if (!entry.ok())
                return s->error(err, FJ_STATUS_PREPARATION_FAILURE,
                                "CUDA call requires the registered owner (unknown runtime handle)");
            if (!s->owner_->accepting())
                return s->error(err, FJ_STATUS_PREPARATION_FAILURE, "CUDA owner admission is closed (runtime is shut down)");
            s->diagnostic_.clear();
            write_error(err, "");
            return make_status(FJ_STATUS_SUCCESS);
        } catch (...) {
            return s->error(err, FJ_STATUS_INTERNAL_FAILURE, "CUDA call precheck failed");
        }
    }

    static FjStatus admit(void* user, FjErrorBuffer* err) {
        CallScope* s = from(user);
        if (!s) return fail(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
        if (!s->owner_ || s->admitted_ || s->admission_attempted_ || s->completed_)
            return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
        s->admission_attempted_ = true;
        try {
            FjStatus refused{};
            std::string text;
            if (!s->take_admission(&refused, &text)) return s->error(err, refused.category, text);
            s->diagnostic_.clear();
            write_error(err, "");
            return make_status(FJ_STATUS_SUCCESS);
        } catch (...) {
            s->drop_admission();
            return s->error(err, FJ_STATUS_INTERNAL_FAILURE, "CUDA call admission failed");
        }
    }

    static void release(void* user) {
        CallScope* s = from(user);
        if (s && s->admitted_ && !s->released_) {
            s->drop_admission();
            s->released_ = true;
        }
    }

    static FjStatus inspect(void* user, const FjFrame* frame, FjCudaContext* context, FjErrorBuffer* err) {
        if (context) *context = FjCudaContext{};
        CallScope* s = from(user);
        if (!s) return fail(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
        if (!s->admitted_) return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, kMissingAdmission);
        try {
            s->diagnostic_.clear();
            s->has_context_ = false;
            if (!context) return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, "out_context is null");
            if (!s->owner_->accepting())
                return s->error(err, FJ_STATUS_PREPARATION_FAILURE, "CUDA owner admission is closed (runtime is shut down)");
            char text[2048];
            FjErrorBuffer local{text, sizeof(text), 0};
            FjCudaContext ctx{};
            const FjStatus st = inspect_frame(frame, &ctx, &local);
            if (st.category != FJ_STATUS_SUCCESS) {
                s->diagnostic_.assign(text, local.length);
                write_error(err, s->diagnostic_);
                return st;
            }
            *context = ctx;
            s->frame_ = *frame;
            s->context_ = ctx;
            s->pixel_size_um_ = frame->pixel_size_um;
            s->has_context_ = true;
            write_error(err, "");
            return st;
        } catch (...) {
            if (context) *context = FjCudaContext{};
            return s->error(err, FJ_STATUS_INTERNAL_FAILURE, "CUDA inspection scope failed");
        }
    }

    static FjStatus preview(void* user, const FjRect* source, const FjRect* meter, uint32_t method, FjAutoExposure* out,
                            FjErrorBuffer* err) {
        if (out) *out = FjAutoExposure{};
        CallScope* s = from(user);
        if (!s) return fail(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
        if (s->completed_ || !out || !source || !meter || method > FJ_METER_HIGHLIGHT_WEIGHTED ||
            source->x2 < source->x1 || source->y2 < source->y1 || meter->x2 < meter->x1 || meter->y2 < meter->y1)
            return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, "invalid auto-exposure preview arguments");
        const int64_t w = static_cast<int64_t>(meter->x2) - meter->x1;
        const int64_t h = static_cast<int64_t>(meter->y2) - meter->y1;
        if (std::max(w, h) > INT32_MAX)
            return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, "auto-exposure preview extent exceeds native limits");
        auto_exposure_preview(*source, *meter, method, *out);
        s->diagnostic_.clear();
        write_error(err, "");
        return make_status(FJ_STATUS_SUCCESS);
    }

    static FjStatus frame_descriptors(void* user, const FjFrameMathInput* in, FjNativeFrameDescriptors* out,
                                      FjErrorBuffer* err) {
        if (out) *out = FjNativeFrameDescriptors{};
        CallScope* s = from(user);
        if (!s) return fail(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
        if (s->completed_ || !s->admitted_ || !s->has_context_ || !in || !out || s->route_ > FJ_ROUTE_POSITIVE_PRINT)
            return s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, kMissingAdmission);
        try {
            std::string diag;
            FjNativeFrameDescriptors d{};
            const uint32_t category = compose_frame_descriptors(s->route_, s->frame_, *in, d, diag);
            if (category != FJ_STATUS_SUCCESS) return s->error(err, category, diag);
            *out = d;
            s->diagnostic_.clear();
            write_error(err, "");
            return make_status(FJ_STATUS_SUCCESS);
        } catch (...) {
            *out = FjNativeFrameDescriptors{};
            return s->error(err, FJ_STATUS_INTERNAL_FAILURE, "native frame descriptor composition failed");
        }
    }

    static FjRenderOutcome render(void* user, const FjCudaContext* context, const FjFrame* frame,
                                  const FjSubmission* submission, const FjPreparedHostData* prepared,
                                  const FjHostPreparation* host, FjAbortCallback abort_callback, FjErrorBuffer* err) {
        FjRenderOutcome outcome{};
        CallScope* s = from(user);
        if (!s) {
            outcome.status = fail(err, FJ_STATUS_UNSUPPORTED_INPUT, kInvalidScope);
            return outcome;
        }
        if (!s->admitted_) {
            outcome.status = s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, kMissingAdmission);
            return outcome;
        }
        if (!s->has_context_ || !context || context->device_id != s->context_.device_id ||
            context->context != s->context_.context) {
            outcome.status = s->error(err, FJ_STATUS_UNSUPPORTED_INPUT, "CUDA render context differs from inspected scope");
            return outcome;
        }
        try {
            char text[2048];
            FjErrorBuffer local{text, sizeof(text), 0};
            outcome = render_locked(s->owner_, context, frame, submission, prepared, abort_callback, &local, host);
            s->diagnostic_.assign(text, local.length);
            write_error(err, s->diagnostic_);
            if (outcome.status.category == FJ_STATUS_CONTEXT_LOSS) s->recovery_pending_ = true;
            return outcome;
        } catch (const std::bad_alloc&) {
            outcome.status = s->error(err, FJ_STATUS_ALLOCATION_FAILURE, "CUDA render allocation failed");
            return outcome;
        } catch (...) {
            outcome.status = s->error(err, FJ_STATUS_INTERNAL_FAILURE, "CUDA render failed");
            return outcome;
        }
    }

    static FjCallCompletionResult complete(void* user, const FjCallCompletionInput* in) {
        CallScope* s = from(user);
        if (!s) return fatal_result(FJ_STATUS_UNSUPPORTED_INPUT);
        if (s->completed_ || !in || s->admitted_ || !s->completion_) return fatal_result(FJ_STATUS_INTERNAL_FAILURE);
        s->completed_ = true;
        try {
            FjCallCompletionInput forwarded = *in;
            if (!s->diagnostic_.empty()) forwarded.diagnostic = FjStringView{s->diagnostic_.data(), s->diagnostic_.size()};
            const FjCudaCompletion report = s->completion_(s->user_, s, &forwarded);
            s->drop_admission();  // a retirement's re-admission ends with the completion
            return report.result;
        } catch (...) {
            s->drop_admission();
            FjCallCompletionResult r{};
            r.outcome = in->outcome;
            r.fatal = 1;
            return r;
        }
    }

    static void trace_submission(void* user, const FjSubmission* submission, uint32_t reused) {
        CallScope* s = from(user);
        if (!s || !submission || !s->has_context_ || (s->frame_.flags & FJ_FRAME_TRACE_VERBOSE) == 0 ||
            trace_level() < 3)
            return;
        char line[256];
        const int n = std::snprintf(line, sizeof(line), "path=hip action=%s frame_token=%llu snapshot_id=%llu instance_token=%llu",
                                    reused ? "reuse" : "new", static_cast<unsigned long long>(submission->frame_token),
                                    static_cast<unsigned long long>(submission->submission_id),
                                    static_cast<unsigned long long>(submission->instance_token));
        if (n > 0) trace_line("MSSNP", line, static_cast<size_t>(std::min<int>(n, sizeof(line) - 1)));
    }

    FjCuda* owner_;
    uint32_t route_;
    FjCudaCompletionContinuation completion_;
    void* user_;
    std::unique_ptr<NativeCall> call_;
    bool admitted_ = false;
    bool admission_attempted_ = false;
    bool released_ = false;
    bool completed_ = false;
    bool has_context_ = false;
    bool recovery_pending_ = false;
    FjFrame frame_{};
    FjCudaContext context_{};
    float pixel_size_um_ = 0.0f;
    std::string diagnostic_;

};

}  // namespace

extern "C" {

FJ_EXPORT FjCallCompletionResult fj_cuda_call(FjCuda* owner, uint32_t route, FjCudaCallContinuation continuation,
                                              FjCudaCompletionContinuation completion, void* user) {
    if (!owner || route > FJ_ROUTE_POSITIVE_PRINT || !continuation || !completion || !user)
        return fatal_result(FJ_STATUS_UNSUPPORTED_INPUT);
    try {
        CallScope scope(owner, route, completion, user);
        return continuation(user, scope.access());
    } catch (...) {
        return fatal_result(FJ_STATUS_INTERNAL_FAILURE);
    }
}

FJ_EXPORT uint32_t fj_cuda_call_retire_context(void* scope, const FjCallCompletionInput* input) {
    if (!scope) return 0;
    try {
        return static_cast<CallScope*>(scope)->retire_context(input);
    } catch (...) {
        return 0;
    }
}

FJ_EXPORT uint32_t fj_cuda_call_trace_flags(void) {
    const int level = trace_level();
    return (level >= 1 ? FJ_FRAME_TRACE_INFO : 0u) | (level >= 3 ? FJ_FRAME_TRACE_VERBOSE : 0u);
}

FJ_EXPORT void fj_cuda_call_trace_cutoff(FjStringView text) {
    if (trace_level() >= 1) trace_line("SPEKTRAFILM", text.data, text.data ? text.count : 0);
}

FJ_EXPORT void fj_cuda_call_trace_failure(FjStringView text) {
    if (trace_level() >= 1) trace_line("HIP", text.data, text.data ? text.count : 0);
}

FJ_EXPORT FjStatus fjx_frame_descriptors(uint32_t route, const FjFrame* frame, const FjFrameMathInput* input,
                                         FjNativeFrameDescriptors* out, FjErrorBuffer* error) {
    try {
        if (out) *out = FjNativeFrameDescriptors{};
        if (!frame || !input || !out) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "null frame descriptor argument");
        std::string diag;
        const uint32_t category = compose_frame_descriptors(route, *frame, *input, *out, diag);
        return fail(error, category, diag);
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "native frame descriptor composition failed");
    }
}

}  // extern "C"
