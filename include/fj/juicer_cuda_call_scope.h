/* juicer_cuda_call_scope.h — the synchronous call scope the host's plug-in
 * edge renders through (C part of the reference's
 * native/juicer_cuda_call_scope.h; layouts pinned by its
 * tests/ffi/prepared_call_abi_facts.inc).
 *
 * fj_cuda_call builds a scope on the calling thread and hands the host's
 * continuation a table of entry points (FjCudaCallAccess). Inside the
 * continuation the host prechecks, admits (takes the owner's lock and the
 * thread's re-entry guard), inspects, composes frame descriptors, renders,
 * releases, and finally completes. Table, user pointer and every borrowed
 * view expire when fj_cuda_call returns.
 */
#ifndef FJ_MI355X_CALL_SCOPE_H
#define FJ_MI355X_CALL_SCOPE_H

#include "juicer_cuda_frame_values.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host-side preparation callbacks borrowed for one render (:12-16). This
 * library builds the scanner LUT and Gaussian weights natively, so a render
 * through the scope never calls them; they are accepted for drop-in use. */
typedef struct FjHostPreparation {
    void* user;
    FjStatus (*scanner_lut)(void*, const FjScannerLutInput*, FjScannerLutView*, FjScannerLutFailure*, FjErrorBuffer*);
    FjStatus (*gaussian)(void*, const FjGaussianKernelInput*, FjGaussianKernelView*, FjGaussianKernelFailure*, FjErrorBuffer*);
} FjHostPreparation;

typedef struct FjCallCompletionInput {
    FjRenderOutcome outcome;
    FjStringView diagnostic;
    uint32_t render_outcome_valid;
    uint32_t abort_failed;
    uint32_t recovery_scope_entered;
    uint32_t has_inspected_context;
    FjCudaContext inspected_context;
    uint32_t force_fatal;
} FjCallCompletionInput;

typedef struct FjCallCompletionResult {
    FjRenderOutcome outcome;
    uint32_t fatal;
} FjCallCompletionResult;

/* Entry points of one scope (:34-46). `user` is the scope. */
typedef struct FjCudaCallAccess {
    void* user;
    FjStatus (*precheck)(void*, FjErrorBuffer*);
    FjStatus (*admit)(void*, FjErrorBuffer*);
    void (*release)(void*);
    FjStatus (*inspect)(void*, const FjFrame*, FjCudaContext*, FjErrorBuffer*);
    FjStatus (*preview)(void*, const FjRect*, const FjRect*, uint32_t, FjAutoExposure*, FjErrorBuffer*);
    FjStatus (*frame_descriptors)(void*, const FjFrameMathInput*, FjNativeFrameDescriptors*, FjErrorBuffer*);
    FjRenderOutcome (*render)(void*, const FjCudaContext*, const FjFrame*, const FjSubmission*,
                              const FjPreparedHostData*, const FjHostPreparation*, FjAbortCallback, FjErrorBuffer*);
    FjCallCompletionResult (*complete)(void*, const FjCallCompletionInput*);
    void (*trace_submission)(void*, const FjSubmission*, uint32_t reused);
} FjCudaCallAccess;

/* Diagnostic sink (:48-52): the trace levels enabled in this process
 * (FJ_FRAME_TRACE_INFO / FJ_FRAME_TRACE_VERBOSE, from the FJ_MI355X_TRACE
 * environment variable: 1 info, 3 verbose) and two borrowed-text sinks that
 * write to stderr when tracing is on. */
uint32_t fj_cuda_call_trace_flags(void);
void fj_cuda_call_trace_cutoff(FjStringView);
void fj_cuda_call_trace_failure(FjStringView);

typedef FjCallCompletionResult (*FjCudaCallContinuation)(void*, FjCudaCallAccess);
typedef struct FjCudaCompletion {
    FjCallCompletionResult result;
    uint32_t retirement; /* bit 0 attempted, bit 1 accepted */
    FjStatus invalidation;
    uint32_t latch_cleared;
} FjCudaCompletion;
typedef FjCudaCompletion (*FjCudaCompletionContinuation)(void*, void*, const FjCallCompletionInput*);

/* One synchronous host call (:58-65): runs continuation(user, access) and
 * returns its result. `complete` forwards to completion(user, scope, input)
 * after the scope's lock is released, on the same thread. */
FjCallCompletionResult fj_cuda_call(FjCuda*, uint32_t route, FjCudaCallContinuation, FjCudaCompletionContinuation,
                                    void*);
/* Only during that completion (:66-69): when the render ended in context
 * loss, retires the native state of the inspected device; returns bit 0
 * (attempted) and bit 1 (accepted). */
uint32_t fj_cuda_call_retire_context(void*, const FjCallCompletionInput*);

#ifdef __cplusplus
}
#endif

#endif
