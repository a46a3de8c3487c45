/* fj_call_harness.c — a host-side driver of the call scope for the tests.
 *
 * fj_cuda_call's continuation and completion return records by value, which
 * Python callbacks cannot, so the tests describe a call as a script of
 * operations and this C driver plays it inside the scope, recording every
 * status, diagnostic and record the scope returns. Test infrastructure only.
 */
#include <stdio.h>
#include <string.h>

#include "fj/fj_mi355x_ext.h"
#include "fj/juicer_cuda_call_scope.h"

enum {
    FJH_PRECHECK = 1,
    FJH_ADMIT = 2,
    FJH_INSPECT = 3,
    FJH_DESCRIPTORS = 4,
    FJH_PREVIEW = 5,
    FJH_RENDER = 6,
    FJH_RELEASE = 7,
    FJH_COMPLETE = 8,
    FJH_PUBLIC_INSPECT = 9, /* fj_cuda_inspect from inside the scope (re-entry) */
    FJH_TRACE = 10,
    FJH_RENDER_OTHER_CONTEXT = 11,
    FJH_PUBLIC_DESTROY = 12,  /* fj_cuda_destroy from inside the scope (re-entry) */
    FJH_PUBLIC_SHUTDOWN = 13, /* fj_cuda_shutdown from inside the scope (re-entry) */
};

#define FJH_MAX_OPS 16
#define FJH_TEXT 256

typedef struct FjhScript {
    /* inputs */
    uint32_t route;
    uint32_t op_count;
    uint32_t ops[FJH_MAX_OPS];
    const FjFrame* frame;
    const FjFrameMathInput* math;
    FjPreparedHostData* prepared; /* receives the composed descriptors when copy_descriptors */
    uint32_t copy_descriptors;
    const FjSubmission* submission;
    FjRect preview_source, preview_meter;
    uint32_t preview_method;
    uint32_t recovery_scope_entered;
    uint32_t force_fatal;
    uint32_t retire_on_completion; /* call fj_cuda_call_retire_context from the completion */
    uint32_t report_context_loss;  /* the completion input reports FJ_STATUS_CONTEXT_LOSS */
    uint32_t host_mode;            /* FJH_HOST_*: the host preparation passed to render */
    /* results */
    FjStatus status[FJH_MAX_OPS];
    char text[FJH_MAX_OPS][FJH_TEXT];
    FjCudaContext context;
    FjNativeFrameDescriptors descriptors;
    FjAutoExposure preview;
    FjRenderOutcome outcome;
    FjCallCompletionResult completed;
    uint32_t completion_calls;
    FjCallCompletionInput completion_seen;
    char completion_diagnostic[FJH_TEXT];
    uint32_t retirement;
    FjCallCompletionResult call_result;
    uint32_t host_calls[2]; /* scanner LUT, Gaussian callbacks made */
} FjhScript;

/* Host preparation callbacks (the reference's Rust side builds the scanner
 * LUT and Gaussian kernels; here the library's own builders stand in):
 *   NONE: no host preparation (native builds);
 *   DELEGATE: both callbacks return the builders' results;
 *   GAUSSIAN_FAIL: the Gaussian callback refuses (sigma, preparation failure);
 *   GAUSSIAN_HALF: the Gaussian callback returns the weights halved;
 *   LUT_FAIL: the scanner-LUT callback refuses (non-finite node);
 *   GAUSSIAN_BROKEN: the Gaussian view names another radius;
 *   LUT_BROKEN: the scanner-LUT view names another resolution. */
enum {
    FJH_HOST_NONE = 0,
    FJH_HOST_DELEGATE = 1,
    FJH_HOST_GAUSSIAN_FAIL = 2,
    FJH_HOST_GAUSSIAN_HALF = 3,
    FJH_HOST_LUT_FAIL = 4,
    FJH_HOST_GAUSSIAN_BROKEN = 5,
    FJH_HOST_LUT_BROKEN = 6,
};

#define FJH_MAX_TAPS 4097
static float g_half[FJH_MAX_TAPS];

static void fjh_put(FjErrorBuffer* e, const char* text) {
    size_t n = strlen(text);
    if (n > e->capacity - 1) n = e->capacity - 1;
    memcpy(e->data, text, n);
    e->data[n] = '\0';
    e->length = n;
}

static FjStatus fjh_host_lut(void* user, const FjScannerLutInput* in, FjScannerLutView* view,
                             FjScannerLutFailure* failure, FjErrorBuffer* error) {
    FjhScript* s = (FjhScript*)user;
    s->host_calls[0]++;
    if (s->host_mode == FJH_HOST_LUT_FAIL) {
        memset(view, 0, sizeof(*view));
        failure->reason = FJ_SCANNER_LUT_REASON_NONFINITE;
        failure->channel = 1;
        failure->index = 7;
        failure->buffer = 0;
        fjh_put(error, "host scanner LUT refused: non-finite node");
        FjStatus st = {FJ_STATUS_PREPARATION_FAILURE, FJ_API_NONE, 0};
        return st;
    }
    FjStatus st = fjx_prepare_scanner_lut(NULL, in, view, failure, error);
    if (st.category == FJ_STATUS_SUCCESS && s->host_mode == FJH_HOST_LUT_BROKEN) view->resolution += 1;
    return st;
}

static FjStatus fjh_host_gaussian(void* user, const FjGaussianKernelInput* in, FjGaussianKernelView* view,
                                  FjGaussianKernelFailure* failure, FjErrorBuffer* error) {
    FjhScript* s = (FjhScript*)user;
    s->host_calls[1]++;
    if (s->host_mode == FJH_HOST_GAUSSIAN_FAIL) {
        memset(view, 0, sizeof(*view));
        failure->reason = FJ_GAUSSIAN_KERNEL_SIGMA;
        fjh_put(error, "host gaussian refused: sigma");
        FjStatus st = {FJ_STATUS_PREPARATION_FAILURE, FJ_API_NONE, 0};
        return st;
    }
    FjStatus st = fjx_prepare_gaussian(NULL, in, view, failure, error);
    if (st.category != FJ_STATUS_SUCCESS) return st;
    if (s->host_mode == FJH_HOST_GAUSSIAN_BROKEN) view->radius += 1;
    if (s->host_mode == FJH_HOST_GAUSSIAN_HALF && view->weights.count <= FJH_MAX_TAPS) {
        for (size_t i = 0; i < view->weights.count; ++i) g_half[i] = 0.5f * view->weights.data[i];
        view->weights.data = g_half;
    }
    return st;
}

typedef struct FjhState {
    FjCuda* cuda;
    FjhScript* s;
} FjhState;

static FjCudaCompletion fjh_completion(void* user, void* scope, const FjCallCompletionInput* input) {
    FjhState* st = (FjhState*)user;
    FjhScript* s = st->s;
    FjCudaCompletion out;
    memset(&out, 0, sizeof(out));
    s->completion_calls++;
    s->completion_seen = *input;
    s->completion_seen.diagnostic.data = NULL;
    s->completion_diagnostic[0] = '\0';
    if (input->diagnostic.data) {
        size_t n = input->diagnostic.count < FJH_TEXT - 1 ? input->diagnostic.count : FJH_TEXT - 1;
        memcpy(s->completion_diagnostic, input->diagnostic.data, n);
        s->completion_diagnostic[n] = '\0';
    }
    if (s->retire_on_completion) {
        out.retirement = fj_cuda_call_retire_context(scope, input);
        s->retirement = out.retirement;
    }
    out.result.outcome = input->outcome;
    out.result.fatal = input->force_fatal;
    return out;
}

static FjCallCompletionResult fjh_continuation(void* user, FjCudaCallAccess a) {
    FjhState* st = (FjhState*)user;
    FjhScript* s = st->s;
    FjCallCompletionResult last;
    memset(&last, 0, sizeof(last));
    int rendered = 0;
    for (uint32_t i = 0; i < s->op_count && i < FJH_MAX_OPS; ++i) {
        FjErrorBuffer err = {s->text[i], FJH_TEXT, 0};
        FjStatus r;
        memset(&r, 0, sizeof(r));
        switch (s->ops[i]) {
            case FJH_PRECHECK: r = a.precheck(a.user, &err); break;
            case FJH_ADMIT: r = a.admit(a.user, &err); break;
            case FJH_INSPECT: r = a.inspect(a.user, s->frame, &s->context, &err); break;
            case FJH_DESCRIPTORS:
                r = a.frame_descriptors(a.user, s->math, &s->descriptors, &err);
                if (r.category == FJ_STATUS_SUCCESS && s->copy_descriptors && s->prepared) {
                    s->prepared->spatial_dir = s->descriptors.spatial_dir;
                    s->prepared->grain = s->descriptors.grain;
                    s->prepared->effects = s->descriptors.effects;
                }
                break;
            case FJH_PREVIEW:
                r = a.preview(a.user, &s->preview_source, &s->preview_meter, s->preview_method, &s->preview, &err);
                break;
            case FJH_RENDER:
            case FJH_RENDER_OTHER_CONTEXT: {
                FjAbortCallback abort_cb = {NULL, NULL};
                FjCudaContext ctx = s->context;
                if (s->ops[i] == FJH_RENDER_OTHER_CONTEXT) ctx.context += 1000;
                FjHostPreparation host = {s, fjh_host_lut, fjh_host_gaussian};
                s->outcome = a.render(a.user, &ctx, s->frame, s->submission, s->prepared,
                                      s->host_mode == FJH_HOST_NONE ? NULL : &host, abort_cb, &err);
                r = s->outcome.status;
                rendered = 1;
                break;
            }
            case FJH_RELEASE:
                a.release(a.user);
                break;
            case FJH_COMPLETE: {
                FjCallCompletionInput in;
                memset(&in, 0, sizeof(in));
                if (rendered) {
                    in.outcome = s->outcome;
                    in.render_outcome_valid = 1;
                }
                if (s->report_context_loss) in.outcome.status.category = FJ_STATUS_CONTEXT_LOSS;
                in.recovery_scope_entered = s->recovery_scope_entered;
                in.has_inspected_context = s->context.context != 0;
                in.inspected_context = s->context;
                in.force_fatal = s->force_fatal;
                s->completed = a.complete(a.user, &in);
                r = s->completed.outcome.status;
                last = s->completed;
                break;
            }
            case FJH_PUBLIC_INSPECT: {
                FjCudaContext ctx;
                r = fj_cuda_inspect(st->cuda, s->frame, &ctx, &err);
                break;
            }
            case FJH_PUBLIC_DESTROY: r = fj_cuda_destroy(st->cuda, &err); break;
            case FJH_PUBLIC_SHUTDOWN: r = fj_cuda_shutdown(st->cuda, &err); break;
            case FJH_TRACE:
                if (a.trace_submission) a.trace_submission(a.user, s->submission, 0);
                break;
            default: break;
        }
        s->status[i] = r;
    }
    return last;
}

/* Plays the script inside fj_cuda_call(cuda, route, ...). */
int fjh_run(FjCuda* cuda, FjhScript* s) {
    FjhState st = {cuda, s};
    s->call_result = fj_cuda_call(cuda, s->route, fjh_continuation, fjh_completion, &st);
    return 0;
}

/* fj_cuda_call with caller-chosen null arguments (argument validation). */
FjCallCompletionResult fjh_call_raw(FjCuda* cuda, uint32_t route, int with_continuation, int with_completion,
                                    int with_user) {
    static FjhScript empty;
    FjhState st = {cuda, &empty};
    return fj_cuda_call(cuda, route, with_continuation ? fjh_continuation : NULL,
                        with_completion ? fjh_completion : NULL, with_user ? (void*)&st : NULL);
}

size_t fjh_script_size(void) { return sizeof(FjhScript); }
