// fj_plan.h — per-device state and the frame executor.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fj/juicer_cuda_api.h"
#include "../../include/fj/fj_mi355x_ext.h"
#include "../../include/fj/juicer_cuda_call_scope.h"
#include "fj_device.h"
#include "fj_fft.h"

namespace fjx {

// Device tables packed from prepared spans, cached by content identity.
struct DeviceTable {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t serial = 0;  // last frame that used it
};

struct DeviceState {
    int device = -1;
    int compute_units = 0;
    int max_line = 4096;  // longest DCT line the LDS transform accepts on this device
    int halation_max_line = 8192;  // longest row the halation row pass holds in LDS
    std::unordered_map<uint64_t, DeviceTable> tables;
    std::list<uint64_t> lru;  // most recent at front
    // Device transfer functions of spatial filters, by descriptor and frame size.
    std::unordered_map<uint64_t, DctTransfer> transfers;
    size_t table_bytes = 0;
    uint64_t frame_serial = 1;  // advanced per rendered frame

    // Error word ring: each frame clears one device word, its kernels OR scan
    // and DIR failure bits into it, and the word is copied to pinned memory.
    static constexpr int kRing = 8;
    unsigned int* error_words = nullptr;       // device [kRing]
    unsigned int* error_host = nullptr;        // pinned [kRing]
    hipEvent_t error_events[kRing] = {};
    bool error_pending[kRing] = {};
    int error_next = 0;

    // Frame workspaces for staged (spatial) schedules. Work that uses a slot
    // on one stream is ordered after the slot's previous user's stream
    // through `done` (recorded after each user's last launch). Single frames
    // use slot 0; a batch spreads its frames over the slots, one worker
    // stream per slot, so that frames overlap.
    struct WorkspaceSlot {
        void* ptr = nullptr;
        size_t bytes = 0;
        hipEvent_t done = nullptr;
        hipStream_t stream = nullptr;
        bool used = false;
    };
    static constexpr int kWorkspaceSlots = 4;
    WorkspaceSlot slots[kWorkspaceSlots];
    int active_slot = 0;
    hipStream_t workers[kWorkspaceSlots] = {};  // batch worker streams, created on first use
    hipEvent_t batch_events[kWorkspaceSlots + 1] = {};
    void* workspace() const { return slots[active_slot].ptr; }

    // One-shot copy of the densities entering the next frame's grain stage
    // (armed by fjx_capture_grain_input).
    float* grain_capture[3] = {};
    long long grain_capture_samples = 0;

    // Upload ring of the differentiable route's per-parameter tables: kFitRing
    // device slots of kFitSlotFloats fed from pinned staging by async copies
    // on the caller's stream; a slot is reused after the event recorded
    // behind its last reader.
    static constexpr int kFitRing = 64;
    static constexpr size_t kFitSlotFloats = size_t(1) << 16;
    float* fit_ring = nullptr;       // device [kFitRing][kFitSlotFloats]
    float* fit_staging = nullptr;    // pinned host, same shape
    hipEvent_t fit_events[kFitRing] = {};
    bool fit_pending[kFitRing] = {};
    int fit_next = 0;

    // Stage timing of the last traced frame (FJ_FRAME_TRACE_VERBOSE).
    static constexpr int kMaxStages = 32;
    hipEvent_t stage_events[kMaxStages + 1] = {};
    const char* stage_names[kMaxStages] = {};
    int stage_count = 0;

    hipError_t init(int dev);
    hipError_t release();
    // Returns the device copy of `bytes` host bytes under `key`, uploading once
    // (host == nullptr: uninitialized device bytes the caller fills). Tables
    // used by the current frame are never evicted during it.
    hipError_t table(uint64_t key, const void* host, size_t bytes, const void** out);
    hipError_t ensure_workspace(size_t bytes);  // of the active slot
    // Orders `stream` after the workspace's previous user; call before the
    // first launch that touches the workspace, then release after the last.
    hipError_t acquire_workspace(hipStream_t stream);
    hipError_t release_workspace(hipStream_t stream);
    // Collects completed error words of earlier frames; returns OR of their bits.
    unsigned int drain_errors(bool wait_all);
};

// The DCT-domain transfer function of a spatial-DIR descriptor at a frame
// size, built once and cached per device (ok false with `diag` on a malformed
// descriptor).
hipError_t dir_transfer_cached(DeviceState& dev, const FjSpatialDir& sd, int width, int height, DctTransfer& out,
                               std::string& diag, bool& ok);
// Whether a scanner post descriptor asks for lens blur, unsharp or glare.
bool scanner_post_active(const FjScannerPostEffects& sp);
// The span and tag checks of a render call's prepared record.
bool record_valid(const FjPreparedHostData& p, uint32_t& category, std::string& text);
// Device tables of one record placed by the caller (the fit's upload ring):
// when given, the assembly uses them instead of uploading and caching the
// record's own film curves and print spectra.
struct GivenTables {
    const float* film_curves = nullptr;     // packed_film_curves layout
    const float4* print_spectra = nullptr;  // packed_print_spectra layout
    int print_k = 0;
};
// The packed layouts the pixel kernels read.
std::vector<float> packed_film_curves(const FjPreparedHostData& p);
std::vector<float> packed_print_spectra(const FjPrint& pr, int& k);
// The per-pixel stage parameters of a prepared record (tables uploaded and
// cached by content); ok false with the status category and diagnostic.
hipError_t assemble_pixel_params(DeviceState& dev, const FjPreparedHostData& p, PixelParams& pp, uint32_t& category,
                                 std::string& text, bool& ok, const GivenTables* given = nullptr);

// `host`: the call scope's host preparation, or null (native builds).
FjRenderOutcome render_frame(DeviceState& dev, const FjFrame& frame, const FjSubmission& submission,
                             const FjPreparedHostData& prepared, FjAbortCallback abort_callback,
                             FjErrorBuffer* error, const FjHostPreparation* host = nullptr);

// Renders `count` frames of one device, frame i on worker stream i % streams
// with workspace slot i % streams, all ordered after the work already queued
// on `caller` and joined back into it. outcomes[i] is frame i's outcome;
// returns the first failing frame's status, else success.
FjStatus render_batch(DeviceState& dev, int count, const FjFrame* frames, const FjSubmission* submissions,
                      const FjPreparedHostData* const* prepared, int streams, hipStream_t caller,
                      FjAbortCallback abort_callback, FjRenderOutcome* outcomes, FjErrorBuffer* error);

// The visual-grain stage of `prepared` applied to three caller-owned C/M/Y
// density planes (device memory, width x height each) on `stream`.
FjStatus grain_apply_planes(DeviceState& dev, const FjPreparedHostData& prepared, float* const density[3], int width,
                            int height, hipStream_t stream, FjErrorBuffer* error);
// One diffusion stage (camera or enlarger descriptor) applied in place to
// three caller-owned R, G, B planes of a width x height frame.
FjStatus diffusion_apply_planes(DeviceState& dev, const FjDiffusion& stage, float* const planes[3], int width,
                                int height, hipStream_t stream, FjErrorBuffer* error);

// Highlight boost of three caller-owned carrier planes of `samples` floats
// each (B/G/R or R/G/B: the operator treats the channels alike), in place on
// `stream`; waits for the disposition and reports a failing one.
FjStatus highlight_boost_planes(DeviceState& dev, const FjHighlightBoost& boost, float* const planes[3],
                                long long samples, hipStream_t stream, FjErrorBuffer* error);
// The scanner-glare field of a width x height window whose first pixel is
// (origin_x, origin_y) of the frame's stream `seed` (medium 1 print, 0 film),
// blurred when the descriptor's glare flag asks for it, into `field`;
// `uniforms` (2 planes, optional) receives the raw FP32 uniforms.
// The differentiable print route (fj_fit.cpp): the host-side checks of a
// record and parameters (no device), one frame's forward, saving
// its state, and the backward from that state.
FjStatus fit_precheck(const FjPreparedHostData& nominal, const FjFitParams& th, int width, int height,
                      FjErrorBuffer* error);
FjStatus fit_forward(DeviceState& dev, const FjPreparedHostData& nominal, const FjFitParams& th, const float* source,
                     int width, int height, const float* delta, float* destination, float* saved, hipStream_t stream,
                     FjErrorBuffer* error);
FjStatus fit_backward(DeviceState& dev, const FjPreparedHostData& nominal, const FjFitParams& th, const float* source,
                      int width, int height, const float* saved, const float* grad_destination, float* grad_source,
                      float* grad_params, hipStream_t stream, FjErrorBuffer* error);
FjStatus glare_field(DeviceState& dev, const FjScannerPostEffects& post, uint64_t seed, uint32_t medium, int origin_x,
                     int origin_y, int width, int height, float* field, float* uniforms, hipStream_t stream,
                     FjErrorBuffer* error);
// The film TC-LUT build (fjx_build_tc_lut) on `dev`'s workspace.
FjStatus build_tc_lut(DeviceState& dev, const FjFilmTcLutInput& input, float* samples, size_t count,
                      hipStream_t stream, FjErrorBuffer* error);
// The checks a boost descriptor must pass (recipe bound, finite fields).
bool highlight_boost_valid(const FjHighlightBoost& boost);

uint64_t hash_bytes(const void* data, size_t bytes, uint64_t seed);

}  // namespace fjx
