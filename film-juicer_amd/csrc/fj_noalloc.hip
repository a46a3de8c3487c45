// fj_noalloc.hip — the highlight-boost stage for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops):
// fjx_highlight_boost_apply_static runs the kernels of
// fjx_highlight_boost_apply with their scratch (reduction partials,
// disposition state, error word) in a fixed file-scope device array instead
// of the device's workspace slot, and allocates nothing.
//
// `scratch` below is a device symbol: in host code its name is a host-side
// stand-in, never a device address. Each device's address of the array is
// taken once with hipGetSymbolAddress (with that device current) and only
// that address is passed to kernels.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/fj/fj_mi355x_ext.h"
#include "fj_highlight.h"
#include "fj_host.h"
#include "fj_owner.h"

#define FJ_EXPORT __attribute__((visibility("default")))

namespace fjx {
namespace {

// Layout of the array: [HbState, padded to 256 B][error word].
constexpr size_t kStateBytes = (sizeof(HbState) + 255) / 256 * 256;
constexpr size_t kScratchBytes = kStateBytes + 256;
// Samples per plane the entry point accepts (the kernels stride over any
// count with at most kHbMaxBlocks blocks; this bounds the call, not the array).
constexpr int64_t kMaxSamples = int64_t(1) << 34;
constexpr int kMaxDevices = 64;

}  // namespace
}  // namespace fjx

static __device__ __align__(256) unsigned char scratch[fjx::kScratchBytes];

namespace fjx {
namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), resolved once.
hipError_t scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = ptr;
    }
    *out = static_cast<unsigned char*>(g_scratch_ptr[device]);
    return hipSuccess;
}

}  // namespace
}  // namespace fjx

using namespace fjx;

extern "C" {

FJ_EXPORT FjStatus fjx_highlight_boost_apply_static(FjCuda* cuda, int32_t device_id, const FjHighlightBoost* boost,
                                                    float* red, float* green, float* blue, int64_t samples,
                                                    uint64_t stream, FjErrorBuffer* error) {
    try {
        if (bad_error_buffer(error)) return make_status(FJ_STATUS_UNSUPPORTED_INPUT);
        NativeCall call(cuda);
        if (!call.ok()) return call.refuse(error);
        if (!boost || boost->recipe_hash == 0 || !highlight_boost_valid(*boost))
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "MalformedPreparedHostData field=optics.highlight_boost");
        if (!red || !green || !blue || samples < 0)
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "UnsupportedInput highlight boost planes");
        if (samples > kMaxSamples)
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT,
                        "UnsupportedInput highlight boost planes above 2^34 samples (static scratch entry point)");
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess) return hip_fail(error, e, "hipGetDeviceCount");
        if (device_id < 0 || device_id >= count || device_id >= kMaxDevices)
            return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "device id out of range");
        e = hipSetDevice(device_id);
        if (e != hipSuccess) return hip_fail(error, e, "hipSetDevice");
        unsigned char* base = nullptr;
        e = scratch_address(device_id, &base);
        if (e != hipSuccess) return hip_fail(error, e, "hipGetSymbolAddress(scratch)");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        HbArgs ha{};
        ha.planes[0] = red;
        ha.planes[1] = green;
        ha.planes[2] = blue;
        ha.n = samples;
        ha.protection = boost->protection;
        ha.sharpness = boost->sharpness;
        ha.peak_gain = boost->peak_gain;
        ha.state = reinterpret_cast<HbState*>(base);
        ha.error_word = reinterpret_cast<unsigned int*>(base + kStateBytes);
        // The array keeps the last call's contents: the error word starts at
        // zero (a kernel, not a memset, so the launch sequence is capturable);
        // the partials and state are written before they are read.
        e = launch_clear_word(ha.error_word, st);
        if (e == hipSuccess) e = launch_highlight_boost(ha, st);
        if (e != hipSuccess) return hip_fail(error, e, "highlight boost");
        unsigned int bits = 0u;
        e = hipMemcpyAsync(&bits, ha.error_word, sizeof(bits), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return hip_fail(error, e, "highlight boost status");
        if (bits & kHbInvalidCarrierBit) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "HighlightBoostInvalidCarrier");
        if (bits & kHbPeakOverflowBit) return fail(error, FJ_STATUS_UNSUPPORTED_INPUT, "HighlightBoostPeakOverflow");
        write_error(error, "");
        return make_status(FJ_STATUS_SUCCESS);
    } catch (...) {
        return fail(error, FJ_STATUS_INTERNAL_FAILURE, "unexpected exception in fjx_highlight_boost_apply_static");
    }
}

}  // extern "C"
