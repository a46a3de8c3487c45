// fj_grain.hip — visual-grain kernels for gfx950.
//
// Sample kernels: one thread per pixel; the development's sub-layer curves
// are staged into LDS (the interpolation's bounded search then hits LDS);
// STBN bytes are random gathers into one 2D slice of the volume per frame
// offset (L2-resident). Blur kernels: LDS-tiled half-sample-reflect FIRs
// accumulating in the reference's order. Sources: grain_apply_simple /
// grain_layer / grain_layer_triplet (src/Cuda/Film/JuicerCudaFilmPipeline.cu:
// 2579-2987), the sampler and noise helpers (:1602-2172), the blur and
// elementwise kernels of JuicerCudaScanPipeline.cu:315-790 and
// JuicerCudaFilmPipeline.cu:2456-2577.
#include <hip/hip_runtime.h>

#include "fj_batch.h"
#include "fj_detmath.h"
#include "fj_grain.h"

// The decision path (Poisson thresholds, rounding of normal samples, Wang
// cells) must not depend on how the compiler fuses multiply-adds.
#pragma clang fp contract(off)

namespace fjx {

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kMix1 = 0xBF58476D1CE4E5B9ull;

__device__ __forceinline__ bool fin(float v) { return isfinite(v); }
__device__ __forceinline__ float san(float v) { return isfinite(v) ? v : 0.0f; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return fminf(fmaxf(v, lo), hi); }
__device__ __forceinline__ float smooth(float t) { return t * t * (3.0f - 2.0f * t); }

__device__ __forceinline__ int wrap_coord(long long c, int size) {
    if ((size & (size - 1)) == 0) return static_cast<int>(static_cast<unsigned long long>(c) & (size - 1));
    c %= size;
    if (c < 0) c += size;
    return static_cast<int>(c);
}

__device__ __forceinline__ float spatial_hash01(int x, int y, uint64_t seed) {
    const uint32_t lo = static_cast<uint32_t>(seed), hi = static_cast<uint32_t>(seed >> 32);
    uint32_t h = lo ^ ((hi << 16) | (hi >> 16));
    h ^= static_cast<uint32_t>(x) * 0x9E3779B9u;
    h ^= static_cast<uint32_t>(y) * 0x85EBCA6Bu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return static_cast<float>(h) * (1.0f / 4294967296.0f);
}

__device__ __forceinline__ uint32_t hash_u32(int x, int y, uint64_t seed, uint64_t salt) {
    const uint64_t h = splitmix64(seed ^ salt ^ (static_cast<uint64_t>(static_cast<long long>(x)) * kGolden) ^
                                  (static_cast<uint64_t>(static_cast<long long>(y)) * kMix1));
    return static_cast<uint32_t>(h);
}

// True when every active lane holds the same (a, b); ua, ub are then that
// pair as wave-uniform values, so integer work on them runs on the scalar
// unit once per wave instead of once per lane (the bits are the same).
__device__ __forceinline__ bool wave_uniform(int a, int b, int& ua, int& ub) {
    ua = __builtin_amdgcn_readfirstlane(a);
    ub = __builtin_amdgcn_readfirstlane(b);
    return __all(a == ua && b == ub);
}

__device__ float value_noise(float x, float y, uint64_t seed) {
    const int ix = static_cast<int>(floorf(x)), iy = static_cast<int>(floorf(y));
    const float fx = x - static_cast<float>(ix), fy = y - static_cast<float>(iy);
    const float sx = smooth(fx), sy = smooth(fy);
    float v00, v10, v01, v11;
    int ux, uy;
    if (wave_uniform(ix, iy, ux, uy)) {  // the wave lies in one lattice cell
        v00 = spatial_hash01(ux, uy, seed);
        v10 = spatial_hash01(ux + 1, uy, seed);
        v01 = spatial_hash01(ux, uy + 1, seed);
        v11 = spatial_hash01(ux + 1, uy + 1, seed);
    } else {
        v00 = spatial_hash01(ix, iy, seed);
        v10 = spatial_hash01(ix + 1, iy, seed);
        v01 = spatial_hash01(ix, iy + 1, seed);
        v11 = spatial_hash01(ix + 1, iy + 1, seed);
    }
    const float v0 = v00 + (v10 - v00) * sx;
    const float v1 = v01 + (v11 - v01) * sx;
    return v0 + (v1 - v0) * sy;
}

__device__ float breathing_factor(const GrainDev& g, uint64_t ax, uint64_t ay) {
    if (!g.breathing_active) return 1.0f;
    const float bx = static_cast<float>(ax) + g.breathing_drift_x;
    const float by = static_cast<float>(ay) + g.breathing_drift_y + g.breathing_roll_y;
    const float xs = bx / g.breathing_cell_small_px, ys = by / g.breathing_cell_small_px;
    const float xl = bx / g.breathing_cell_large_px, yl = by / g.breathing_cell_large_px;
    const float sa = value_noise(xs, ys, g.breathing_seed_a ^ kGolden);
    const float sb = value_noise(xs, ys, g.breathing_seed_b ^ kGolden);
    const float la = value_noise(xl, yl, g.breathing_seed_a ^ kMix1);
    const float lb = value_noise(xl, yl, g.breathing_seed_b ^ kMix1);
    const float ns = sa + (sb - sa) * g.breathing_t;
    const float nl = la + (lb - la) * g.breathing_t;
    const float noise = g.breathing_mix * ns + (1.0f - g.breathing_mix) * nl;
    const float n = noise * 2.0f - 1.0f;
    const float f = 1.0f + g.breathing_amp * n;
    return f > 0.0f ? f : 0.0f;
}

__device__ float clump_factor(const GrainDev& g, uint64_t ax, uint64_t ay) {
    if (!g.clump_active) return 1.0f;
    const float xs = static_cast<float>(ax) / g.clump_cell_px;
    const float ys = (static_cast<float>(ay) + g.clump_roll_y) / g.clump_cell_px;
    const float u1 = clampf(value_noise(xs, ys, g.clump_seed ^ kGolden), 1e-6f, 1.0f - 1e-6f);
    const float u2 = clampf(value_noise(xs + 19.19f, ys + 7.23f, g.clump_seed ^ kMix1), 0.0f, 1.0f);
    const float r = sqrtf(-2.0f * det_logf(u1));
    const float n = r * det_cosf(6.28318530717958647692f * u2);
    float v = det_expf(g.clump_mu + g.clump_sigma * n);
    if (!fin(v)) v = 1.0f;
    return v * g.clump_rms_norm;
}

__device__ __forceinline__ float wang_byte(const GrainDev& g, int tile, int x, int y) {
    if (tile < 0 || tile >= g.wang_count) return 0.0f;
    x = min(max(x, 0), g.wang_w - 1);
    y = min(max(y, 0), g.wang_h - 1);
    const uint8_t v = g.wang_tiles[(static_cast<size_t>(tile) * g.wang_h + y) * g.wang_w + x];
    return (static_cast<float>(v) + 0.5f) * (1.0f / 256.0f);
}

__device__ void wang_offsets(const GrainDev& g, uint64_t ax, uint64_t ay, int& ox, int& oy) {
    ox = oy = 0;
    if (!(g.pixel_size_um > 0.0f) || !(g.wang_cell_mm > 0.0f)) return;
    const float xmm = static_cast<float>(ax) * (g.pixel_size_um * 0.001f);
    const float ymm = static_cast<float>(ay) * (g.pixel_size_um * 0.001f);
    const float inv = 1.0f / g.wang_cell_mm;
    const float cx = xmm * inv, cy = ymm * inv;
    const float fx = cx - floorf(cx), fy = cy - floorf(cy);
    const long long mx = static_cast<long long>(floorf(cx)), my = static_cast<long long>(floorf(cy));
    const uint64_t s = g.clip_token ? g.clip_token : g.session_seed;
    const uint64_t base = s ? s : 1ull;
    const uint32_t c = static_cast<uint32_t>(g.wang_colors);
    auto cell_tile = [&](int cx, int cy) {
        const uint32_t l = hash_u32(cx, cy, base, 0xA5A5A5A5u) % c;
        const uint32_t r = hash_u32(cx + 1, cy, base, 0x5A5A5A5Au) % c;
        const uint32_t t = hash_u32(cx, cy, base, 0xC3C3C3C3u) % c;
        const uint32_t b = hash_u32(cx, cy + 1, base, 0x3C3C3C3Cu) % c;
        return static_cast<int>(g.wang_lut[((static_cast<size_t>(l) * c + r) * c + t) * c + b]);
    };
    int ux, uy;
    const int tile = wave_uniform(static_cast<int>(mx), static_cast<int>(my), ux, uy)
                         ? cell_tile(ux, uy)
                         : cell_tile(static_cast<int>(mx), static_cast<int>(my));
    const int w = g.wang_w, h = g.wang_h;
    int tx = static_cast<int>(floorf(fx * static_cast<float>(w)));
    int ty = static_cast<int>(floorf(fy * static_cast<float>(h)));
    if (tx >= w) tx = w - 1;
    if (ty >= h) ty = h - 1;
    const float v0 = wang_byte(g, tile, tx, ty);
    const float v1 = wang_byte(g, tile, (tx + (w >> 1)) % w, (ty + (h >> 1)) % h);
    const int warp_x = (g.stbn_w / 8) > 4 ? g.stbn_w / 8 : 4;
    const int warp_y = (g.stbn_h / 8) > 4 ? g.stbn_h / 8 : 4;
    ox = static_cast<int>(floorf((v0 - 0.5f) * 2.0f * static_cast<float>(warp_x)));
    oy = static_cast<int>(floorf((v1 - 0.5f) * 2.0f * static_cast<float>(warp_y)));
}

// One pixel's uniform stream over the STBN volume (GrainRngDevice,
// JuicerCudaFilmPipeline.cu:1892-1972).
struct Stream {
    const uint8_t* slice;   // the frame's 2D slice
    uint64_t seed;
    long long x0, y0;       // absolute pixel + warp + session offsets + pitch rows
    uint32_t draw;
    int w, h, mask;

    __device__ float uniform() {
        const uint64_t hh = splitmix64(seed + static_cast<uint64_t>(draw++) * kGolden);
        const long long xc = x0 + static_cast<long long>(hh & 0xFFFFu);
        const long long yc = y0 + static_cast<long long>((hh >> 16) & 0xFFFFu);
        int x, y;
        if (mask) {
            x = static_cast<int>(static_cast<unsigned long long>(xc) & (w - 1));
            y = static_cast<int>(static_cast<unsigned long long>(yc) & (h - 1));
        } else {
            x = wrap_coord(xc, w);
            y = wrap_coord(yc, h);
        }
        return (static_cast<float>(slice[static_cast<size_t>(y) * w + x]) + 0.5f) * (1.0f / 256.0f);
    }
    __device__ float normal() {
        float u1 = uniform();
        u1 = fminf(fmaxf(u1, 1e-7f), 1.0f);
        const float u2 = uniform();
        const float r = sqrtf(-2.0f * det_logf(u1));
        return r * det_cosf(6.28318530717958647692f * u2);
    }
};

__device__ __forceinline__ Stream make_stream(const GrainDev& g, uint64_t seed, uint64_t ax, uint64_t ay,
                                              int frame_offset, int wx, int wy) {
    Stream s;
    const long long pitch = g.pitch_px > 0 ? static_cast<long long>(g.pitch_px) * (g.frame_index + frame_offset) : 0;
    const int frame = wrap_coord(static_cast<long long>(g.stbn_frame) + frame_offset, g.stbn_frames);
    s.slice = g.stbn + static_cast<size_t>(frame) * g.stbn_h * g.stbn_w;
    s.seed = seed;
    s.x0 = static_cast<long long>(ax) + wx + g.stbn_off_x;
    s.y0 = static_cast<long long>(ay) + wy + g.stbn_off_y + pitch;
    s.draw = 0;
    s.w = g.stbn_w;
    s.h = g.stbn_h;
    s.mask = g.stbn_mask;
    return s;
}

__device__ int poisson(float lambda, Stream& s) {
    if (!fin(lambda) || !(lambda > 0.0f)) return 0;
    if (lambda < 10.0f) {
        // Knuth's product loop (fast_poisson_device,
        // JuicerCudaFilmPipeline.cu:1997-2025). A draw's hash depends only on
        // the seed and the draw index, the same for every pixel of the wave,
        // so four draws' STBN gathers are issued together and consumed in
        // order; a pixel stops at the same draw as the one-at-a-time loop.
        const float L = det_expf(-lambda);
        float p = 1.0f;
        int k = 0;
        bool live = p > L;
        while (live) {
            float u[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) u[j] = s.uniform();
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                if (live) {
                    ++k;
                    p *= fminf(fmaxf(u[j], 1e-7f), 1.0f);
                    live = p > L && k < 1024;
                }
            }
        }
        return k - 1;
    }
    const float z = s.normal();
    float sample = lambda + sqrtf(lambda) * z;
    if (!fin(sample)) sample = 0.0f;
    const int n = static_cast<int>(floorf(sample + 0.5f));
    return n < 0 ? 0 : n;
}

__device__ float particle(const GrainDev& g, float density, float dmax, float n_particles, float od, float uniformity,
                          uint64_t seed, uint64_t ax, uint64_t ay, int frame_offset, int wx, int wy) {
    if (!fin(density) || density < 0.0f) density = 0.0f;
    float p = density / dmax;
    p = fminf(fmaxf(p, 1e-6f), 1.0f - 1e-6f);
    const float u = fin(uniformity) ? fminf(fmaxf(uniformity, 0.0f), 1.0f) : 0.0f;
    float sat = 1.0f - p * u * (1.0f - 1e-6f);
    if (!fin(sat) || sat <= 0.0f) sat = 1e-6f;
    Stream s = make_stream(g, seed, ax, ay, frame_offset, wx, wy);
    const float lambda = n_particles / sat;
    const int count = poisson(lambda * p, s);
    return san(static_cast<float>(count) * od * sat);
}

// Layer density from the total density through the development curves
// (interp_density_layer_device, JuicerCudaFilmPipeline.cu:2077-2172); x and y
// are LDS copies.
__device__ float interp_layer(float density, const float* x, const float* y, int n, bool positive, bool bounded,
                              const float* prefix) {
    if (!fin(density)) return y[0];
    const float q = positive ? -density : density;
    auto ax = [&](int i) { return positive ? -x[i] : x[i]; };
    int begin = 0, end = n - 1;
    if (!bounded) {
        while (begin < n && !fin(ax(begin))) ++begin;
        if (begin >= n) return density;
        while (end > begin && !fin(ax(end))) --end;
    }
    const float xmin = ax(begin), xmax = ax(end);
    if (!fin(xmin) || !fin(xmax) || !(xmax >= xmin)) return y[begin];
    if (q <= xmin) return y[begin];
    if (q >= xmax) return y[end];
    int i1 = begin + 1;
    if (bounded) {
        int bb = 0, be = 15;
#pragma unroll
        for (int step = 0; step < 4; ++step) {
            const int mid = bb + (be - bb) / 2;
            if (prefix[mid] < q)
                bb = mid + 1;
            else
                be = mid;
        }
        // First i in [start, search_end] with !(ax(i) < q), else
        // search_end + 1: the 16 entries of block bb are read at once and
        // compared in parallel (x is 16-byte aligned in LDS).
        const int base = bb * 16;
        const int start = max(1, base);
        const int search_end = min(end, base + 15);
        const float4* blk = reinterpret_cast<const float4*>(x + base);
        unsigned hit = 0;
#pragma unroll
        for (int qd = 0; qd < 4; ++qd) {
            const float4 v4 = blk[qd];
            const float e4[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int idx = base + qd * 4 + t;
                const float av = positive ? -e4[t] : e4[t];
                if (idx >= start && idx <= search_end && !(av < q)) hit |= 1u << (qd * 4 + t);
            }
        }
        i1 = hit ? base + __builtin_ctz(hit) : max(start, search_end + 1);
    } else {
        while (i1 <= end && ax(i1) < q) ++i1;
    }
    if (i1 > end) return y[end];
    const int i0 = i1 - 1;
    const float x0 = ax(i0), x1 = ax(i1), y0 = y[i0], y1 = y[i1];
    const float den = x1 - x0;
    if (!(den > 0.0f) || !fin(den)) return y0;
    const float t = (q - x0) / den;
    return y0 + t * (y1 - y0);
}

__device__ __forceinline__ float retime(const GrainDev& g, float density, float s0, float s1, float bf, float cf,
                                        bool want_next) {
    float a = density + (s0 - density) * bf * cf;
    if (want_next) {
        const float b = density + (s1 - density) * bf * cf;
        const float d0 = a - density, d1 = b - density;
        const float blend = d0 + (d1 - d0) * g.time_alpha;
        const float w0 = 1.0f - g.time_alpha;
        const float den = w0 * w0 + g.time_alpha * g.time_alpha + 1e-6f;
        a = density + blend * (1.0f / sqrtf(den));
    }
    return a;
}

}  // namespace

template <int kMode>
__global__ __launch_bounds__(256) void grain_sample_kernel(GrainSampleArgs a) {
    __shared__ __align__(16) float s_axis[kGrainCurveMax];
    __shared__ float s_layer[3][kGrainCurveMax];
    __shared__ float s_prefix[16];
    const GrainDev& g = a.g;
    const int c = a.channel;
    const int n = g.curve_n;
    if constexpr (kMode != kGrainSimple) {
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            s_axis[i] = g.base_curve[c][i];
            for (int l = 0; l < 3; ++l) s_layer[l][i] = (kMode == kGrainTriplet || l == a.layer) ? g.layer_curve[l][c][i] : 0.0f;
        }
        if (threadIdx.x < 16) s_prefix[threadIdx.x] = g.axis_prefix[c][threadIdx.x];
        __syncthreads();
    }
    const long long total = static_cast<long long>(a.width) * a.height;
    const long long idx = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= total) return;
    const int y = static_cast<int>(idx / a.width);
    const int x = static_cast<int>(idx - static_cast<long long>(y) * a.width);
    const uint64_t ax = static_cast<uint64_t>(static_cast<long long>(g.origin_x) + x);
    const uint64_t ay = static_cast<uint64_t>(static_cast<long long>(g.origin_y) + y);
    int wx = 0, wy = 0;
    wang_offsets(g, ax, ay, wx, wy);
    const float cf = clump_factor(g, ax, ay);
    const float bf = breathing_factor(g, ax, ay);
    const bool want_next = g.time_alpha > 1e-6f && g.time_alpha < 0.999999f;
    const float in_density = a.density[idx];
    float value;
    if constexpr (kMode == kGrainSimple) {
        const float density = in_density + g.density_min[c];
        const int nsub = g.n_sublayers > 0 ? g.n_sublayers : 1;
        float acc = 0.0f, acc_next = 0.0f;
        for (int sl = 0; sl < nsub; ++sl) {
            const uint64_t salt = static_cast<uint64_t>(c) + static_cast<uint64_t>(sl) * 10ull;
            acc += particle(g, density, g.density_max[c], g.n_particles[c], g.od_particle[c], g.uniformity[c],
                            g.seed_base ^ salt, ax, ay, 0, wx, wy);
            if (want_next)
                acc_next += particle(g, density, g.density_max[c], g.n_particles[c], g.od_particle[c], g.uniformity[c],
                                     g.seed_base_next ^ salt, ax, ay, 1, wx, wy);
        }
        acc /= static_cast<float>(nsub);
        if (want_next) acc_next /= static_cast<float>(nsub);
        value = san(retime(g, density, acc, acc_next, bf, cf, want_next));
    } else {
        const bool bounded = g.axis_finite[c] && n == 256;
        float sum = 0.0f;
        const int l0 = kMode == kGrainTriplet ? 0 : a.layer;
        const int l1 = kMode == kGrainTriplet ? 3 : a.layer + 1;
        for (int l = l0; l < l1; ++l) {
            float density = interp_layer(in_density, s_axis, s_layer[l], n, g.positive_film != 0, bounded, s_prefix);
            density += g.dmin_l[l][c];
            const uint64_t salt = static_cast<uint64_t>(c) + static_cast<uint64_t>(l) * 10ull;
            const float s0 = particle(g, density, g.dmax_l[l][c], g.n_l[l][c], g.od_l[l][c], g.uniformity[c],
                                      g.seed_base ^ salt, ax, ay, 0, wx, wy);
            const float s1 = want_next ? particle(g, density, g.dmax_l[l][c], g.n_l[l][c], g.od_l[l][c],
                                                  g.uniformity[c], g.seed_base_next ^ salt, ax, ay, 1, wx, wy)
                                       : 0.0f;
            sum += san(retime(g, density, s0, s1, bf, cf, want_next));
        }
        value = san(sum);
    }
    if (a.form_delta) {
        // accumulate (san(0 + v)), then grain_form_delta_kernel
        const float acc = san(0.0f + value);
        const float biased = san(acc + a.bias);
        value = san((biased - in_density) * 1.0f);
    }
    a.out[idx] = value;
}

// Horizontal pass: 256 outputs of one row per block plus the halo in LDS.
__global__ __launch_bounds__(256) void grain_blur_h_kernel(GrainBlurArgs a) {
    extern __shared__ float lds[];
    const int r = a.radius, taps = 2 * r + 1;
    float* w = lds;
    float* tile = lds + taps;
    for (int i = threadIdx.x; i < taps; i += blockDim.x) w[i] = a.weights[i];
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * 256;
    const float* row = a.in + static_cast<long long>(y) * a.width;
    for (int i = threadIdx.x; i < 256 + 2 * r; i += blockDim.x) {
        long long xi = static_cast<long long>(x0) + i - r;
        // half-sample mirror, repeated for halos wider than the row
        while (xi < 0 || xi >= a.width) xi = xi < 0 ? -xi - 1 : 2LL * a.width - xi - 1;
        tile[i] = a.width > 1 ? row[xi] : row[0];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= a.width) return;
    float acc = 0.0f;
    for (int j = 0; j < taps; ++j) acc = fmaf(tile[threadIdx.x + j], w[j], acc);
    a.out[static_cast<long long>(y) * a.width + x] = san(acc);
}

// Vertical pass with the accumulate / delta epilogues: 64 columns x 16 rows
// per block, a (16 + 2r) x 64 tile in LDS.
__global__ __launch_bounds__(256) void grain_blur_v_kernel(GrainBlurArgs a) {
    extern __shared__ float lds[];
    const int r = a.radius, taps = 2 * r + 1;
    float* w = lds;
    float* tile = lds + taps;
    for (int i = threadIdx.x; i < taps; i += blockDim.x) w[i] = a.weights[i];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    const int x = blockIdx.x * 64 + tx;
    const int y0 = blockIdx.y * 16;
    const int rows = 16 + 2 * r;
    batched<8, float>(
        ty, rows, 4,
        [&](int i) {
            long long yi = static_cast<long long>(y0) + i - r;
            while (yi < 0 || yi >= a.height) yi = yi < 0 ? -yi - 1 : 2LL * a.height - yi - 1;
            if (a.height <= 1) yi = 0;
            return x < a.width ? a.in[yi * a.width + x] : 0.0f;
        },
        [&](int i, float v) { tile[i * 64 + tx] = v; });
    __syncthreads();
    if (x >= a.width) return;
    const float wf = isfinite(a.weight) ? a.weight : 0.0f;
    for (int yy = ty; yy < 16; yy += 4) {
        const int y = y0 + yy;
        if (y >= a.height) break;
        float acc = 0.0f;
        for (int j = 0; j < taps; ++j) acc = fmaf(tile[(yy + j) * 64 + tx], w[j], acc);
        const float blurred = san(acc);
        const long long i = static_cast<long long>(y) * a.width + x;
        if (a.unsharp) {
            // Scanner unsharp mask (optics_unsharp_vertical_combine_kernel,
            // JuicerCudaScanPipeline.cu:586-660).
            const float v0 = a.out[i];
            a.out[i] = san(fmaf(a.weight, v0 - blurred, v0));
            continue;
        }
        const float prev = a.initialize ? 0.0f : a.out[i];
        float v = a.weighted ? prev + wf * blurred : prev + blurred;
        v = san(v);
        if (!a.weighted && a.sub) {
            if (a.bias != 0.0f) v = san(v + a.bias);
            v = san(v - a.sub[i]);
        }
        if (a.gain != 1.0f) v = san(v * a.gain);
        a.out[i] = v;
    }
}

__global__ __launch_bounds__(256) void grain_elt_kernel(GrainEltArgs a) {
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        switch (a.op) {
            case kEltClear: a.a[i] = 0.0f; break;
            case kEltAccumulate: a.a[i] = san((a.init ? 0.0f : a.a[i]) + a.b[i]); break;
            case kEltAccWeighted: {
                const float wv = isfinite(a.w) ? a.w : 0.0f;
                a.a[i] = san((a.init ? 0.0f : a.a[i]) + wv * a.b[i]);
                break;
            }
            case kEltFormDelta: {
                const float biased = san(a.a[i] + a.bias);
                a.a[i] = san((biased - a.b[i]) * 1.0f);
                break;
            }
            case kEltSubtract: a.a[i] = san((a.a[i] - a.b[i]) * 1.0f); break;
            case kEltScale: {
                const float s = isfinite(a.s) ? a.s : 0.0f;
                a.a[i] = san(a.a[i] * s);
                break;
            }
            case kEltMixShared: {
                const float sw = clampf(a.sw, 0.0f, 1.0f), iw = clampf(a.iw, 0.0f, 1.0f);
                const float shared = (a.shared && sw > 0.0f) ? a.shared[i] : 0.0f;
                const float ind = (a.use_ind && iw > 0.0f) ? a.a[i] : 0.0f;
                a.a[i] = san(a.amp * (sw * shared + iw * ind));
                break;
            }
            case kEltReconstruct: {
                const float s = isfinite(a.s) ? a.s : 0.0f;
                a.a[i] = san((a.add_mean ? a.a[i] : 0.0f) + s * a.b[i]);
                break;
            }
            case kEltMixReconstruct: {
                const float sw = clampf(a.sw, 0.0f, 1.0f), iw = clampf(a.iw, 0.0f, 1.0f);
                const float shared = (a.shared && sw > 0.0f) ? a.shared[i] : 0.0f;
                const float ind = (a.use_ind && iw > 0.0f) ? a.b[i] : 0.0f;
                const float mixed = san(a.amp * (sw * shared + iw * ind));
                const float s = isfinite(a.s) ? a.s : 0.0f;
                a.a[i] = san((a.add_mean ? a.a[i] : 0.0f) + s * mixed);
                break;
            }
            default: a.a[i] = a.b[i]; break;
        }
    }
}

__global__ __launch_bounds__(256) void grain_avg3_kernel(GrainAvgArgs a) {
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    const float s = isfinite(a.scale) ? a.scale : 1.0f, o = isfinite(a.offset) ? a.offset : 0.0f;
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < a.n; i += stride) {
        float m = (a.d[0][i] + a.d[1][i] + a.d[2][i]) * (1.0f / 3.0f);
        if (!isfinite(m)) m = 0.0f;
        const float v = o + m * s;
        a.d[0][i] = v;
        a.d[1][i] = v;
        a.d[2][i] = v;
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_grain_samples(const GrainSampleArgs& a, int mode, hipStream_t s) {
    const long long total = static_cast<long long>(a.width) * a.height;
    if (total <= 0) return hipSuccess;
    if (mode != kGrainSimple && (a.g.curve_n <= 0 || a.g.curve_n > kGrainCurveMax)) return hipErrorInvalidValue;
    const long long blocks = (total + 255) / 256;
    if (blocks > 0x7fffffffLL) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(blocks));
    switch (mode) {
        case kGrainSimple: hipLaunchKernelGGL(grain_sample_kernel<kGrainSimple>, grid, dim3(256), 0, s, a); break;
        case kGrainTriplet: hipLaunchKernelGGL(grain_sample_kernel<kGrainTriplet>, grid, dim3(256), 0, s, a); break;
        default: hipLaunchKernelGGL(grain_sample_kernel<kGrainLayer>, grid, dim3(256), 0, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_grain_blur_h(const GrainBlurArgs& a, hipStream_t s) {
    if (a.width <= 0 || a.height <= 0 || a.radius <= 0) return hipSuccess;
    const size_t lds = static_cast<size_t>(2 * a.radius + 1 + 256 + 2 * a.radius) * sizeof(float);
    const dim3 grid((a.width + 255) / 256, a.height);
    hipLaunchKernelGGL(grain_blur_h_kernel, grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_grain_blur_v(const GrainBlurArgs& a, hipStream_t s) {
    if (a.width <= 0 || a.height <= 0 || a.radius <= 0) return hipSuccess;
    const size_t lds = static_cast<size_t>(2 * a.radius + 1 + (16 + 2 * a.radius) * 64) * sizeof(float);
    const dim3 grid((a.width + 63) / 64, (a.height + 15) / 16);
    hipLaunchKernelGGL(grain_blur_v_kernel, grid, dim3(256), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_grain_elt(const GrainEltArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const long long blocks = (a.n + 255) / 256;
    const int grid = static_cast<int>(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(grain_elt_kernel, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_grain_avg3(const GrainAvgArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    const long long blocks = (a.n + 255) / 256;
    const int grid = static_cast<int>(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(grain_avg3_kernel, dim3(grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace fjx
