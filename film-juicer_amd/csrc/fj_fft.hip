// fj_fft.hip — LDS Stockham FFT and the DCT-domain spatial filters built on it.
//
// Every line transform runs inside one workgroup: the line (a pair of real
// lines packed as re/im of one complex line) is staged into LDS, transformed
// with a self-sorting mixed-radix Stockham FFT (radix 4, 2, 3, 5 butterflies
// in registers, any other prime factor through a generic LDS butterfly), and
// written back. Twiddles come from a per-length table built on the host in
// double precision.
//
// DCT-II / DCT-III of length N use Makhoul's reordering onto one length-N
// complex FFT. The DCT-II basis diagonalizes convolution of the half-sample
// symmetric periodic extension of a line (the reference's reflect boundary,
// JuicerCudaFilmPipeline.cu:82-92) with any symmetric kernel, so a filter is
// applied as IDCT(T ⊙ DCT(x)) with T the kernel's frequency response.
#include <hip/hip_runtime.h>

#include "fj_fft.h"

namespace fjx {

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) : (x, y) -> (y, -x)
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }

// Twiddle w_N^e = exp(-2*pi*i*e/N) for 0 <= e < N. Every caller's exponent
// is below N: r * k * (N / (ns R)) < R ns N / (ns R) = N for r < R, k < ns.
__device__ __forceinline__ float2 tw(const float2* table, int e) { return table[e]; }

// One Stockham pass of radix R over the whole line: reads src, writes dst.
// Ns = product of radices already applied. Forward transform (e^{-i}).
// `stw` holds this pass's twiddles, R-1 per butterfly index k < ns, and
// `magic` = ceil(2^32 / ns) (exact division for j ns < 2^32).
template <int R>
__device__ __forceinline__ void pass_fixed(const float2* src, float2* dst, int n, int ns, const float2* stw,
                                           unsigned magic) {
    const int count = n / R;
    for (int j = threadIdx.x; j < count; j += blockDim.x) {
        const int q = ns > 1 ? static_cast<int>(__umulhi(static_cast<unsigned>(j), magic)) : j;
        const int k = j - q * ns;
        float2 a[R];
#pragma unroll
        for (int r = 0; r < R; ++r) a[r] = src[j + r * count];
        if (ns > 1) {
            const float2* t = stw + k * (R - 1);
#pragma unroll
            for (int r = 1; r < R; ++r) a[r] = cmul(a[r], t[r - 1]);
        }
        float2 b[R];
        if constexpr (R == 2) {
            b[0] = cadd(a[0], a[1]);
            b[1] = csub(a[0], a[1]);
        } else if constexpr (R == 4) {
            const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
            const float2 s13 = cadd(a[1], a[3]), d13 = mul_mi(csub(a[1], a[3]));
            b[0] = cadd(s02, s13);
            b[2] = csub(s02, s13);
            b[1] = cadd(d02, d13);
            b[3] = csub(d02, d13);
        } else if constexpr (R == 3) {
            const float c = -0.5f, s = -0.86602540378443864676f;  // cos/sin(-2pi/3)
            const float2 t = cadd(a[1], a[2]);
            const float2 u = csub(a[1], a[2]);
            b[0] = cadd(a[0], t);
            const float2 m = make_float2(fmaf(c, t.x, a[0].x), fmaf(c, t.y, a[0].y));
            const float2 v = make_float2(-s * u.y, s * u.x);  // i*s*u
            b[1] = cadd(m, v);
            b[2] = csub(m, v);
        } else if constexpr (R == 5) {
            const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
            const float s1 = -0.95105651629515357212f, s2 = -0.58778525229247312917f;  // sin(-2pi/5), sin(-4pi/5)
            const float2 t1 = cadd(a[1], a[4]), u1 = csub(a[1], a[4]);
            const float2 t2 = cadd(a[2], a[3]), u2 = csub(a[2], a[3]);
            b[0] = cadd(a[0], cadd(t1, t2));
            const float2 m1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
            const float2 m2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
            // i*(s1*u1 + s2*u2) and i*(s2*u1 - s1*u2)
            const float2 v1 = make_float2(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
            const float2 v2 = make_float2(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
            b[1] = cadd(m1, v1);
            b[4] = csub(m1, v1);
            b[2] = cadd(m2, v2);
            b[3] = csub(m2, v2);
        }
        const int base = q * ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[base + r * ns] = b[r];
    }
}

// Generic prime radix: each output is a length-R DFT evaluated directly.
__device__ void pass_generic(const float2* src, float2* dst, int n, int ns, int R, const float2* table) {
    const int count = n / R;
    const int stride_tw = n / (ns * R);
    const int root = n / R;  // w_R = w_N^(N/R)
    for (int idx = threadIdx.x; idx < count * R; idx += blockDim.x) {
        const int j = idx / R;
        const int q = idx - j * R;  // output index within the butterfly
        const int k = j % ns;
        float2 acc = make_float2(0.0f, 0.0f);
        for (int r = 0; r < R; ++r) {
            float2 a = src[j + r * count];
            if (ns > 1 && r > 0) a = cmul(a, tw(table, r * k * stride_tw));
            acc = cadd(acc, cmul(a, tw(table, ((r * q) % R) * root)));
        }
        dst[(j / ns) * ns * R + k + q * ns] = acc;
    }
}

// Forward FFT of buf[0..n) in LDS; returns the buffer holding the result.
__device__ float2* fft_lds(float2* a, float2* b, const FftPlanDev& plan, const float2* table) {
    int ns = 1;
    float2* src = a;
    float2* dst = b;
    for (int p = 0; p < plan.passes; ++p) {
        const int r = plan.radix[p];
        const float2* stw = table + plan.n + plan.tw_offset[p];
        const unsigned magic = plan.ns_magic[p];
        switch (r) {
            case 4: pass_fixed<4>(src, dst, plan.n, ns, stw, magic); break;
            case 2: pass_fixed<2>(src, dst, plan.n, ns, stw, magic); break;
            case 3: pass_fixed<3>(src, dst, plan.n, ns, stw, magic); break;
            case 5: pass_fixed<5>(src, dst, plan.n, ns, stw, magic); break;
            default: pass_generic(src, dst, plan.n, ns, r, table); break;
        }
        __syncthreads();
        ns *= r;
        float2* t = src;
        src = dst;
        dst = t;
    }
    return src;
}

// Whole-sample mirror with period 2 (n - 1), n >= 2 (reflect-101).
__device__ __forceinline__ int reflect101(int i, int n) {
    if (i >= 0 && i < n) return i;
    const int period = 2 * (n - 1);
    int v = i % period;
    if (v < 0) v += period;
    return v < n ? v : period - v;
}

// Makhoul permutation index: position of x[i] in v.
__device__ __forceinline__ int makhoul_pos(int i, int n) { return (i & 1) ? n - 1 - (i >> 1) : (i >> 1); }

// DCT-II of two real lines packed as (re, im) of one complex line.
// On return out_a / out_b (global) hold X_a[k], X_b[k].
__device__ void dct2_pair(float2* a, float2* b, const FftPlanDev& plan, const float2* table, const float2* half_tw,
                          float* out_a, float* out_b, float scale_a, float scale_b) {
    const int n = plan.n;
    float2* V = fft_lds(a, b, plan, table);
    // V = Va + i Vb with Va, Vb spectra of real sequences.
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const float2 z = V[k];
        const float2 zc = V[k == 0 ? 0 : n - k];
        const float2 va = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        const float2 vb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
        const float2 w = half_tw[k];  // exp(-i*pi*k/(2n))
        if (out_a) out_a[k] = (w.x * va.x - w.y * va.y) * scale_a;
        if (out_b) out_b[k] = (w.x * vb.x - w.y * vb.y) * scale_b;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Rows [2*blockIdx.x, 2*blockIdx.x+1] of each plane (blockIdx.y) -> DCT-II rows.
__global__ __launch_bounds__(kMaxRowThreads) void dct2_rows_kernel(DctRowsArgs args) {
    extern __shared__ float2 lds2[];
    const FftPlanDev& plan = args.plan;
    const int n = plan.n;
    float2* a = lds2;
    float2* b = lds2 + n;
    const int row0 = 2 * blockIdx.x;
    const int plane = blockIdx.y;
    if (row0 >= args.rows) return;
    const bool has1 = row0 + 1 < args.rows;
    const float* in = args.in[plane];
    float* out = args.out[plane];
    if (args.pad >= 0) {
        const float* r0 = in + static_cast<long long>(reflect101(row0 - args.pad, args.src_h)) * args.in_stride;
        const float* r1 = in + static_cast<long long>(reflect101(row0 + 1 - args.pad, args.src_h)) * args.in_stride;
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            const int si = reflect101(i - args.pad, args.src_w);
            a[makhoul_pos(i, n)] = make_float2(r0[si], has1 ? r1[si] : 0.0f);
        }
        __syncthreads();
        dct2_pair(a, b, plan, args.table, args.half_tw, out + static_cast<long long>(row0) * args.out_stride,
                  has1 ? out + static_cast<long long>(row0 + 1) * args.out_stride : nullptr, 1.0f, 1.0f);
        return;
    }
    const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
    const float* r1 = has1 ? in + static_cast<long long>(row0 + 1) * args.in_stride : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float v0 = r0[i];
        const float v1 = has1 ? r1[i] : 0.0f;
        a[makhoul_pos(i, n)] = make_float2(v0, v1);
    }
    __syncthreads();
    dct2_pair(a, b, plan, args.table, args.half_tw, out + static_cast<long long>(row0) * args.out_stride,
              has1 ? out + static_cast<long long>(row0 + 1) * args.out_stride : nullptr, 1.0f, 1.0f);
}

// Inverse (DCT-III scaled so that it undoes dct2) of row pairs, optionally
// multiplied by a separable-sum transfer function first:
//   T(k, row) = sum_c w_c * tk[c][k] * trow[c][row]
__global__ __launch_bounds__(kMaxRowThreads) void dct3_rows_kernel(DctRowsArgs args) {
    extern __shared__ float2 lds2[];
    const FftPlanDev& plan = args.plan;
    const int n = plan.n;
    float2* a = lds2;
    float2* b = lds2 + n;
    const int row0 = args.row_begin + 2 * blockIdx.x;
    const int plane = blockIdx.y;
    if (row0 >= args.rows) return;
    const bool has1 = row0 + 1 < args.rows;
    const float* in = args.in[plane];
    float* out = args.out[plane];
    const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
    const float* r1 = has1 ? in + static_cast<long long>(row0 + 1) * args.in_stride : r0;
    // V_a[k] = e^{i pi k/2n} (X_a[k] - i X_a[n-k]); pack V = V_a + i V_b.
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int km = k == 0 ? -1 : n - k;
        const float xa = r0[k], xam = km < 0 ? 0.0f : r0[km];
        const float xb = has1 ? r1[k] : 0.0f, xbm = (has1 && km >= 0) ? r1[km] : 0.0f;
        const float2 w = args.half_tw[k];  // e^{-i pi k/2n}; use conj
        const float2 wc = make_float2(w.x, -w.y);
        const float2 va = cmul(wc, make_float2(xa, -xam));
        const float2 vb = cmul(wc, make_float2(xb, -xbm));
        // conj(V) for an inverse via the forward FFT: conj(Va + i Vb)
        const float2 v = make_float2(va.x - vb.y, va.y + vb.x);
        a[k] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    float2* res = fft_lds(a, b, plan, args.table);
    const float inv = 1.0f / static_cast<float>(n);
    if (args.pad >= 0) {
        // Rows [pad, pad + src_h) of the padded plane hold the plane's rows.
        const int pad = args.pad;
        const float f = args.fraction, keep = 1.0f - args.fraction;
        const long long y0 = row0 - pad;
        const float* m = args.mix[plane];
        for (int i = threadIdx.x; i < args.src_w; i += blockDim.x) {
            const float2 z = res[makhoul_pos(i + pad, n)];
            const long long o = y0 * args.out_stride + i;
            out[o] = keep * m[o] + f * (z.x * inv);
            if (has1) out[o + args.out_stride] = keep * m[o + args.out_stride] + f * (-z.y * inv);
        }
        return;
    }
    float* o0 = out + static_cast<long long>(row0) * args.out_stride;
    float* o1 = has1 ? out + static_cast<long long>(row0 + 1) * args.out_stride : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float2 z = res[makhoul_pos(i, n)];
        // result = conj(z) / n = (va_time + i vb_time)
        o0[i] = z.x * inv;
        if (o1) o1[i] = -z.y * inv;
    }
}

// Forward DCT-II along rows, multiply by T, inverse along rows, in one pass.
// Used on the transposed plane (rows = original columns): the row index is the
// first-axis frequency kx, the in-row index the second-axis frequency.
// kFullTable: T is a full per-plane table (diffusion); otherwise a weighted sum
// of separable component factors (spatial DIR).
template <bool kFullTable>
__global__ __launch_bounds__(kMaxRowThreads) void dct_filter_rows_kernel(DctRowsArgs args, DctTransfer tf) {
    extern __shared__ float2 lds2[];
    const FftPlanDev& plan = args.plan;
    const int n = plan.n;
    float2* a = lds2;
    float2* b = lds2 + n;
    const int row0 = 2 * blockIdx.x;
    const int plane = blockIdx.y;
    if (row0 >= args.rows) return;
    const bool has1 = row0 + 1 < args.rows;
    const float* in = args.in[plane];
    float* out = args.out[plane];
    const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
    const float* r1 = has1 ? in + static_cast<long long>(row0 + 1) * args.in_stride : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        a[makhoul_pos(i, n)] = make_float2(r0[i], has1 ? r1[i] : 0.0f);
    }
    __syncthreads();
    // Per-row component factors for rows row0 and row0 + 1, or the plane's
    // full transfer table.
    const float* full = tf.full[plane];
    float fa[kMaxDctComponents], fb[kMaxDctComponents];
    if (!kFullTable) {
        for (int c = 0; c < tf.count; ++c) {
            fa[c] = tf.weight[c] * tf.t_rows[c][row0];
            fb[c] = has1 ? tf.weight[c] * tf.t_rows[c][row0 + 1] : 0.0f;
        }
    }
    float2* V = fft_lds(a, b, plan, args.table);
    float2* other = V == a ? b : a;
    float* spec = reinterpret_cast<float*>(other);  // n floats X_a, then n floats X_b
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const float2 z = V[k];
        const float2 zc = V[k == 0 ? 0 : n - k];
        const float2 va = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        const float2 vb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
        const float2 w = args.half_tw[k];
        float ta = 0.0f, tb = 0.0f;
        if (kFullTable) {
            ta = full[static_cast<long long>(row0) * n + k];
            tb = has1 ? full[static_cast<long long>(row0 + 1) * n + k] : 0.0f;
        } else {
            for (int c = 0; c < tf.count; ++c) {
                const float tk = tf.t_cols[c][k];
                ta = fmaf(fa[c], tk, ta);
                tb = fmaf(fb[c], tk, tb);
            }
        }
        spec[k] = (w.x * va.x - w.y * va.y) * ta;
        spec[n + k] = (w.x * vb.x - w.y * vb.y) * tb;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int km = k == 0 ? -1 : n - k;
        const float xa = spec[k], xam = km < 0 ? 0.0f : spec[km];
        const float xb = spec[n + k], xbm = km < 0 ? 0.0f : spec[n + km];
        const float2 w = args.half_tw[k];
        const float2 wc = make_float2(w.x, -w.y);
        const float2 va = cmul(wc, make_float2(xa, -xam));
        const float2 vb = cmul(wc, make_float2(xb, -xbm));
        const float2 v = make_float2(va.x - vb.y, va.y + vb.x);
        V[k] = make_float2(v.x, -v.y);
    }
    __syncthreads();
    float2* res = fft_lds(V, other, plan, args.table);
    const float inv = 1.0f / static_cast<float>(n);
    float* o0 = out + static_cast<long long>(row0) * args.out_stride;
    float* o1 = has1 ? out + static_cast<long long>(row0 + 1) * args.out_stride : nullptr;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const float2 z = res[makhoul_pos(i, n)];
        o0[i] = z.x * inv;
        if (o1) o1[i] = -z.y * inv;
    }
}

// 64x64 tiled transpose through LDS (padded against bank conflicts).
__global__ __launch_bounds__(256) void transpose_kernel(TransposeArgs t) {
    __shared__ float tile[64][65];
    const int plane = blockIdx.z;
    const float* in = t.in[plane];
    float* out = t.out[plane];
    const int bx = blockIdx.x * 64, by = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int j = ty; j < 64; j += 4) {
        const int x = bx + tx, y = by + j;
        if (x < t.width && y < t.height) tile[j][tx] = in[static_cast<long long>(y) * t.in_stride + x];
    }
    __syncthreads();
    for (int j = ty; j < 64; j += 4) {
        const int x = by + tx, y = bx + j;  // out is width x height transposed
        if (x < t.height && y < t.width) out[static_cast<long long>(y) * t.out_stride + x] = tile[tx][j];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t dct_lds_bytes(int n, bool) { return static_cast<size_t>(n) * sizeof(float2) * 2; }

// Threads per line-pair block: about one radix-4 butterfly per thread, whole waves.
static int row_threads(int n) {
    const int t = (n / 4 + 63) / 64 * 64;
    return t < 64 ? 64 : (t > kMaxRowThreads ? kMaxRowThreads : t);
}

hipError_t launch_dct2_rows(const DctRowsArgs& a, int planes, hipStream_t s) {
    const dim3 grid((a.rows + 1) / 2, planes);
    hipLaunchKernelGGL(dct2_rows_kernel, grid, dim3(row_threads(a.plan.n)), dct_lds_bytes(a.plan.n, false), s, a);
    return hipGetLastError();
}

hipError_t launch_dct3_rows(const DctRowsArgs& a, int planes, hipStream_t s) {
    if (a.rows <= a.row_begin) return hipSuccess;
    const dim3 grid((a.rows - a.row_begin + 1) / 2, planes);
    hipLaunchKernelGGL(dct3_rows_kernel, grid, dim3(row_threads(a.plan.n)), dct_lds_bytes(a.plan.n, false), s, a);
    return hipGetLastError();
}

hipError_t launch_dct_filter_rows(const DctRowsArgs& a, const DctTransfer& tf, int planes, hipStream_t s) {
    const dim3 grid((a.rows + 1) / 2, planes);
    if (tf.full[0]) {
        for (int p = 1; p < planes; ++p)
            if (!tf.full[p]) return hipErrorInvalidValue;
        hipLaunchKernelGGL(dct_filter_rows_kernel<true>, grid, dim3(row_threads(a.plan.n)), dct_lds_bytes(a.plan.n, true),
                           s, a, tf);
    } else {
        hipLaunchKernelGGL(dct_filter_rows_kernel<false>, grid, dim3(row_threads(a.plan.n)),
                           dct_lds_bytes(a.plan.n, true), s, a, tf);
    }
    return hipGetLastError();
}

hipError_t launch_transpose(const TransposeArgs& t, int planes, hipStream_t s) {
    const dim3 grid((t.width + 63) / 64, (t.height + 63) / 64, planes);
    hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, s, t);
    return hipGetLastError();
}

}  // namespace fjx

namespace fjx {

hipError_t fft_kernels_init(int* max_lds_n) {
    // Line transforms beyond 4096 points need more than the default 64 KiB of
    // LDS; fall back to 4096-point lines if the runtime refuses the request.
    const int bytes = static_cast<int>(dct_lds_bytes(kMaxLdsFft, true));
    const void* kernels[4] = {reinterpret_cast<const void*>(dct2_rows_kernel),
                              reinterpret_cast<const void*>(dct3_rows_kernel),
                              reinterpret_cast<const void*>(dct_filter_rows_kernel<false>),
                              reinterpret_cast<const void*>(dct_filter_rows_kernel<true>)};
    *max_lds_n = kMaxLdsFft;
    for (const void* k : kernels) {
        if (hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes) != hipSuccess) {
            (void)hipGetLastError();
            *max_lds_n = 4096;
        }
    }
    return hipSuccess;
}

}  // namespace fjx
