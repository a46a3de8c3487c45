// fj_fft.hip — LDS Stockham FFT and the DCT-domain spatial filters built on it.
//
// Every line transform runs inside one workgroup: the line (a pair of real
// lines packed as re/im of one complex line) is staged into LDS, transformed
// with a self-sorting mixed-radix Stockham FFT, and written back. Twiddles come
// from per-length tables built on the host in double precision.
//
// Lines whose length has only the prime factors 2, 3 and 5 (every padded
// diffusion line, and the frame sides of the usual formats) run the in-place
// engine: one LDS buffer (odd radices first, so the strided writes of the
// first passes spread over the banks); each pass is one butterfly per thread, of a
// composite radix up to 25 (16 = 4x4, 15 = 3x5, 24 = 3x8, 25 = 5x5, ...) evaluated in
// registers as a two-level Cooley-Tukey DFT, so a 3840- or 5625-point line
// takes three passes. Other lengths run the ping-pong engine (radix 4, 2, 3, 5
// in registers, other primes through a generic LDS butterfly) over two buffers.
//
// DCT-II / DCT-III of length N use Makhoul's reordering onto one length-N
// complex FFT. The DCT-II basis diagonalizes convolution of the half-sample
// symmetric periodic extension of a line (the reference's reflect boundary,
// JuicerCudaFilmPipeline.cu:82-92) with any symmetric kernel, so a filter is
// applied as IDCT(T ⊙ DCT(x)) with T the kernel's frequency response.
#include <hip/hip_runtime.h>

#include "fj_batch.h"
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

// ---------------------------------------------------------------------------
// In-place engine
// ---------------------------------------------------------------------------

// cos / sin (2 pi e / N) of the composite radices' inner twiddles.
template <int N>
struct Roots;
template <> struct Roots<6> {
    static constexpr float c[6] = {1.000000000f, 5.000000000e-1f, -5.000000000e-1f, -1.000000000f, -5.000000000e-1f, 5.000000000e-1f};
    static constexpr float s[6] = {0.000000000f, 8.660254038e-1f, 8.660254038e-1f, 0.000000000f, -8.660254038e-1f, -8.660254038e-1f};
};
template <> struct Roots<8> {
    static constexpr float c[8] = {1.000000000f, 7.071067812e-1f, 0.000000000f, -7.071067812e-1f, -1.000000000f, -7.071067812e-1f, 0.000000000f, 7.071067812e-1f};
    static constexpr float s[8] = {0.000000000f, 7.071067812e-1f, 1.000000000f, 7.071067812e-1f, 0.000000000f, -7.071067812e-1f, -1.000000000f, -7.071067812e-1f};
};
template <> struct Roots<9> {
    static constexpr float c[9] = {1.000000000f, 7.660444431e-1f, 1.736481777e-1f, -5.000000000e-1f, -9.396926208e-1f, -9.396926208e-1f, -5.000000000e-1f, 1.736481777e-1f, 7.660444431e-1f};
    static constexpr float s[9] = {0.000000000f, 6.427876097e-1f, 9.848077530e-1f, 8.660254038e-1f, 3.420201433e-1f, -3.420201433e-1f, -8.660254038e-1f, -9.848077530e-1f, -6.427876097e-1f};
};
template <> struct Roots<10> {
    static constexpr float c[10] = {1.000000000f, 8.090169944e-1f, 3.090169944e-1f, -3.090169944e-1f, -8.090169944e-1f, -1.000000000f, -8.090169944e-1f, -3.090169944e-1f, 3.090169944e-1f, 8.090169944e-1f};
    static constexpr float s[10] = {0.000000000f, 5.877852523e-1f, 9.510565163e-1f, 9.510565163e-1f, 5.877852523e-1f, 0.000000000f, -5.877852523e-1f, -9.510565163e-1f, -9.510565163e-1f, -5.877852523e-1f};
};
template <> struct Roots<12> {
    static constexpr float c[12] = {1.000000000f, 8.660254038e-1f, 5.000000000e-1f, 0.000000000f, -5.000000000e-1f, -8.660254038e-1f, -1.000000000f, -8.660254038e-1f, -5.000000000e-1f, 0.000000000f, 5.000000000e-1f, 8.660254038e-1f};
    static constexpr float s[12] = {0.000000000f, 5.000000000e-1f, 8.660254038e-1f, 1.000000000f, 8.660254038e-1f, 5.000000000e-1f, 0.000000000f, -5.000000000e-1f, -8.660254038e-1f, -1.000000000f, -8.660254038e-1f, -5.000000000e-1f};
};
template <> struct Roots<15> {
    static constexpr float c[15] = {1.000000000f, 9.135454576e-1f, 6.691306064e-1f, 3.090169944e-1f, -1.045284633e-1f, -5.000000000e-1f, -8.090169944e-1f, -9.781476007e-1f, -9.781476007e-1f, -8.090169944e-1f, -5.000000000e-1f, -1.045284633e-1f, 3.090169944e-1f, 6.691306064e-1f, 9.135454576e-1f};
    static constexpr float s[15] = {0.000000000f, 4.067366431e-1f, 7.431448255e-1f, 9.510565163e-1f, 9.945218954e-1f, 8.660254038e-1f, 5.877852523e-1f, 2.079116908e-1f, -2.079116908e-1f, -5.877852523e-1f, -8.660254038e-1f, -9.945218954e-1f, -9.510565163e-1f, -7.431448255e-1f, -4.067366431e-1f};
};
template <> struct Roots<16> {
    static constexpr float c[16] = {1.000000000f, 9.238795325e-1f, 7.071067812e-1f, 3.826834324e-1f, 0.000000000f, -3.826834324e-1f, -7.071067812e-1f, -9.238795325e-1f, -1.000000000f, -9.238795325e-1f, -7.071067812e-1f, -3.826834324e-1f, 0.000000000f, 3.826834324e-1f, 7.071067812e-1f, 9.238795325e-1f};
    static constexpr float s[16] = {0.000000000f, 3.826834324e-1f, 7.071067812e-1f, 9.238795325e-1f, 1.000000000f, 9.238795325e-1f, 7.071067812e-1f, 3.826834324e-1f, 0.000000000f, -3.826834324e-1f, -7.071067812e-1f, -9.238795325e-1f, -1.000000000f, -9.238795325e-1f, -7.071067812e-1f, -3.826834324e-1f};
};
template <> struct Roots<18> {
    static constexpr float c[18] = {1.000000000f, 9.396926208e-1f, 7.660444431e-1f, 5.000000000e-1f, 1.736481777e-1f, -1.736481777e-1f, -5.000000000e-1f, -7.660444431e-1f, -9.396926208e-1f, -1.000000000f, -9.396926208e-1f, -7.660444431e-1f, -5.000000000e-1f, -1.736481777e-1f, 1.736481777e-1f, 5.000000000e-1f, 7.660444431e-1f, 9.396926208e-1f};
    static constexpr float s[18] = {0.000000000f, 3.420201433e-1f, 6.427876097e-1f, 8.660254038e-1f, 9.848077530e-1f, 9.848077530e-1f, 8.660254038e-1f, 6.427876097e-1f, 3.420201433e-1f, 0.000000000f, -3.420201433e-1f, -6.427876097e-1f, -8.660254038e-1f, -9.848077530e-1f, -9.848077530e-1f, -8.660254038e-1f, -6.427876097e-1f, -3.420201433e-1f};
};
template <> struct Roots<24> {
    static constexpr float c[24] = {1.000000000f, 9.659258263e-1f, 8.660254038e-1f, 7.071067812e-1f, 5.000000000e-1f, 2.588190451e-1f, 0.000000000f, -2.588190451e-1f, -5.000000000e-1f, -7.071067812e-1f, -8.660254038e-1f, -9.659258263e-1f, -1.000000000f, -9.659258263e-1f, -8.660254038e-1f, -7.071067812e-1f, -5.000000000e-1f, -2.588190451e-1f, 0.000000000f, 2.588190451e-1f, 5.000000000e-1f, 7.071067812e-1f, 8.660254038e-1f, 9.659258263e-1f};
    static constexpr float s[24] = {0.000000000f, 2.588190451e-1f, 5.000000000e-1f, 7.071067812e-1f, 8.660254038e-1f, 9.659258263e-1f, 1.000000000f, 9.659258263e-1f, 8.660254038e-1f, 7.071067812e-1f, 5.000000000e-1f, 2.588190451e-1f, 0.000000000f, -2.588190451e-1f, -5.000000000e-1f, -7.071067812e-1f, -8.660254038e-1f, -9.659258263e-1f, -1.000000000f, -9.659258263e-1f, -8.660254038e-1f, -7.071067812e-1f, -5.000000000e-1f, -2.588190451e-1f};
};
template <> struct Roots<20> {
    static constexpr float c[20] = {1.000000000f, 9.510565163e-1f, 8.090169944e-1f, 5.877852523e-1f, 3.090169944e-1f, 0.000000000f, -3.090169944e-1f, -5.877852523e-1f, -8.090169944e-1f, -9.510565163e-1f, -1.000000000f, -9.510565163e-1f, -8.090169944e-1f, -5.877852523e-1f, -3.090169944e-1f, 0.000000000f, 3.090169944e-1f, 5.877852523e-1f, 8.090169944e-1f, 9.510565163e-1f};
    static constexpr float s[20] = {0.000000000f, 3.090169944e-1f, 5.877852523e-1f, 8.090169944e-1f, 9.510565163e-1f, 1.000000000f, 9.510565163e-1f, 8.090169944e-1f, 5.877852523e-1f, 3.090169944e-1f, 0.000000000f, -3.090169944e-1f, -5.877852523e-1f, -8.090169944e-1f, -9.510565163e-1f, -1.000000000f, -9.510565163e-1f, -8.090169944e-1f, -5.877852523e-1f, -3.090169944e-1f};
};
template <> struct Roots<25> {
    static constexpr float c[25] = {1.000000000f, 9.685831611e-1f, 8.763066800e-1f, 7.289686274e-1f, 5.358267950e-1f, 3.090169944e-1f, 6.279051953e-2f, -1.873813146e-1f, -4.257792916e-1f, -6.374239897e-1f, -8.090169944e-1f, -9.297764859e-1f, -9.921147013e-1f, -9.921147013e-1f, -9.297764859e-1f, -8.090169944e-1f, -6.374239897e-1f, -4.257792916e-1f, -1.873813146e-1f, 6.279051953e-2f, 3.090169944e-1f, 5.358267950e-1f, 7.289686274e-1f, 8.763066800e-1f, 9.685831611e-1f};
    static constexpr float s[25] = {0.000000000f, 2.486898872e-1f, 4.817536741e-1f, 6.845471059e-1f, 8.443279255e-1f, 9.510565163e-1f, 9.980267284e-1f, 9.822872507e-1f, 9.048270525e-1f, 7.705132428e-1f, 5.877852523e-1f, 3.681245527e-1f, 1.253332336e-1f, -1.253332336e-1f, -3.681245527e-1f, -5.877852523e-1f, -7.705132428e-1f, -9.048270525e-1f, -9.822872507e-1f, -9.980267284e-1f, -9.510565163e-1f, -8.443279255e-1f, -6.845471059e-1f, -4.817536741e-1f, -2.486898872e-1f};
};

// x * exp(-2 pi i e / N) for 0 <= e < N (e is a constant once unrolled).
template <int N>
__device__ __forceinline__ float2 rot(float2 x, int e) {
    if (e == 0) return x;
    if (4 * e == N) return mul_mi(x);
    if (2 * e == N) return make_float2(-x.x, -x.y);
    if (4 * e == 3 * N) return make_float2(-x.y, x.x);
    return cmul(x, make_float2(Roots<N>::c[e], -Roots<N>::s[e]));
}

// Forward DFT of R points held in registers, in place: a[k] = sum_r a[r] w_R^(r k).
template <int R>
struct Dft;

template <>
struct Dft<2> {
    static __device__ __forceinline__ void run(float2* a) {
        const float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    }
};

template <>
struct Dft<3> {
    static __device__ __forceinline__ void run(float2* a) {
        const float c = -0.5f, s = -0.86602540378443864676f;  // cos/sin(-2pi/3)
        const float2 t = cadd(a[1], a[2]);
        const float2 u = csub(a[1], a[2]);
        const float2 m = make_float2(fmaf(c, t.x, a[0].x), fmaf(c, t.y, a[0].y));
        const float2 v = make_float2(-s * u.y, s * u.x);  // i*s*u
        a[0] = cadd(a[0], t);
        a[1] = cadd(m, v);
        a[2] = csub(m, v);
    }
};

template <>
struct Dft<4> {
    static __device__ __forceinline__ void run(float2* a) {
        const float2 s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        const float2 s13 = cadd(a[1], a[3]), d13 = mul_mi(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[2] = csub(s02, s13);
        a[1] = cadd(d02, d13);
        a[3] = csub(d02, d13);
    }
};

template <>
struct Dft<5> {
    static __device__ __forceinline__ void run(float2* a) {
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = -0.95105651629515357212f, s2 = -0.58778525229247312917f;  // sin(-2pi/5), sin(-4pi/5)
        const float2 t1 = cadd(a[1], a[4]), u1 = csub(a[1], a[4]);
        const float2 t2 = cadd(a[2], a[3]), u2 = csub(a[2], a[3]);
        const float2 m1 = make_float2(a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y);
        const float2 m2 = make_float2(a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y);
        const float2 v1 = make_float2(-(s1 * u1.y + s2 * u2.y), s1 * u1.x + s2 * u2.x);
        const float2 v2 = make_float2(-(s2 * u1.y - s1 * u2.y), s2 * u1.x - s1 * u2.x);
        a[0] = cadd(a[0], cadd(t1, t2));
        a[1] = cadd(m1, v1);
        a[4] = csub(m1, v1);
        a[2] = cadd(m2, v2);
        a[3] = csub(m2, v2);
    }
};

// N = R1 R2: input n = R2 n1 + n2, output k = k1 + R1 k2;
// X[k1 + R1 k2] = sum_n2 w_N^(n2 k1) [sum_n1 x[R2 n1 + n2] w_R1^(n1 k1)] w_R2^(n2 k2).
template <int R1, int R2>
struct DftCT {
    static __device__ __forceinline__ void run(float2* a) {
        constexpr int N = R1 * R2;
        float2 y[N];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            float2 t[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) t[n1] = a[R2 * n1 + n2];
            Dft<R1>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) y[n2 * R1 + k1] = rot<N>(t[k1], n2 * k1);
        }
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            float2 t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) t[n2] = y[n2 * R1 + k1];
            Dft<R2>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) a[k1 + R1 * k2] = t[k2];
        }
    }
};

template <> struct Dft<6> : DftCT<2, 3> {};
template <> struct Dft<8> : DftCT<2, 4> {};
template <> struct Dft<9> : DftCT<3, 3> {};
template <> struct Dft<10> : DftCT<2, 5> {};
template <> struct Dft<12> : DftCT<4, 3> {};
template <> struct Dft<15> : DftCT<3, 5> {};
template <> struct Dft<16> : DftCT<4, 4> {};
template <> struct Dft<18> : DftCT<2, 9> {};
template <> struct Dft<20> : DftCT<4, 5> {};
template <> struct Dft<24> : DftCT<3, 8> {};
template <> struct Dft<25> : DftCT<5, 5> {};

// One in-place Stockham pass of radix R: thread j < n / R owns butterfly j.
// All reads happen before the first barrier and all writes after it, so the
// butterfly may overwrite the slots other threads read.
// Powers w^1 .. w^(R-1) of a loaded w, applied to v[1..R-1]: even powers
// square w^(r/2), odd ones multiply w^(r-1) by w.
template <int R>
__device__ __forceinline__ void twiddle_powers(float2* v, float2 w) {
    float2 t[R];
    t[1] = w;
#pragma unroll
    for (int r = 2; r < R; ++r) t[r] = (r & 1) ? cmul(t[r - 1], w) : cmul(t[r / 2], t[r / 2]);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], t[r]);
}

// Radices up to 16 form their twiddles from one loaded base power (fewer
// loads, registers to spare); larger radices load all R - 1 (`stw`, R - 1
// per butterfly) to stay within the register budget.
template <int R>
__device__ __forceinline__ void pass_ip(float2* buf, int n, int ns, const float2* stw, const float2* stw1,
                                        unsigned magic) {
    const int count = n / R;
    // A fresh value per pass: the pass's R addresses are computed inside it
    // rather than hoisted out of the pass loop and held live across passes.
    int j = threadIdx.x;
    asm volatile("" : "+v"(j));
    const bool active = j < count;
    float2 v[R];
    int base = 0;
    if (active) {
        int q = j, k = 0;
        if (ns > 1) {
            q = static_cast<int>(__umulhi(static_cast<unsigned>(j), magic));
            k = j - q * ns;
        }
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = buf[j + r * count];
        if (ns > 1) {
            if constexpr (R <= 16) {
                twiddle_powers<R>(v, stw1[k]);
            } else {
                const float2* t = stw + k * (R - 1);
#pragma unroll
                for (int r = 1; r < R; ++r) v[r] = cmul(v[r], t[r - 1]);
            }
        }
        Dft<R>::run(v);
        base = q * ns * R + k;
    }
    __syncthreads();
    if (active) {
#pragma unroll
        for (int r = 0; r < R; ++r) buf[base + r * ns] = v[r];
    }
    __syncthreads();
}

// Forward FFT of buf[0..n) in place. blockDim.x >= n / R for every radix of
// the plan (the plan's thread count); kMaxR bounds the plan's radices, so a
// plan without 20 or 25 runs code with fewer live registers.
template <int kMaxR>
__device__ void fft_inplace(float2* buf, const FftPlanDev& plan, const float2* table) {
    const int len = fft_length(plan);
    int ns = 1;
    for (int p = 0; p < plan.passes; ++p) {
        const int r = plan.radix[p];
        const float2* stw = table + len + plan.tw_offset[p];
        const float2* stw1 = table + len + plan.tw1_offset[p];
        const unsigned magic = plan.ns_magic[p];
        switch (r) {
            case 2: pass_ip<2>(buf, len, ns, stw, stw1, magic); break;
            case 3: pass_ip<3>(buf, len, ns, stw, stw1, magic); break;
            case 4: pass_ip<4>(buf, len, ns, stw, stw1, magic); break;
            case 5: pass_ip<5>(buf, len, ns, stw, stw1, magic); break;
            case 6: pass_ip<6>(buf, len, ns, stw, stw1, magic); break;
            case 8: pass_ip<8>(buf, len, ns, stw, stw1, magic); break;
            case 9: pass_ip<9>(buf, len, ns, stw, stw1, magic); break;
            case 10: pass_ip<10>(buf, len, ns, stw, stw1, magic); break;
            case 12: pass_ip<12>(buf, len, ns, stw, stw1, magic); break;
            case 15: pass_ip<15>(buf, len, ns, stw, stw1, magic); break;
            case 16: pass_ip<16>(buf, len, ns, stw, stw1, magic); break;
            case 18:
                if constexpr (kMaxR >= 18) pass_ip<18>(buf, len, ns, stw, stw1, magic);
                break;
            case 20:
                if constexpr (kMaxR >= 20) pass_ip<20>(buf, len, ns, stw, stw1, magic);
                break;
            case 24:
                if constexpr (kMaxR >= 24) pass_ip<24>(buf, len, ns, stw, stw1, magic);
                break;
            case 25:
                if constexpr (kMaxR >= 25) pass_ip<25>(buf, len, ns, stw, stw1, magic);
                break;
            default: break;  // not produced by the planner for in-place plans
        }
        ns *= r;
    }
}

// Engines: 0 = ping-pong over two buffers; 16 / 25 = in place with radices up
// to 16 / 25; kBluestein = a Bluestein plan, in place over its length m.
constexpr int kBluestein = 125;
template <int kEng>
constexpr bool kInplace = kEng != 0;

// Length-n DFT of buf[0..n) by the chirp-z convolution of a Bluestein plan,
// in place in buf[0..m): x c, forward m-point transform, x the kernel
// spectrum, conjugate, forward again (the inverse up to conjugation and 1/m),
// then x c. Callers synchronize before (buf written) as for any line transform.
__device__ void bluestein_dft(float2* buf, const FftPlanDev& plan, const float2* table) {
    const int n = plan.n, m = plan.m;
    const float2* chirp = table + plan.chirp_offset;
    const float2* spectrum = table + plan.spectrum_offset;
    for (int j = threadIdx.x; j < m; j += blockDim.x) buf[j] = j < n ? cmul(buf[j], chirp[j]) : make_float2(0.0f, 0.0f);
    __syncthreads();
    fft_inplace<25>(buf, plan, table);
    for (int k = threadIdx.x; k < m; k += blockDim.x) {
        const float2 z = cmul(buf[k], spectrum[k]);
        buf[k] = make_float2(z.x, -z.y);
    }
    __syncthreads();
    fft_inplace<25>(buf, plan, table);
    const float inv = 1.0f / static_cast<float>(m);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const float2 z = buf[k];
        buf[k] = cmul(make_float2(z.x * inv, -z.y * inv), chirp[k]);
    }
    __syncthreads();
}

// Forward FFT with the plan's engine; returns the buffer holding the result.
template <int kEng>
__device__ __forceinline__ float2* fft_line(float2* a, float2* b, const FftPlanDev& plan, const float2* table) {
    if constexpr (kEng == kBluestein) {
        bluestein_dft(a, plan, table);
        return a;
    } else if constexpr (kInplace<kEng>) {
        fft_inplace<kEng>(a, plan, table);
        return a;
    } else {
        return fft_lds(a, b, plan, table);
    }
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

// DCT-II of two real lines packed as (re, im) of one complex line held in
// LDS (Makhoul order). On return out_a / out_b (global) hold X_a[k], X_b[k].
template <int kEng>
__device__ void dct2_pair(float2* a, float2* b, const FftPlanDev& plan, const float2* table, const float2* half_tw,
                          float* out_a, float* out_b) {
    const int n = plan.n;
    float2* V = fft_line<kEng>(a, b, plan, table);
    // V = Va + i Vb with Va, Vb spectra of real sequences.
    batched<8, float2>(
        threadIdx.x, n, blockDim.x, [&](int k) { return half_tw[k]; },  // exp(-i*pi*k/(2n))
        [&](int k, float2 w) {
            const float2 z = V[k];
            const float2 zc = V[k == 0 ? 0 : n - k];
            const float2 va = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
            const float2 vb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
            out_a[k] = w.x * va.x - w.y * va.y;
            if (out_b) out_b[k] = w.x * vb.x - w.y * vb.y;
        });
}

// Packed inverse input for one pair position: conj(e^{i pi k/2n}(X_a[k] - i X_a[n-k])
// + i e^{i pi k/2n}(X_b[k] - i X_b[n-k])); xam / xbm are 0 for k = 0.
__device__ __forceinline__ float2 dct3_pack(float xa, float xam, float xb, float xbm, float2 w) {
    const float2 wc = make_float2(w.x, -w.y);
    const float2 va = cmul(wc, make_float2(xa, -xam));
    const float2 vb = cmul(wc, make_float2(xb, -xbm));
    return make_float2(va.x - vb.y, -(va.y + vb.x));
}

}  // namespace

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// Rows [2*blockIdx.x, 2*blockIdx.x+1] of each plane (blockIdx.y) -> DCT-II rows.
template <int kEng>
__global__ __launch_bounds__(kMaxRowThreads) void dct2_rows_kernel(DctRowsArgs args) {
    extern __shared__ float2 lds2[];
    const FftPlanDev& plan = args.plan;
    const int n = plan.n;
    float2* a = lds2;
    float2* b = kInplace<kEng> ? nullptr : lds2 + n;
    const int row0 = 2 * blockIdx.x;
    const int plane = blockIdx.y;
    if (row0 >= args.rows) return;
    const bool has1 = row0 + 1 < args.rows;
    const float* in = args.in[plane];
    float* out = args.out[plane];
    auto put = [&](int i, float2 v) { a[makhoul_pos(i, n)] = v; };
    if (args.pad >= 0) {
        const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
        const float* r1 = has1 ? r0 + args.in_stride : r0;
        batched<8, float2>(
            threadIdx.x, n, blockDim.x,
            [&](int i) {
                const int si = reflect101(i - args.pad, args.src_w);
                return make_float2(r0[si], has1 ? r1[si] : 0.0f);
            },
            put);
    } else {
        const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
        const float* r1 = has1 ? in + static_cast<long long>(row0 + 1) * args.in_stride : nullptr;
        batched<8, float2>(
            threadIdx.x, n, blockDim.x, [&](int i) { return make_float2(r0[i], has1 ? r1[i] : 0.0f); }, put);
    }
    __syncthreads();
    dct2_pair<kEng>(a, b, plan, args.table, args.half_tw, out + static_cast<long long>(row0) * args.out_stride,
                   has1 ? out + static_cast<long long>(row0 + 1) * args.out_stride : nullptr);
}

// Inverse (DCT-III scaled so that it undoes dct2) of row pairs.
template <int kEng>
__global__ __launch_bounds__(kMaxRowThreads) void dct3_rows_kernel(DctRowsArgs args) {
    extern __shared__ float2 lds2[];
    const FftPlanDev& plan = args.plan;
    const int n = plan.n;
    float2* a = lds2;
    float2* b = kInplace<kEng> ? nullptr : lds2 + n;
    const int row0 = args.row_begin + 2 * blockIdx.x;
    const int plane = blockIdx.y;
    if (row0 >= args.rows) return;
    const bool has1 = row0 + 1 < args.rows;
    const float* in = args.in[plane];
    float* out = args.out[plane];
    const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
    const float* r1 = has1 ? in + static_cast<long long>(row0 + 1) * args.in_stride : r0;
    struct Pack {
        float xa, xam, xb, xbm;
        float2 w;
    };
    batched<8, Pack>(
        threadIdx.x, n, blockDim.x,
        [&](int k) {
            const int km = k == 0 ? -1 : n - k;
            Pack v;
            v.xa = r0[k];
            v.xam = km < 0 ? 0.0f : r0[km];
            v.xb = has1 ? r1[k] : 0.0f;
            v.xbm = (has1 && km >= 0) ? r1[km] : 0.0f;
            v.w = args.half_tw[k];
            return v;
        },
        [&](int k, const Pack& v) { a[k] = dct3_pack(v.xa, v.xam, v.xb, v.xbm, v.w); });
    __syncthreads();
    float2* res = fft_line<kEng>(a, b, plan, args.table);
    const float inv = 1.0f / static_cast<float>(n);
    if (args.pad >= 0) {
        // Elements [pad, pad + src_w) of the padded line are the row's own.
        const int pad = args.pad;
        const float f = args.fraction, keep = 1.0f - args.fraction;
        const long long y0 = row0;
        const float* m = args.mix[plane];
        const float* m0 = m + y0 * args.out_stride;
        float* q0 = out + y0 * args.out_stride;
        batched<8, float2>(
            threadIdx.x, args.src_w, blockDim.x,
            [&](int i) { return make_float2(m0[i], has1 ? m0[i + args.out_stride] : 0.0f); },
            [&](int i, float2 mv) {
                const float2 z = res[makhoul_pos(i + pad, n)];
                q0[i] = keep * mv.x + f * (z.x * inv);
                if (has1) q0[i + args.out_stride] = keep * mv.y + f * (-z.y * inv);
            });
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

// Transfer value T(row, k) of one row of a filter launch.
template <bool kFullTable>
__device__ __forceinline__ float transfer_at(const DctTransfer& tf, const float* full, const float* f, int row, int n,
                                             int k) {
    if (kFullTable) return full[static_cast<long long>(row) * n + k];
    float t = 0.0f;
    for (int c = 0; c < tf.count; ++c) t = fmaf(f[c], tf.t_cols[c][k], t);
    return t;
}

// Forward DCT-II along rows, multiply by T, inverse along rows, in one pass.
// Used on the transposed plane (rows = original columns): the row index is the
// first-axis frequency kx, the in-row index the second-axis frequency.
// kFullTable: T is a full per-plane table (diffusion); otherwise a weighted sum
// of separable component factors (spatial DIR).
template <bool kFullTable, int kEng>
__device__ __forceinline__ void dct_filter_rows_body(const DctRowsArgs& args, const DctTransfer& tf) {
    extern __shared__ float2 lds2[];
    const FftPlanDev& plan = args.plan;
    const int n = plan.n;
    float2* a = lds2;
    float2* b = kInplace<kEng> ? nullptr : lds2 + n;
    const int row0 = 2 * blockIdx.x;
    const int plane = blockIdx.y;
    if (row0 >= args.rows) return;
    const bool has1 = row0 + 1 < args.rows;
    const float* in = args.in[plane];
    float* out = args.out[plane];
    const float* r0 = in + static_cast<long long>(row0) * args.in_stride;
    const float* r1 = has1 ? in + static_cast<long long>(row0 + 1) * args.in_stride : nullptr;
    batched<8, float2>(
        threadIdx.x, n, blockDim.x, [&](int i) { return make_float2(r0[i], has1 ? r1[i] : 0.0f); },
        [&](int i, float2 v) { a[makhoul_pos(i, n)] = v; });
    __syncthreads();
    // Per-row component factors for rows row0 and row0 + 1, or the plane's
    // full transfer table.
    const float* full = tf.full[plane];
    float fa[kMaxDctComponents], fb[kMaxDctComponents];
    if (!kFullTable) {
#pragma unroll
        for (int c = 0; c < kMaxDctComponents; ++c) {
            fa[c] = c < tf.count ? tf.weight[c] * tf.t_rows[c][row0] : 0.0f;
            fb[c] = c < tf.count && has1 ? tf.weight[c] * tf.t_rows[c][row0 + 1] : 0.0f;
        }
    }
    const int row1 = has1 ? row0 + 1 : row0;
    float2* V = fft_line<kEng>(a, b, plan, args.table);
    if constexpr (kInplace<kEng>) {
        // Spectrum product and inverse packing, in place: positions k and
        // n - k depend only on V[k] and V[n - k], so each thread owns a pair.
        struct Coef {
            float2 w1, w2;
            float ta1, ta2, tb1, tb2;
        };
        batched<kFullTable ? 4 : 2, Coef>(
            threadIdx.x, n / 2 + 1, blockDim.x,
            [&](int p) {
                const int k1 = p, k2 = p == 0 ? 0 : n - p;
                Coef c;
                c.w1 = args.half_tw[k1];
                c.w2 = args.half_tw[k2];
                if constexpr (kFullTable) {
                    c.ta1 = transfer_at<true>(tf, full, fa, row0, n, k1);
                    c.ta2 = transfer_at<true>(tf, full, fa, row0, n, k2);
                    c.tb1 = has1 ? transfer_at<true>(tf, full, fb, row1, n, k1) : 0.0f;
                    c.tb2 = has1 ? transfer_at<true>(tf, full, fb, row1, n, k2) : 0.0f;
                } else {
                    // transfer_at's sums, with each column factor loaded once
                    // (both rows use it) and all of them before the first sum.
                    float v1[kMaxDctComponents], v2[kMaxDctComponents];
#pragma unroll
                    for (int q = 0; q < kMaxDctComponents; ++q) {
                        v1[q] = q < tf.count ? tf.t_cols[q][k1] : 0.0f;
                        v2[q] = q < tf.count ? tf.t_cols[q][k2] : 0.0f;
                    }
                    c.ta1 = c.ta2 = c.tb1 = c.tb2 = 0.0f;
#pragma unroll
                    for (int q = 0; q < kMaxDctComponents; ++q) {
                        if (q < tf.count) {
                            c.ta1 = fmaf(fa[q], v1[q], c.ta1);
                            c.ta2 = fmaf(fa[q], v2[q], c.ta2);
                            c.tb1 = fmaf(fb[q], v1[q], c.tb1);
                            c.tb2 = fmaf(fb[q], v2[q], c.tb2);
                        }
                    }
                    if (!has1) c.tb1 = c.tb2 = 0.0f;
                }
                return c;
            },
            [&](int p, const Coef& c) {
                const int k1 = p, k2 = p == 0 ? 0 : n - p;
                const float2 z1 = V[k1], z2 = V[k2];
                const float2 w1 = c.w1, w2 = c.w2;
                // DCT-II coefficients of both lines at k1 and k2, times T.
                const float2 va1 = make_float2(0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y));
                const float2 vb1 = make_float2(0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x));
                const float2 va2 = make_float2(0.5f * (z2.x + z1.x), 0.5f * (z2.y - z1.y));
                const float2 vb2 = make_float2(0.5f * (z2.y + z1.y), -0.5f * (z2.x - z1.x));
                const float xa1 = (w1.x * va1.x - w1.y * va1.y) * c.ta1;
                const float xb1 = (w1.x * vb1.x - w1.y * vb1.y) * c.tb1;
                const float xa2 = (w2.x * va2.x - w2.y * va2.y) * c.ta2;
                const float xb2 = (w2.x * vb2.x - w2.y * vb2.y) * c.tb2;
                if (p == 0) {
                    V[0] = dct3_pack(xa1, 0.0f, xb1, 0.0f, w1);
                } else {
                    V[k1] = dct3_pack(xa1, xa2, xb1, xb2, w1);
                    if (k2 != k1) V[k2] = dct3_pack(xa2, xa1, xb2, xb1, w2);
                }
            });
    } else {
        float2* other = V == a ? b : a;
        float* spec = reinterpret_cast<float*>(other);  // n floats X_a, then n floats X_b
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const float2 z = V[k];
            const float2 zc = V[k == 0 ? 0 : n - k];
            const float2 va = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
            const float2 vb = make_float2(0.5f * (z.y + zc.y), -0.5f * (z.x - zc.x));
            const float2 w = args.half_tw[k];
            const float ta = transfer_at<kFullTable>(tf, full, fa, row0, n, k);
            const float tb = has1 ? transfer_at<kFullTable>(tf, full, fb, row1, n, k) : 0.0f;
            spec[k] = (w.x * va.x - w.y * va.y) * ta;
            spec[n + k] = (w.x * vb.x - w.y * vb.y) * tb;
        }
        __syncthreads();
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const int km = k == 0 ? -1 : n - k;
            const float xam = km < 0 ? 0.0f : spec[km];
            const float xbm = km < 0 ? 0.0f : spec[n + km];
            V[k] = dct3_pack(spec[k], xam, spec[n + k], xbm, args.half_tw[k]);
        }
    }
    __syncthreads();
    float2* res = fft_line<kEng>(V, kInplace<kEng> ? nullptr : (V == a ? b : a), plan, args.table);
    const float inv = 1.0f / static_cast<float>(n);
    const int begin = args.out_begin, count = args.out_count > 0 ? args.out_count : n;
    float* o0 = out + static_cast<long long>(row0) * args.out_stride - begin;
    float* o1 = has1 ? o0 + args.out_stride : nullptr;
    for (int i = begin + threadIdx.x; i < begin + count; i += blockDim.x) {
        const float2 z = res[makhoul_pos(i, n)];
        o0[i] = z.x * inv;
        if (o1) o1[i] = -z.y * inv;
    }
}

template <bool kFullTable, int kEng>
__global__ __launch_bounds__(kMaxRowThreads) void dct_filter_rows_kernel(DctRowsArgs args, DctTransfer tf) {
    dct_filter_rows_body<kFullTable, kEng>(args, tf);
}

// The diffusion column filter on radix <= 16 plans: capped at 96 VGPRs
// (five waves per SIMD) it still needs no scratch, and five 3840-point
// blocks then fit a CU by registers as they do by LDS.
template <>
__global__ __launch_bounds__(kMaxRowThreads) __attribute__((amdgpu_waves_per_eu(5, 8))) void
dct_filter_rows_kernel<true, 16>(DctRowsArgs args, DctTransfer tf) {
    dct_filter_rows_body<true, 16>(args, tf);
}

// 64x64 tiled transpose through LDS (padded against bank conflicts).
__global__ __launch_bounds__(256) void transpose_kernel(TransposeArgs t) {
    __shared__ float tile[64][65];
    const int plane = blockIdx.z;
    const float* in = t.in[plane];
    float* out = t.out[plane];
    const int bx = blockIdx.x * 64, by = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    // All 16 loads of a thread are issued before the first is used.
    float v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int j = ty + 4 * u;
        const int x = bx + tx, y = by + j;
        v[u] = 0.0f;  // outside the plane: never read back
        if (x < t.width && y < t.height) {
            const int sy = t.src_rows > 0 ? reflect101(y - t.row_pad, t.src_rows) : y;
            v[u] = in[static_cast<long long>(sy) * t.in_stride + x];
        }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) tile[ty + 4 * u][tx] = v[u];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        const int j = ty + 4 * u;
        const int x = by + tx, y = bx + j;  // out is width x height transposed
        if (x < t.height && y < t.width) out[static_cast<long long>(y) * t.out_stride + x] = tile[tx][j];
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
size_t dct_lds_bytes(const FftPlanDev& plan) {
    const size_t n = static_cast<size_t>(fft_length(plan));
    return (plan.inplace ? n : 2 * n) * sizeof(float2);
}

// Threads per line-pair block: the in-place plan's butterfly count, else
// about one radix-4 butterfly per thread; whole waves.
static int row_threads(const FftPlanDev& plan) {
    if (plan.inplace) return plan.threads;
    const int t = (plan.n / 4 + 63) / 64 * 64;
    return t < 64 ? 64 : (t > kMaxRowThreads ? kMaxRowThreads : t);
}

static int engine(const FftPlanDev& plan) {
    if (plan.bluestein) return kBluestein;
    if (!plan.inplace) return 0;
    int max_r = 0;
    for (int p = 0; p < plan.passes; ++p) max_r = plan.radix[p] > max_r ? plan.radix[p] : max_r;
    return max_r <= 16 ? 16 : 25;
}

hipError_t launch_dct2_rows(const DctRowsArgs& a, int planes, hipStream_t s) {
    const dim3 grid((a.rows + 1) / 2, planes), block(row_threads(a.plan));
    const size_t lds = dct_lds_bytes(a.plan);
    switch (engine(a.plan)) {
        case 16: hipLaunchKernelGGL(dct2_rows_kernel<16>, grid, block, lds, s, a); break;
        case 25: hipLaunchKernelGGL(dct2_rows_kernel<25>, grid, block, lds, s, a); break;
        case kBluestein: hipLaunchKernelGGL(dct2_rows_kernel<kBluestein>, grid, block, lds, s, a); break;
        default: hipLaunchKernelGGL(dct2_rows_kernel<0>, grid, block, lds, s, a); break;
    }
    return hipGetLastError();
}

hipError_t launch_dct3_rows(const DctRowsArgs& a, int planes, hipStream_t s) {
    if (a.rows <= a.row_begin) return hipSuccess;
    const dim3 grid((a.rows - a.row_begin + 1) / 2, planes), block(row_threads(a.plan));
    const size_t lds = dct_lds_bytes(a.plan);
    switch (engine(a.plan)) {
        case 16: hipLaunchKernelGGL(dct3_rows_kernel<16>, grid, block, lds, s, a); break;
        case 25: hipLaunchKernelGGL(dct3_rows_kernel<25>, grid, block, lds, s, a); break;
        case kBluestein: hipLaunchKernelGGL(dct3_rows_kernel<kBluestein>, grid, block, lds, s, a); break;
        default: hipLaunchKernelGGL(dct3_rows_kernel<0>, grid, block, lds, s, a); break;
    }
    return hipGetLastError();
}

template <bool kFull>
static void launch_filter(const DctRowsArgs& a, const DctTransfer& tf, dim3 grid, dim3 block, size_t lds,
                          hipStream_t s) {
    switch (engine(a.plan)) {
        case 16: hipLaunchKernelGGL((dct_filter_rows_kernel<kFull, 16>), grid, block, lds, s, a, tf); break;
        case 25: hipLaunchKernelGGL((dct_filter_rows_kernel<kFull, 25>), grid, block, lds, s, a, tf); break;
        case kBluestein:
            hipLaunchKernelGGL((dct_filter_rows_kernel<kFull, kBluestein>), grid, block, lds, s, a, tf);
            break;
        default: hipLaunchKernelGGL((dct_filter_rows_kernel<kFull, 0>), grid, block, lds, s, a, tf); break;
    }
}

hipError_t launch_dct_filter_rows(const DctRowsArgs& a, const DctTransfer& tf, int planes, hipStream_t s) {
    const dim3 grid((a.rows + 1) / 2, planes), block(row_threads(a.plan));
    const size_t lds = dct_lds_bytes(a.plan);
    const bool full = tf.full[0] != nullptr;
    for (int p = 1; p < planes; ++p)
        if ((tf.full[p] != nullptr) != full) return hipErrorInvalidValue;
    if (full)
        launch_filter<true>(a, tf, grid, block, lds, s);
    else
        launch_filter<false>(a, tf, grid, block, lds, s);
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
    // LDS (the ping-pong engine holds two lines); fall back to 4096-point
    // lines if the runtime refuses the request.
    const int bytes = static_cast<int>(kMaxLdsFft * sizeof(float2));  // = 2 x kMaxPingPongFft float2
    const void* kernels[] = {reinterpret_cast<const void*>(dct2_rows_kernel<0>),
                             reinterpret_cast<const void*>(dct2_rows_kernel<16>),
                             reinterpret_cast<const void*>(dct2_rows_kernel<25>),
                             reinterpret_cast<const void*>(dct3_rows_kernel<0>),
                             reinterpret_cast<const void*>(dct3_rows_kernel<16>),
                             reinterpret_cast<const void*>(dct3_rows_kernel<25>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<false, 0>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<false, 16>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<false, 25>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<true, 0>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<true, 16>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<true, 25>),
                             reinterpret_cast<const void*>(dct2_rows_kernel<kBluestein>),
                             reinterpret_cast<const void*>(dct3_rows_kernel<kBluestein>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<false, kBluestein>),
                             reinterpret_cast<const void*>(dct_filter_rows_kernel<true, kBluestein>)};
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
