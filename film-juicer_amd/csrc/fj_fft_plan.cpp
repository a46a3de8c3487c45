// fj_fft_plan.cpp — host planning for the LDS Stockham FFT: factorization
// into butterfly radices and double-precision twiddle tables.
#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <vector>

#include "fj_fft.h"

namespace fjx {

namespace {

struct Factoring {
    bool ok = false;
    int passes = 0;
    int min_radix = 0;
    std::vector<int> radices;
};

// Composite radices of the in-place engine (largest first).
constexpr int kInplaceRadices[] = {25, 24, 20, 18, 16, 15, 12, 10, 9, 8, 6, 5, 4, 3, 2};

// Fewest passes, then the largest smallest radix, over factorings of m into
// in-place radices no smaller than rmin.
Factoring best_factoring(int m, int rmin, std::map<int, Factoring>& memo) {
    if (m == 1) {
        Factoring f;
        f.ok = true;
        f.min_radix = 1 << 30;
        return f;
    }
    auto it = memo.find(m);
    if (it != memo.end()) return it->second;
    Factoring best;
    for (int r : kInplaceRadices) {
        if (r < rmin || m % r != 0) continue;
        const Factoring sub = best_factoring(m / r, rmin, memo);
        if (!sub.ok) continue;
        const int passes = sub.passes + 1;
        const int min_radix = std::min(sub.min_radix, r);
        if (!best.ok || passes < best.passes || (passes == best.passes && min_radix > best.min_radix)) {
            best.ok = true;
            best.passes = passes;
            best.min_radix = min_radix;
            best.radices = sub.radices;
            best.radices.push_back(r);
        }
    }
    memo.emplace(m, best);
    return best;
}

bool inplace_radices(int n, std::vector<int>& radices, int& threads) {
    int m = n;
    for (int p : {2, 3, 5})
        while (m % p == 0) m /= p;
    if (m != 1) return false;
    std::map<int, Factoring> memo;
    const int rmin = (n + kMaxRowThreads - 1) / kMaxRowThreads;
    const Factoring f = best_factoring(n, rmin, memo);
    if (!f.ok) return false;
    radices = f.radices;
    // Odd radices first: a first pass of odd radix R writes with an odd
    // stride, which spreads a wave's stores over the LDS banks.
    std::sort(radices.begin(), radices.end(), [](int a, int b) {
        if ((a & 1) != (b & 1)) return (a & 1) != 0;
        return a > b;
    });
    int count = 1;
    for (int r : radices) count = std::max(count, n / r);
    threads = std::max(64, (count + 63) / 64 * 64);
    return threads <= kMaxRowThreads;
}

}  // namespace

bool make_fft_plan(int n, FftPlanDev& plan) {
    plan = FftPlanDev{};
    if (n < 1 || n > kMaxLdsFft) return false;
    plan.n = n;
    std::vector<int> radices;
    int threads = 0;
    if (!inplace_radices(n, radices, threads)) {
        // Chirp-z over the smallest smooth length that holds the linear
        // convolution of two n-point sequences.
        const int m = inplace_fft_length(2 * n - 1);
        if (m == 0 || !inplace_radices(m, radices, threads)) return false;
        plan.bluestein = 1;
        plan.m = m;
    }
    plan.inplace = 1;
    plan.threads = threads;
    if (static_cast<int>(radices.size()) > kMaxFftPasses) return false;
    plan.passes = static_cast<int>(radices.size());
    int ns = 1, offset = 0;
    for (int i = 0; i < plan.passes; ++i) {
        plan.radix[i] = radices[i];
        plan.tw_offset[i] = offset;
        plan.ns_magic[i] = ns > 1 ? static_cast<unsigned>((0x100000000ull + ns - 1) / ns) : 0u;
        offset += ns * (radices[i] - 1);
        ns *= radices[i];
    }
    ns = 1;
    for (int i = 0; i < plan.passes; ++i) {
        plan.tw1_offset[i] = offset;
        offset += ns;
        ns *= radices[i];
    }
    if (plan.bluestein) {
        plan.chirp_offset = plan.m + offset;
        plan.spectrum_offset = plan.chirp_offset + n;
    }
    return true;
}

int inplace_fft_length(int n) {
    std::vector<int> radices;
    int threads = 0;
    for (int m = std::max(n, 1); m <= kMaxLdsFft; ++m)
        if (inplace_radices(m, radices, threads)) return m;
    return 0;
}

namespace {

// Forward DFT of a 2,3,5-smooth length in double precision (recursive
// decimation in time), for the Bluestein kernel spectrum.
void dft_smooth(std::vector<std::complex<double>>& x) {
    const size_t n = x.size();
    if (n <= 1) return;
    size_t r = 0;
    for (size_t p : {size_t(2), size_t(3), size_t(5)})
        if (n % p == 0) {
            r = p;
            break;
        }
    const size_t q = n / r;
    std::vector<std::vector<std::complex<double>>> sub(r, std::vector<std::complex<double>>(q));
    for (size_t s = 0; s < r; ++s)
        for (size_t j = 0; j < q; ++j) sub[s][j] = x[j * r + s];
    for (auto& v : sub) dft_smooth(v);
    const double two_pi = 6.283185307179586476925286766559;
    for (size_t k = 0; k < n; ++k) {
        std::complex<double> acc = 0.0;
        for (size_t s = 0; s < r; ++s) {
            const double a = -two_pi * static_cast<double>((s * k) % n) / static_cast<double>(n);
            acc += sub[s][k % q] * std::complex<double>(std::cos(a), std::sin(a));
        }
        x[k] = acc;
    }
}

}  // namespace

std::vector<float> fft_twiddles(int n) {
    FftPlanDev plan;
    make_fft_plan(n, plan);
    const int len = fft_length(plan);
    size_t stage = 0;
    for (int p = 0, ns = 1; p < plan.passes; ns *= plan.radix[p], ++p) stage += static_cast<size_t>(ns) * plan.radix[p];
    const size_t extra = plan.bluestein ? static_cast<size_t>(n) + static_cast<size_t>(plan.m) : 0;
    std::vector<float> t(2 * (static_cast<size_t>(len) + stage + extra));
    const double two_pi = 6.283185307179586476925286766559;
    auto put = [&](size_t i, long long e) {
        const double a = -two_pi * static_cast<double>(e % len) / len;
        t[2 * i] = static_cast<float>(std::cos(a));
        t[2 * i + 1] = static_cast<float>(std::sin(a));
    };
    for (int e = 0; e < len; ++e) put(static_cast<size_t>(e), e);
    for (int p = 0, ns = 1; p < plan.passes; ns *= plan.radix[p], ++p) {
        const int R = plan.radix[p];
        const int stride = len / (ns * R);
        for (int k = 0; k < ns; ++k)
            for (int r = 1; r < R; ++r)
                put(static_cast<size_t>(len) + plan.tw_offset[p] + static_cast<size_t>(k) * (R - 1) + (r - 1),
                    static_cast<long long>(r) * k * stride);
        for (int k = 0; k < ns; ++k)
            put(static_cast<size_t>(len) + plan.tw1_offset[p] + static_cast<size_t>(k), static_cast<long long>(k) * stride);
    }
    if (plan.bluestein) {
        // c_j = exp(-i pi j^2 / n) with j^2 reduced mod 2n (exact in integers).
        const double pi = 3.14159265358979323846264338327950;
        std::vector<std::complex<double>> c(n);
        for (int j = 0; j < n; ++j) {
            const long long e = (static_cast<long long>(j) * j) % (2LL * n);
            const double a = -pi * static_cast<double>(e) / n;
            c[j] = std::complex<double>(std::cos(a), std::sin(a));
            t[2 * (plan.chirp_offset + static_cast<size_t>(j))] = static_cast<float>(c[j].real());
            t[2 * (plan.chirp_offset + static_cast<size_t>(j)) + 1] = static_cast<float>(c[j].imag());
        }
        // b = conj(c) at offsets 0..n-1 and m-1..m-n+1 (the negative ones).
        std::vector<std::complex<double>> b(plan.m, 0.0);
        for (int j = 0; j < n; ++j) {
            b[j] = std::conj(c[j]);
            if (j) b[plan.m - j] = std::conj(c[j]);
        }
        dft_smooth(b);
        for (int k = 0; k < plan.m; ++k) {
            t[2 * (plan.spectrum_offset + static_cast<size_t>(k))] = static_cast<float>(b[k].real());
            t[2 * (plan.spectrum_offset + static_cast<size_t>(k)) + 1] = static_cast<float>(b[k].imag());
        }
    }
    return t;
}

std::vector<float> dct_half_twiddles(int n) {
    std::vector<float> t(2 * static_cast<size_t>(n));
    const double pi = 3.14159265358979323846264338327950;
    for (int k = 0; k < n; ++k) {
        const double a = -pi * static_cast<double>(k) / (2.0 * n);
        t[2 * k] = static_cast<float>(std::cos(a));
        t[2 * k + 1] = static_cast<float>(std::sin(a));
    }
    return t;
}

}  // namespace fjx
