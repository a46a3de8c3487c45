// fj_fft_plan.cpp — host planning for the LDS Stockham FFT: factorization
// into butterfly radices and double-precision twiddle tables.
#include <cmath>
#include <vector>

#include "fj_fft.h"

namespace fjx {

bool make_fft_plan(int n, FftPlanDev& plan) {
    plan = FftPlanDev{};
    if (n < 1 || n > kMaxLdsFft) return false;
    plan.n = n;
    int m = n;
    std::vector<int> radices;
    while (m % 4 == 0) {
        radices.push_back(4);
        m /= 4;
    }
    while (m % 2 == 0) {
        radices.push_back(2);
        m /= 2;
    }
    for (int p = 3; p * p <= m || p <= m; p += 2) {
        while (m % p == 0) {
            radices.push_back(p);
            m /= p;
        }
        if (m == 1) break;
    }
    if (m != 1) radices.push_back(m);
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
    return true;
}

std::vector<float> fft_twiddles(int n) {
    FftPlanDev plan;
    make_fft_plan(n, plan);
    size_t stage = 0;
    for (int p = 0, ns = 1; p < plan.passes; ns *= plan.radix[p], ++p) stage += static_cast<size_t>(ns) * (plan.radix[p] - 1);
    std::vector<float> t(2 * (static_cast<size_t>(n) + stage));
    const double two_pi = 6.283185307179586476925286766559;
    auto put = [&](size_t i, long long e) {
        const double a = -two_pi * static_cast<double>(e % n) / n;
        t[2 * i] = static_cast<float>(std::cos(a));
        t[2 * i + 1] = static_cast<float>(std::sin(a));
    };
    for (int e = 0; e < n; ++e) put(static_cast<size_t>(e), e);
    for (int p = 0, ns = 1; p < plan.passes; ns *= plan.radix[p], ++p) {
        const int R = plan.radix[p];
        const int stride = n / (ns * R);
        for (int k = 0; k < ns; ++k)
            for (int r = 1; r < R; ++r)
                put(static_cast<size_t>(n) + plan.tw_offset[p] + static_cast<size_t>(k) * (R - 1) + (r - 1),
                    static_cast<long long>(r) * k * stride);
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
