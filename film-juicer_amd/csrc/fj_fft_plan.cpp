// fj_fft_plan.cpp — host planning for the LDS Stockham FFT: factorization
// into butterfly radices and double-precision twiddle tables.
#include <algorithm>
#include <cmath>
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
    if (inplace_radices(n, radices, threads)) {
        plan.inplace = 1;
        plan.threads = threads;
    } else {
        if (n > kMaxPingPongFft) return false;
        radices.clear();
        int m = n;
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
    }
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
    return true;
}

int inplace_fft_length(int n) {
    std::vector<int> radices;
    int threads = 0;
    for (int m = std::max(n, 1); m <= kMaxLdsFft; ++m)
        if (inplace_radices(m, radices, threads)) return m;
    return 0;
}

std::vector<float> fft_twiddles(int n) {
    FftPlanDev plan;
    make_fft_plan(n, plan);
    size_t stage = 0;
    for (int p = 0, ns = 1; p < plan.passes; ns *= plan.radix[p], ++p) stage += static_cast<size_t>(ns) * plan.radix[p];
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
        for (int k = 0; k < ns; ++k)
            put(static_cast<size_t>(n) + plan.tw1_offset[p] + static_cast<size_t>(k), static_cast<long long>(k) * stride);
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
