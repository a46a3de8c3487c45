// fj_batch.h — strided device loops whose global loads are issued in groups.
#pragma once

#include <hip/hip_runtime.h>

namespace fjx {

// The strided loop i = first, first + step, ... < end in groups of kU: a
// group's global loads (load(i) -> T) are all issued before any of them is
// used (use(i, T)), so a thread waits for memory once per group rather than
// once per element (a plain loop that stores each load to LDS waits on every
// iteration).
template <int kU, typename T, typename Load, typename Use>
__device__ __forceinline__ void batched(int first, int end, int step, Load load, Use use) {
    for (int i0 = first; i0 < end; i0 += kU * step) {
        T v[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u * step;
            if (i < end) v[u] = load(i);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const int i = i0 + u * step;
            if (i < end) use(i, v[u]);
        }
    }
}

}  // namespace fjx
