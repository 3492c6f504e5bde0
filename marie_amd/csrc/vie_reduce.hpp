// Deterministic block reductions shared by the Krylov solvers: wave64
// shuffles, then the per-wave sums added by thread 0 in wave order.
#pragma once
#include <hip/hip_runtime.h>

namespace marie {

// v[q] of every thread of an NTH-thread block summed into thread 0's v[q];
// sh: NTH/64 * W doubles of LDS
template <int W, int NTH = 256>
__device__ __forceinline__ void block_reduce(double (&v)[W], double* sh) {
#pragma unroll
    for (int q = 0; q < W; ++q) {
        double s = v[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        v[q] = s;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < W; ++q) sh[wid * W + q] = v[q];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < W; ++q) {
            double s = 0.0;
            for (int w = 0; w < NTH / 64; ++w) s += sh[w * W + q];
            v[q] = s;
        }
    }
    __syncthreads();
}

}  // namespace marie
