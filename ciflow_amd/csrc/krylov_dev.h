// Device helpers shared by the Krylov step kernels (krylov.hip) and the
// batched Davidson step kernels (davbatch.hip).
#pragma once

#include "common.h"

namespace cf {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Selective second Gram-Schmidt pass (Daniel-Gragg-Kaufman-Stewart): the pass
// is skipped when the first one kept more than eta of the vector,
// sum(after) > eta^2 sum(before), both given as np <= 256 per-block parts.
// Every thread of the block gets the same answer (fixed summation order).
__device__ __forceinline__ bool dgks_skip(const double* __restrict__ before, const double* __restrict__ after, int np,
                                          double eta2) {
    __shared__ double red[2][16];
    const int t = threadIdx.x;
    double b = t < np ? before[t] : 0.0, a = t < np ? after[t] : 0.0;
    b = wave_sum(b);
    a = wave_sum(a);
    if ((t & 63) == 0) {
        red[0][t >> 6] = b;
        red[1][t >> 6] = a;
    }
    __syncthreads();
    double sb = 0.0, sa = 0.0;
    const int waves = (blockDim.x + 63) >> 6;
    for (int w = 0; w < waves; ++w) {
        sb += red[0][w];
        sa += red[1][w];
    }
    __syncthreads();
    return sa > eta2 * sb;
}

// cs[i] = sum_b partial[i*G + b] for i < k (G a power of two <= 64, k <= 64),
// the same arithmetic in every workgroup of 256 threads; all loads are issued
// before the first shuffle.
__device__ __forceinline__ void reduce_coeffs(int k, int G, const double* __restrict__ partial, double* cs) {
    const int per_round = 256 / G;
    const int gi = threadIdx.x / G, b = threadIdx.x % G;
    double v[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int i = r * per_round + gi;
        v[r] = (i < k) ? partial[i * G + b] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r * per_round >= k) break;
        double x = v[r];
        for (int off = G >> 1; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        const int i = r * per_round + gi;
        if (b == 0 && i < k) cs[i] = x;
    }
}

// Element q of the deterministic start vector: 1 + 0.5 (u(seed, q) - 0.5),
// u in [0, 1) from a splitmix64-style mix of seed and q.
__device__ __forceinline__ double start_element(uint64_t seed, uint64_t q) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (q + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    return 1.0 + 0.5 * (u - 0.5);
}

}  // namespace cf
