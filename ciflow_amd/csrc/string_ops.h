// Device helpers on occupation bit strings shared by the string tables, the
// CRS builds and the determinant-list build.
#pragma once

#include "kernels.h"

namespace cf {

// Rank of a string in increasing-integer (Gosper) order: the combinatorial
// number system, rank(s) = sum_m C(k_m, m) with k_m the position of the m-th
// set bit (Permutator.cpp:113-153 enumerates, CIMethod.cpp:663-747 indexes).
__device__ __forceinline__ uint32_t rank_bits(str_t s, const uint64_t* binom) {
    uint32_t r = 0;
    int m = 1;
    while (s) {
        r += (uint32_t)binom[__builtin_ctzll(s) * kBinomCols + m];
        ++m;
        s &= s - 1;
    }
    return r;
}

// One-string part of CIMethod::diagonal: get_diag_cont_one_bitstring
// (CIMethod.cpp:329-341), same summation order (h terms, then pairs k<l).
__device__ __forceinline__ double diag_one_string(str_t s, const HamDev& H) {
    const int L = H.L;
    double acc = 0.0;
    for (str_t t = s; t; t &= t - 1) {
        int k = __builtin_ctzll(t);
        acc += H.h[k * L + k];
    }
    for (str_t t = s; t; t &= t - 1) {
        int k = __builtin_ctzll(t);
        for (str_t u = t & (t - 1); u; u &= u - 1) {
            int l = __builtin_ctzll(u);
            const size_t kl = ((size_t)k * L + l) * L;
            acc += H.V[(kl + k) * L + l] - H.V[(kl + l) * L + k];
        }
    }
    return acc;
}

// get_nstacked (CIMethod.cpp:382-396): common occupied orbitals strictly
// between p and q.
__device__ __forceinline__ int stacked(str_t common, int p, int q) {
    int lo = p < q ? p : q, hi = p < q ? q : p;
    str_t mask = (((str_t)1 << hi) - 1) & ~(((str_t)1 << lo) - 1);
    return __popcll(common & mask);
}

}  // namespace cf
