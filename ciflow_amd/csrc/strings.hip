// Determinant-string kernels: ranking/unranking, one-string diagonal parts and
// sorted per-string excitation lists.
//
// Ranking replaces Permutator_Bit + determine_weight (Permutator.cpp:113-153,
// CIMethod.cpp:663-747): the reference enumerates n-of-L bitstrings in
// increasing integer order (Gosper's next-combination) and indexes them with
// vertex weights vw[k][m] = C(k,m).  That order is the combinatorial number
// system, rank(s) = sum_m C(k_m, m) with k_m the position of the m-th set bit,
// so every lane ranks or unranks its own string directly with popcount/ctz and
// a binomial table: no serial Gosper walk and no start_l skipping
// (CIMethod.cpp:619-620, 763-764).
#include "kernels.h"
#include "string_ops.h"

namespace cf {

__device__ __forceinline__ uint64_t dbinom(const uint64_t* binom, int k, int m) { return binom[k * kBinomCols + m]; }

__device__ __forceinline__ str_t unrank_bits(uint64_t r, int L, int nel, const uint64_t* binom) {
    str_t s = 0;
    int hi = L;  // bits of lower m lie strictly below the previous choice
    for (int m = nel; m >= 1; --m) {
        int k = m - 1;
        while (k + 1 < hi && dbinom(binom, k + 1, m) <= r) ++k;
        s |= (str_t)1 << k;
        r -= dbinom(binom, k, m);
        hi = k;
    }
    return s;
}

__global__ void k_unrank(int L, int nel, uint64_t first, uint64_t count, const uint64_t* __restrict__ binom,
                         str_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = unrank_bits(first + i, L, nel, binom);
}

__global__ void k_rank(const str_t* __restrict__ s, uint64_t count, const uint64_t* __restrict__ binom,
                       uint32_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) out[i] = rank_bits(s[i], binom);
}

void launch_unrank(int L, int nel, uint64_t first, uint64_t count, const uint64_t* binom, str_t* out,
                   hipStream_t stream) {
    if (!count) return;
    unsigned grid = (unsigned)((count + 255) / 256);
    k_unrank<<<grid, 256, 0, stream>>>(L, nel, first, count, binom, out);
    CF_HIP(hipGetLastError());
}

void launch_rank(int L, int nel, const str_t* strings, uint64_t count, const uint64_t* binom, uint32_t* out,
                 hipStream_t stream) {
    (void)L;
    (void)nel;
    if (!count) return;
    unsigned grid = (unsigned)((count + 255) / 256);
    k_rank<<<grid, 256, 0, stream>>>(strings, count, binom, out);
    CF_HIP(hipGetLastError());
}

__global__ void k_diag1(int nel, const str_t* __restrict__ strs, uint64_t d, HamDev H, double* __restrict__ out) {
    (void)nel;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d) out[i] = diag_one_string(strs[i], H);
}

void launch_diag1(int nel, const str_t* strs, uint64_t d, HamDev H, double* out, hipStream_t stream) {
    if (!d) return;
    k_diag1<<<(unsigned)((d + 255) / 256), 256, 0, stream>>>(nel, strs, d, H, out);
    CF_HIP(hipGetLastError());
}


// Excitation lists.  One 64-lane workgroup per string: enumerate self,
// singles (occ f -> virt t) and same-spin doubles, bitonic-sort them in LDS by
// target bitstring (= target rank), then evaluate each entry with the
// reference's element formulas:
//   single: one_diff_orbital without the other-spin sum (CIMethod.cpp:409-427)
//   double: two_diff_orbitals(up1, up2) with up1 = the smaller string, i.e.
//           the orientation in which the reference evaluates the element for
//           its upper triangle (CIMethod.cpp:429-442, 808-812, 850-854).
__global__ void __launch_bounds__(64) k_exc_lists(int L, int nel, const str_t* __restrict__ strs,
                                                  const uint64_t* __restrict__ binom, HamDev H,
                                                  const double* __restrict__ diag1, int E, int S, int P,
                                                  Exc* __restrict__ full, Exc* __restrict__ sub) {
    extern __shared__ unsigned char smem[];
    uint64_t* key = (uint64_t*)smem;
    uint32_t* code = (uint32_t*)(key + P);
    __shared__ int occ[kMaxL];
    __shared__ int vir[kMaxL];
    const uint64_t idx = blockIdx.x;
    const int lane = threadIdx.x;
    const str_t I = strs[idx];
    const int nv = L - nel;
    if (lane < L) {
        str_t below = I & (((str_t)1 << lane) - 1);
        if ((I >> lane) & 1) occ[__popcll(below)] = lane;
        else vir[lane - __popcll(below)] = lane;
    }
    __syncthreads();
    const int ns = nel * nv;
    const int cv2 = nv * (nv - 1) / 2;
    for (int e = lane; e < P; e += 64) {
        uint64_t k = ~0ull;
        uint32_t c = 0xffffffffu;
        if (e == 0) {
            k = I;
            c = 0;
        } else if (e < 1 + ns) {
            int s = e - 1;
            int f = occ[s / nv], t = vir[s % nv];
            k = I ^ ((str_t)1 << f) ^ ((str_t)1 << t);
            c = 1u | (uint32_t)f << 2 | (uint32_t)t << 8;
        } else if (e < E) {
            int q = e - 1 - ns;
            int po = q / cv2, pv = q % cv2;
            int a = 0;
            while (po >= nel - 1 - a) { po -= nel - 1 - a; ++a; }
            int b = a + 1 + po;
            int x = 0;
            while (pv >= nv - 1 - x) { pv -= nv - 1 - x; ++x; }
            int y = x + 1 + pv;
            int f1 = occ[a], f2 = occ[b], t1 = vir[x], t2 = vir[y];
            k = I ^ ((str_t)1 << f1) ^ ((str_t)1 << f2) ^ ((str_t)1 << t1) ^ ((str_t)1 << t2);
            c = 2u | (uint32_t)f1 << 2 | (uint32_t)f2 << 8 | (uint32_t)t1 << 14 | (uint32_t)t2 << 20;
        }
        key[e] = k;
        code[e] = c;
    }
    __syncthreads();
    // bitonic sort, ascending
    for (int kk = 2; kk <= P; kk <<= 1) {
        for (int j = kk >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < P; i += 64) {
                int ixj = i ^ j;
                if (ixj > i) {
                    bool up = (i & kk) == 0;
                    uint64_t ki = key[i], kj = key[ixj];
                    if ((ki > kj) == up) {
                        key[i] = kj;
                        key[ixj] = ki;
                        uint32_t t = code[i];
                        code[i] = code[ixj];
                        code[ixj] = t;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int Lc = L;
    Exc* out = full + idx * (uint64_t)E;
    for (int e = lane; e < E; e += 64) {
        const str_t J = key[e];
        const uint32_t c = code[e];
        Exc x;
        x.J = rank_bits(J, binom);
        x.type = (uint8_t)(c & 3u);
        x.sign = 1;
        x.f = 0;
        x.t = 0;
        x.x = 0.0;
        const str_t common = I & J;
        if (x.type == 0) {
            x.x = diag1[idx];
        } else if (x.type == 1) {
            int f = (int)((c >> 2) & 63u), t = (int)((c >> 8) & 63u);
            int p = f < t ? f : t, q = f < t ? t : f;
            double s = H.h[p * Lc + q];
            for (str_t u = common; u; u &= u - 1) {
                int z = __builtin_ctzll(u);
                s += H.V[(((size_t)p * Lc + z) * Lc + q) * Lc + z] - H.V[(((size_t)p * Lc + z) * Lc + z) * Lc + q];
            }
            x.f = (uint8_t)f;
            x.t = (uint8_t)t;
            x.sign = (stacked(common, p, q) & 1) ? -1 : 1;
            x.x = s;
        } else {
            int f1 = (int)((c >> 2) & 63u), f2 = (int)((c >> 8) & 63u);
            int t1 = (int)((c >> 14) & 63u), t2 = (int)((c >> 20) & 63u);
            int a, b, cc, d;
            if (I < J) { a = f1; b = f2; cc = t1; d = t2; }
            else { a = t1; b = t2; cc = f1; d = f2; }
            int diff = stacked(common, a, b) + stacked(common, cc, d);
            double v = H.V[(((size_t)a * Lc + b) * Lc + cc) * Lc + d] - H.V[(((size_t)a * Lc + b) * Lc + d) * Lc + cc];
            if (diff & 1) v *= -1;
            x.x = v;
        }
        out[e] = x;
    }
    __syncthreads();
    // self + singles sub-list, keeps the sorted order
    if (lane == 0) {
        Exc* so = sub + idx * (uint64_t)S;
        int w = 0;
        for (int e = 0; e < E && w < S; ++e)
            if (out[e].type != 2) so[w++] = out[e];
    }
}

void launch_exc_lists(int L, int nel, const str_t* strs, uint64_t d, const uint64_t* binom, HamDev H,
                      const double* diag1, int E, int S, Exc* full, Exc* sub, hipStream_t stream) {
    int P = 1;
    while (P < E) P <<= 1;
    if (P < 64) P = 64;
    size_t lds = (size_t)P * (sizeof(uint64_t) + sizeof(uint32_t));
    if (lds > 96 * 1024) throw Error(ERR_UNSUPPORTED, "excitation list too long for LDS sort");
    if (!d) return;
    k_exc_lists<<<(unsigned)d, 64, lds, stream>>>(L, nel, strs, binom, H, diag1, E, S, P, full, sub);
    CF_HIP(hipGetLastError());
}

}  // namespace cf
