// DOCI SpMV on position-major pair-hop columns (replaces SparseMatrix_CRS_Sym::mvprod,
// SparseMatrix_CRS.cpp:700-723, for DOCI matrices built by DOCI::construct_CI_matrix,
// CIMethod.cpp:611-660).
//
// A DOCI row is one pair string s; its off-diagonal elements are the hops of a
// pair from an occupied orbital j to an empty orbital k, with element
// V(m, m, M, M) (m = min(j, k), M = max(j, k)) and phase +1, so the value
// depends on the hop only.  The CRS stores 12 (FP64) or 8 (FP32) bytes per
// element; here a thread per row walks its np * nv hops in a fixed order (a-th
// occupied orbital from the top, b-th empty orbital from the top), takes the
// value from an L x L pair table in LDS and the target rank from a
// position-major column array, col[t * n + i], read as one contiguous
// wavefront load per slot.  Elements the CRS build drops (|v| <= zero) are
// zero in the table, so the product equals the CRS product up to summation
// order.  Rows in rank order share their high orbitals, so a slot's targets
// across a wavefront fall into a few nearby x segments.
//
// HBM-bound.  Algorithmic bytes per call:
//   B = n * (4 * np * nv + 8 (string) + 8 (diagonal) + 8 (x once) + 8 (y)).
#include "spmv_doci.h"
#include "string_ops.h"

namespace cf {

__global__ void __launch_bounds__(256) k_doci_ell_fill(DociEllArgs a, const uint64_t* __restrict__ binom,
                                                       uint32_t* __restrict__ col) {
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= a.n) return;
    const str_t full = a.L == 64 ? ~0ull : (((str_t)1 << a.L) - 1);
    const str_t s = a.s[i], vir = ~s & full;
    uint64_t t = 0;
    str_t o = s;
    for (int p = 0; p < a.np; ++p) {
        const int j = 63 - __builtin_clzll(o);
        o ^= (str_t)1 << j;
        str_t v = vir;
        for (int q = 0; q < a.nv; ++q) {
            const int k = 63 - __builtin_clzll(v);
            v ^= (str_t)1 << k;
            col[t * a.n + i] = rank_bits(s ^ ((str_t)1 << j) ^ ((str_t)1 << k), binom);
            ++t;
        }
    }
}

void launch_doci_ell_fill(const DociEllArgs& a, const uint64_t* binom, uint32_t* col, hipStream_t s) {
    if (!a.n || !a.np || !a.nv) return;
    k_doci_ell_fill<<<(unsigned)((a.n + 255) / 256), 256, 0, s>>>(a, binom, col);
    CF_HIP(hipGetLastError());
}

__global__ void k_doci_pair_table(int L, HamDev H, double zero, double* __restrict__ vk64,
                                  float* __restrict__ vk32) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L * L) return;
    const int j = t / L, k = t - j * L;
    const size_t lo = j < k ? j : k, hi = j < k ? k : j;
    double v = H.V[((lo * L + lo) * L + hi) * L + hi];
    if (!(fabs(v) > zero)) v = 0.0;
    vk64[t] = v;
    vk32[t] = (float)v;
}

void launch_doci_pair_table(const DociEllArgs& a, HamDev H, double* vk64, float* vk32, hipStream_t s) {
    const int LL = a.L * a.L;
    k_doci_pair_table<<<(LL + 255) / 256, 256, 0, s>>>(a.L, H, a.zero, vk64, vk32);
    CF_HIP(hipGetLastError());
}

// SUB lanes share a row, lane `sub` taking the occupied orbitals a = sub,
// sub + SUB, ... (small spaces need several lanes per row to fill the GPU);
// U slots per batch: U column loads, then U x gathers, then U FMAs, so each
// lane keeps U independent gathers in flight.
template <typename VT, int SUB, int U>
__global__ void __launch_bounds__(1024) k_spmv_doci(DociEllArgs a, const VT* __restrict__ vk,
                                                    const double* __restrict__ x, double* __restrict__ y) {
    extern __shared__ double smem_doci[];
    VT* tab = reinterpret_cast<VT*>(smem_doci);
    const int L = a.L;
    for (int t = threadIdx.x; t < L * L; t += blockDim.x) tab[t] = vk[t];
    __syncthreads();
    const uint64_t n = a.n;
    const uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / SUB;
    const int sub = (int)(threadIdx.x % SUB);
    const bool live = i < n;
    double acc = 0.0;
    if (live) {
        const str_t full = L == 64 ? ~0ull : (((str_t)1 << L) - 1);
        const str_t s = a.s[i], vir = ~s & full;
        if (sub == 0) {
            double d = a.diag[i];
            if (!(fabs(d) > a.zero)) d = 0.0;
            acc = (double)(VT)d * x[i];
        }
        const int nv = a.nv;
        const uint32_t na = a.np > sub ? (uint32_t)((a.np - sub + SUB - 1) / SUB) : 0u;
        const uint32_t R1 = na * (uint32_t)nv;
        const uint32_t* cp = a.col + i;
        str_t o = s, v = 0;
        for (int q = 0; q < sub && o; ++q) o ^= (str_t)1 << (63 - __builtin_clzll(o));
        int j = 0;
        uint32_t slot = 0;  // a * nv + b of the next hop
        for (uint32_t t0 = 0; t0 < R1; t0 += U) {
            uint32_t c[U];
            double w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                c[u] = 0;
                w[u] = 0.0;
                if (t0 + u < R1) {
                    if (!v) {  // next occupied orbital of this lane: a += SUB
                        if (t0 + u) {
                            slot += (uint32_t)(SUB - 1) * (uint32_t)nv;
                            for (int q = 1; q < SUB && o; ++q) o ^= (str_t)1 << (63 - __builtin_clzll(o));
                        } else {
                            slot = (uint32_t)sub * (uint32_t)nv;
                        }
                        j = 63 - __builtin_clzll(o);
                        o ^= (str_t)1 << j;
                        v = vir;
                    }
                    const int k = 63 - __builtin_clzll(v);
                    v ^= (str_t)1 << k;
                    w[u] = (double)tab[j * L + k];
                    c[u] = __builtin_nontemporal_load(cp + (uint64_t)slot * n);
                    ++slot;
                }
            }
            double xv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) xv[u] = t0 + u < R1 ? x[c[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < U; ++u) acc += w[u] * xv[u];
        }
    }
#pragma unroll
    for (int off = SUB / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, SUB);
    if (live && sub == 0) y[i] = acc;
}

DociShape doci_default_shape(uint64_t n) {
    // enough lanes for ~2 waves per SIMD on 256 CUs: 131,072 lanes
    DociShape sh;
    sh.threads = 256;
    sh.sub = 1;
    while (sh.sub < 8 && n * (uint64_t)sh.sub < 131072) sh.sub *= 2;
    return sh;
}

template <typename VT>
static void launch_doci(const DociEllArgs& a, DociShape sh, const VT* vk, const double* x, double* y,
                        hipStream_t s) {
    if (!a.n) return;
    const int T = sh.threads;
    const uint64_t lanes = a.n * (uint64_t)sh.sub;
    const uint64_t nb = (lanes + T - 1) / T;
    CF_REQUIRE(nb < (1ull << 31), "too many rows for one DOCI SpMV launch");
    const size_t shm = (size_t)a.L * a.L * sizeof(VT);
    switch (sh.sub) {
        case 1: k_spmv_doci<VT, 1, 8><<<(uint32_t)nb, T, shm, s>>>(a, vk, x, y); break;
        case 2: k_spmv_doci<VT, 2, 8><<<(uint32_t)nb, T, shm, s>>>(a, vk, x, y); break;
        case 4: k_spmv_doci<VT, 4, 8><<<(uint32_t)nb, T, shm, s>>>(a, vk, x, y); break;
        case 8: k_spmv_doci<VT, 8, 8><<<(uint32_t)nb, T, shm, s>>>(a, vk, x, y); break;
        default: CF_REQUIRE(false, "DOCI SpMV: lanes per row must be 1, 2, 4 or 8");
    }
    CF_HIP(hipGetLastError());
}

void launch_spmv_doci_f64(const DociEllArgs& a, DociShape sh, const double* vk, const double* x, double* y,
                          hipStream_t s) {
    launch_doci<double>(a, sh, vk, x, y, s);
}

void launch_spmv_doci_f32(const DociEllArgs& a, DociShape sh, const float* vk, const double* x, double* y,
                          hipStream_t s) {
    launch_doci<float>(a, sh, vk, x, y, s);
}

}  // namespace cf
