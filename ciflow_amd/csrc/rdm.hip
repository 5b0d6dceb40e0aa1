// 1- and 2-RDMs as reductions over the string excitation lists.
//
// The reference walks determinant pairs: DensFCI::construct_CI_two_dens is an
// O(n^2/2) pair loop with a popcount classification per pair
// (CIDens.cpp:1582-1631, 855-881), DensDOCI loops over pair hops
// (:1165-1224).  Here:
//
// FCI.  For every determinant K and spin s the one-particle transition
// densities X^s_ab(K) = <K| a+_as a_bs |Psi> are gathered from the single-
// excitation lists of K's strings (diagonal: n_a(K) c_K; a != b: the single
// a -> b of K's s-string, with its phase, times c of the target).  Then
//   gamma^s_ps              = sum_K c_K X^s_ps(K)
//   Gamma^{ab}(pq,rs)       = sum_K X^a_rp(K) X^b_qs(K)
//   Gamma^{ss}(pq,rs)       = sum_K X^s_rp(K) X^s_qs(K) - delta_qr gamma^s_ps
// (from a+_p a+_q a_s a_r = E_pr E_qs - delta_qr E_ps, and for opposite spins
// a+_pa a+_qb a_sb a_ra = E^a_pr E^b_qs), i.e. three Gram products X^T X
// reduced over K in fixed-order chunks (deterministic).
//
// DOCI.  gamma_jj = sum_I c_I^2 n_j(I); D_jk = sum_I c_I^2 n_j n_k;
// P_jk = sum_{J: k in J, j not in J} c_{J-k+j} c_J; each thread owns one (j,k)
// bin and reduces over a chunk of determinants, chunks summed in order.
#include "rdm.h"
#include "string_ops.h"

#include <hipcub/hipcub.hpp>
#include <algorithm>

#include <vector>

namespace cf {

// X^a and X^b rows of determinant K (row-major n x L^2).  Rows must be zeroed.
__global__ void k_build_x(uint64_t da, uint64_t db, int L, const str_t* __restrict__ sa, const str_t* __restrict__ sb,
                          const Exc* __restrict__ suba, const Exc* __restrict__ subb, int Sa, int Sb,
                          const double* __restrict__ c, double* __restrict__ Xa, double* __restrict__ Xb) {
    const uint64_t K = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (K >= da * db) return;
    const uint64_t ka = K / db, kb = K - ka * db;
    const size_t P = (size_t)L * L;
    double* xa = Xa + K * P;
    double* xb = Xb + K * P;
    const double cK = c[K];
    for (str_t s = sa[ka]; s; s &= s - 1) {
        int a = __builtin_ctzll(s);
        xa[a * L + a] = cK;
    }
    for (str_t s = sb[kb]; s; s &= s - 1) {
        int a = __builtin_ctzll(s);
        xb[a * L + a] = cK;
    }
    const Exc* la = suba + ka * (uint64_t)Sa;
    for (int e = 0; e < Sa; ++e) {
        const Exc x = la[e];
        if (x.type != 1) continue;
        double v = c[(uint64_t)x.J * db + kb];
        xa[x.f * L + x.t] = x.sign < 0 ? -v : v;
    }
    const Exc* lb = subb + kb * (uint64_t)Sb;
    for (int e = 0; e < Sb; ++e) {
        const Exc x = lb[e];
        if (x.type != 1) continue;
        double v = c[ka * db + x.J];
        xb[x.f * L + x.t] = x.sign < 0 ? -v : v;
    }
}

// part[chunk][p] = sum_{K in chunk} c_K X[K][p]
__global__ void k_gamma_part(uint64_t n, int P, const double* __restrict__ c, const double* __restrict__ X,
                             uint64_t rows, double* __restrict__ part) {
    const uint64_t k0 = (uint64_t)blockIdx.x * rows;
    uint64_t k1 = k0 + rows;
    if (k1 > n) k1 = n;
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
        double acc = 0.0;
        for (uint64_t K = k0; K < k1; ++K) acc += c[K] * X[K * P + p];
        part[(uint64_t)blockIdx.x * P + p] = acc;
    }
}

// Split-K Gram product: part[chunk][p][q] = sum_{K in chunk} A[K][p] B[K][q].
constexpr int kGT = 32;  // output tile
constexpr int kGK = 32;  // K rows per LDS stage
__global__ void __launch_bounds__(256) k_gram(uint64_t n, int P, const double* __restrict__ A,
                                              const double* __restrict__ B, uint64_t rows, double* __restrict__ part) {
    __shared__ double As[kGK][kGT + 1];
    __shared__ double Bs[kGK][kGT + 1];
    const int nt = (P + kGT - 1) / kGT;
    const int tp = blockIdx.y / nt, tq = blockIdx.y % nt;
    const int p0 = tp * kGT, q0 = tq * kGT;
    const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
    const uint64_t k0 = (uint64_t)blockIdx.x * rows;
    uint64_t k1 = k0 + rows;
    if (k1 > n) k1 = n;
    double acc00 = 0, acc01 = 0, acc10 = 0, acc11 = 0;
    for (uint64_t kb = k0; kb < k1; kb += kGK) {
        for (int i = threadIdx.x; i < kGK * kGT; i += 256) {
            const int kk = i / kGT, pp = i % kGT;
            const uint64_t K = kb + kk;
            const bool ok = K < k1;
            As[kk][pp] = (ok && p0 + pp < P) ? A[K * P + p0 + pp] : 0.0;
            Bs[kk][pp] = (ok && q0 + pp < P) ? B[K * P + q0 + pp] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int kk = 0; kk < kGK; ++kk) {
            const double a0 = As[kk][ty * 2], a1 = As[kk][ty * 2 + 1];
            const double b0 = Bs[kk][tx * 2], b1 = Bs[kk][tx * 2 + 1];
            acc00 += a0 * b0;
            acc01 += a0 * b1;
            acc10 += a1 * b0;
            acc11 += a1 * b1;
        }
        __syncthreads();
    }
    const int p = p0 + ty * 2, q = q0 + tx * 2;
    double* o = part + (uint64_t)blockIdx.x * P * P;
    if (p < P && q < P) o[p * P + q] = acc00;
    if (p < P && q + 1 < P) o[p * P + q + 1] = acc01;
    if (p + 1 < P && q < P) o[(p + 1) * P + q] = acc10;
    if (p + 1 < P && q + 1 < P) o[(p + 1) * P + q + 1] = acc11;
}

// The same split-K Gram product on FP64 MFMA (v_mfma_f64_16x16x4f64), for
// P = L^2 <= 256.  A workgroup of 16 wavefronts owns the P x P output of its
// chunk of K rows (or one of `gridDim.y` interleaved parts of it, so that a
// wavefront holds at most TPW = 8 accumulator tiles): 32 rows of A and B are
// staged in LDS per step and every wavefront accumulates its 16 x 16 output
// tiles (tile = 16 (y + gridDim.y t) + wave), so each element of A and B is
// read from HBM gridDim.y times per product.
// 2 n P^2 flop per product.
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int kGramRows = 32;

template <int TPW>
__global__ void __launch_bounds__(1024) k_gram_mfma(uint64_t n, int P, const double* __restrict__ A,
                                                    const double* __restrict__ B, uint64_t rows,
                                                    double* __restrict__ part) {
    extern __shared__ double gsm[];
    const int SP = P + 1;
    double* As = gsm;
    double* Bs = gsm + kGramRows * SP;
    const int T16 = (P + 15) / 16, ntiles = T16 * T16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, kq = lane >> 4;
    v4d acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = v4d{0.0, 0.0, 0.0, 0.0};
    const uint64_t k0 = (uint64_t)blockIdx.x * rows;
    uint64_t k1 = k0 + rows;
    if (k1 > n) k1 = n;
    for (uint64_t kb = k0; kb < k1; kb += kGramRows) {
        for (int i = threadIdx.x; i < kGramRows * P; i += 1024) {
            const int kk = i / P, pp = i - kk * P;
            const uint64_t K = kb + kk;
            const bool ok = K < k1;
            As[kk * SP + pp] = ok ? A[K * P + pp] : 0.0;
            Bs[kk * SP + pp] = ok ? B[K * P + pp] : 0.0;
        }
        __syncthreads();
        for (int ks = 0; ks < kGramRows; ks += 4) {
            const double* ar = As + (ks + kq) * SP;
            const double* br = Bs + (ks + kq) * SP;
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int tile = 16 * ((int)blockIdx.y + (int)gridDim.y * t) + wave;
                if (tile < ntiles) {
                    const int ti = tile / T16, tj = tile - ti * T16;
                    const int p = ti * 16 + r, q = tj * 16 + r;
                    const double av = p < P ? ar[p] : 0.0;  // A^T[p][K]
                    const double bv = q < P ? br[q] : 0.0;  // B[K][q]
                    acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
    double* o = part + (uint64_t)blockIdx.x * P * P;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = 16 * ((int)blockIdx.y + (int)gridDim.y * t) + wave;
        if (tile < ntiles) {
            const int ti = tile / T16, tj = tile - ti * T16;
#pragma unroll
            for (int v = 0; v < 4; ++v) {  // D[i][j]: lane j + 16 (i % 4), element i / 4
                const int p = ti * 16 + 4 * v + kq, q = tj * 16 + r;
                if (p < P && q < P) o[(uint64_t)p * P + q] = acc[t][v];
            }
        }
    }
}

// part[chunk] = A^T B over the chunk's rows; MFMA for P <= 256, else k_gram.
static void launch_gram(uint64_t n, int P, const double* A, const double* B, uint64_t rows, uint64_t nchunk,
                        double* part, hipStream_t stream) {
    if (P <= 256) {
        const int T16 = (P + 15) / 16, per = (T16 * T16 + 15) / 16;
        const int split = (per + 7) / 8;  // parts of the output per chunk
        const int tpw = (per + split - 1) / split;
        const size_t lds = 2 * (size_t)kGramRows * (P + 1) * sizeof(double);
        const dim3 grid((unsigned)nchunk, (unsigned)split);
#define CF_GRAM(T) k_gram_mfma<T><<<grid, 1024, lds, stream>>>(n, P, A, B, rows, part)
        if (tpw <= 1) CF_GRAM(1);
        else if (tpw <= 2) CF_GRAM(2);
        else if (tpw <= 4) CF_GRAM(4);
        else CF_GRAM(8);
#undef CF_GRAM
    } else {
        const int nt = (P + kGT - 1) / kGT;
        dim3 grid((unsigned)nchunk, (unsigned)(nt * nt));
        k_gram<<<grid, 256, 0, stream>>>(n, P, A, B, rows, part);
    }
    CF_HIP(hipGetLastError());
}

// out[i] = sum_chunk part[chunk][i]
__global__ void k_sum_chunks(int nchunk, uint64_t len, const double* __restrict__ part, double* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += part[(uint64_t)c * len + i];
    out[i] = s;
}

// DOCI bins.  part layout per chunk: [0, L^2) D, [L^2, 2L^2) P, [2L^2, 2L^2+L) gamma.
__global__ void k_doci_rdm(uint64_t d, int L, const str_t* __restrict__ s, const uint64_t* __restrict__ binom,
                           const double* __restrict__ c, uint64_t rows, bool two, double* __restrict__ part) {
    const uint64_t i0 = (uint64_t)blockIdx.x * rows;
    uint64_t i1 = i0 + rows;
    if (i1 > d) i1 = d;
    const int L2 = L * L;
    double* o = part + (uint64_t)blockIdx.x * (2 * L2 + L);
    for (int bin = threadIdx.x; bin < L2; bin += blockDim.x) {
        const int j = bin / L, k = bin % L;
        double D = 0.0, Pj = 0.0;
        for (uint64_t I = i0; I < i1; ++I) {
            const str_t x = s[I];
            const bool hj = (x >> j) & 1, hk = (x >> k) & 1;
            const double cI = c[I];
            if (hj && hk) D += cI * cI;
            if (two && j != k && hk && !hj) {
                // J = x holds k, the partner I' = x - k + j holds j
                const str_t y = x ^ ((str_t)1 << k) ^ ((str_t)1 << j);
                Pj += c[rank_bits(y, binom)] * cI;
            }
        }
        o[bin] = D;
        o[L2 + bin] = Pj;
        if (j == k) o[2 * L2 + j] = D;
    }
}

// gamma and Gamma from the transition densities X^a, X^b (n x L^2 rows) and
// the coefficients c_K of the same n determinants (0 for a determinant
// outside the CI space): gamma^s = sum_K c_K X^s(K), Gamma = X^T X Gram
// products minus the contraction term, chunks reduced in fixed order.
static void rdm_from_x(hipStream_t stream, int L, uint64_t n, const double* c, const double* Xa, const double* Xb,
                       double* gamma, double* Gamma) {
    const int P = L * L;
    uint64_t nchunk = (n + 1023) / 1024;
    if (nchunk > 512) nchunk = 512;
    if (nchunk < 1) nchunk = 1;
    const uint64_t rows = (n + nchunk - 1) / nchunk;
    {
        DBuf<double> part(nchunk * (uint64_t)P), tot(2 * (uint64_t)P);
        for (int s = 0; s < 2; ++s) {
            k_gamma_part<<<(unsigned)nchunk, 256, 0, stream>>>(n, P, c, s == 0 ? Xa : Xb, rows, part.p);
            CF_HIP(hipGetLastError());
            k_sum_chunks<<<(unsigned)((P + 255) / 256), 256, 0, stream>>>((int)nchunk, P, part.p, tot.p + s * P);
            CF_HIP(hipGetLastError());
        }
        CF_HIP(hipMemcpyAsync(gamma, tot.p, 2 * P * sizeof(double), hipMemcpyDeviceToHost, stream));
        CF_HIP(hipStreamSynchronize(stream));
    }
    if (!Gamma) return;
    const uint64_t PP = (uint64_t)P * P;
    DBuf<double> part(nchunk * PP), G(3 * PP);
    const double* ops[3][2] = {{Xa, Xa}, {Xa, Xb}, {Xb, Xb}};
    for (int b = 0; b < 3; ++b) {
        launch_gram(n, P, ops[b][0], ops[b][1], rows, nchunk, part.p, stream);
        k_sum_chunks<<<(unsigned)((PP + 255) / 256), 256, 0, stream>>>((int)nchunk, PP, part.p, G.p + b * PP);
        CF_HIP(hipGetLastError());
    }
    std::vector<double> hG(3 * PP);
    CF_HIP(hipMemcpyAsync(hG.data(), G.p, 3 * PP * sizeof(double), hipMemcpyDeviceToHost, stream));
    CF_HIP(hipStreamSynchronize(stream));
    const size_t L4 = PP;
    for (int b = 0; b < 3; ++b) {
        const double* Gb = hG.data() + b * PP;
        const int s = b == 2 ? 1 : 0;
        for (int p = 0; p < L; ++p)
            for (int q = 0; q < L; ++q)
                for (int r = 0; r < L; ++r)
                    for (int t = 0; t < L; ++t) {
                        double v = Gb[(size_t)(r * L + p) * P + (q * L + t)];
                        if (b != 1 && q == r) v -= gamma[s * P + p * L + t];
                        Gamma[b * L4 + (((size_t)p * L + q) * L + r) * L + t] = v;
                    }
    }
}

void rdm_dense(hipStream_t stream, int kind, int L, int nup, int ndown, uint64_t da, uint64_t db, const str_t* sa,
               const str_t* sb, const Exc* suba, const Exc* subb, int Sa, int Sb, const uint64_t* binom,
               const double* c_host, uint64_t n, double* gamma, double* Gamma) {
    (void)nup;
    (void)ndown;
    const int P = L * L;
    DBuf<double> c(n);
    CF_HIP(hipMemcpyAsync(c.p, c_host, n * sizeof(double), hipMemcpyHostToDevice, stream));
    std::vector<double> g(2 * (size_t)P, 0.0);
    if (kind == 1) {  // DOCI
        uint64_t nchunk = (n + 2047) / 2048;
        if (nchunk > 1024) nchunk = 1024;
        if (nchunk < 1) nchunk = 1;
        const uint64_t rows = (n + nchunk - 1) / nchunk;
        const uint64_t len = 2 * (uint64_t)P + L;
        DBuf<double> part(nchunk * len), tot(len);
        int thr = P < 1024 ? ((P + 63) / 64) * 64 : 1024;
        k_doci_rdm<<<(unsigned)nchunk, thr, 0, stream>>>(n, L, sa, binom, c.p, rows, Gamma != nullptr, part.p);
        CF_HIP(hipGetLastError());
        k_sum_chunks<<<(unsigned)((len + 255) / 256), 256, 0, stream>>>((int)nchunk, len, part.p, tot.p);
        CF_HIP(hipGetLastError());
        std::vector<double> h(len);
        CF_HIP(hipMemcpyAsync(h.data(), tot.p, len * sizeof(double), hipMemcpyDeviceToHost, stream));
        CF_HIP(hipStreamSynchronize(stream));
        for (int j = 0; j < L; ++j) {
            gamma[j * L + j] = h[2 * P + j];
            gamma[P + j * L + j] = h[2 * P + j];
        }
        for (int p = 0; p < L; ++p)
            for (int q = 0; q < L; ++q)
                if (p != q) gamma[p * L + q] = gamma[P + p * L + q] = 0.0;
        if (Gamma) {
            const size_t L4 = (size_t)P * P;
            std::fill(Gamma, Gamma + 3 * L4, 0.0);
            auto at = [&](int b, int i, int j, int k, int l) -> double& {
                return Gamma[b * L4 + (((size_t)i * L + j) * L + k) * L + l];
            };
            for (int j = 0; j < L; ++j)
                for (int k = 0; k < L; ++k) {
                    const double D = h[j * L + k];
                    at(1, j, k, j, k) = D;
                    if (j != k) {
                        at(0, j, k, j, k) = D;
                        at(0, j, k, k, j) = -D;
                        at(2, j, k, j, k) = D;
                        at(2, j, k, k, j) = -D;
                        at(1, j, j, k, k) = h[P + j * L + k];
                    }
                }
        }
        return;
    }
    // FCI
    DBuf<double> Xa(n * (uint64_t)P), Xb(n * (uint64_t)P);
    CF_HIP(hipMemsetAsync(Xa.p, 0, Xa.bytes(), stream));
    CF_HIP(hipMemsetAsync(Xb.p, 0, Xb.bytes(), stream));
    k_build_x<<<(unsigned)((n + 127) / 128), 128, 0, stream>>>(da, db, L, sa, sb, suba, subb, Sa, Sb, c.p, Xa.p, Xb.p);
    CF_HIP(hipGetLastError());
    rdm_from_x(stream, L, n, c.p, Xa.p, Xb.p, gamma, Gamma);
}

// ---------------------------------------------------------------------------
// Determinant lists (DensFILE, CIDens.cpp:1468-1526).  The same transition
// densities and Gram products as for FCI, over the determinants K that one
// single excitation reaches from the list (the list included): K is keyed
// up | down << L, the keys are sorted and made unique (hipCUB), and every
// list determinant J scatters c_J into X^s[K][a*L + b] for K = J - b + a
// (b occupied in J's s-string, a = b or empty), with the single's phase.
// (K, s, a, b) fixes J, so the scatter has no conflicts and is deterministic.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t det_key(str_t u, str_t d, int L) { return (uint64_t)u | ((uint64_t)d << L); }

// keys[J*E + e]: J itself, then its alpha singles, then its beta singles.
__global__ void k_det_closure_keys(uint64_t n, int L, int E, const str_t* __restrict__ up,
                                   const str_t* __restrict__ down, uint64_t* __restrict__ keys) {
    const uint64_t J = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (J >= n) return;
    const str_t u = up[J], d = down[J];
    const str_t full = L >= 64 ? ~(str_t)0 : (((str_t)1 << L) - 1);
    uint64_t* o = keys + J * (uint64_t)E;
    int e = 0;
    o[e++] = det_key(u, d, L);
    for (str_t t = u; t; t &= t - 1) {
        const str_t bb = t & (~t + 1);
        for (str_t v = full & ~u; v; v &= v - 1) o[e++] = det_key((u ^ bb) | (v & (~v + 1)), d, L);
    }
    for (str_t t = d; t; t &= t - 1) {
        const str_t bb = t & (~t + 1);
        for (str_t v = full & ~d; v; v &= v - 1) o[e++] = det_key(u, (d ^ bb) | (v & (~v + 1)), L);
    }
}

__device__ __forceinline__ uint64_t find_key(const uint64_t* __restrict__ ks, uint64_t nk, uint64_t key) {
    uint64_t lo = 0, hi = nk;  // first index with ks[i] >= key
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (ks[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_det_scatter_x(uint64_t n, int L, const str_t* __restrict__ up, const str_t* __restrict__ down,
                                const double* __restrict__ c, const uint64_t* __restrict__ ks, uint64_t nk,
                                double* __restrict__ cK, double* __restrict__ Xa, double* __restrict__ Xb) {
    const uint64_t J = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (J >= n) return;
    const str_t u = up[J], d = down[J];
    const str_t full = L >= 64 ? ~(str_t)0 : (((str_t)1 << L) - 1);
    const size_t P = (size_t)L * L;
    const double cj = c[J];
    const uint64_t self = find_key(ks, nk, det_key(u, d, L));
    cK[self] = cj;
    for (str_t t = u; t; t &= t - 1) {  // diagonal: X^a[J][b*L + b] = c_J
        const int b = __builtin_ctzll(t);
        Xa[self * P + b * L + b] = cj;
    }
    for (str_t t = d; t; t &= t - 1) {
        const int b = __builtin_ctzll(t);
        Xb[self * P + b * L + b] = cj;
    }
    for (int s = 0; s < 2; ++s) {
        const str_t own = s == 0 ? u : d;
        double* X = s == 0 ? Xa : Xb;
        for (str_t t = own; t; t &= t - 1) {
            const int b = __builtin_ctzll(t);
            for (str_t v = full & ~own; v; v &= v - 1) {
                const int a = __builtin_ctzll(v);
                const str_t ks_ = (own ^ ((str_t)1 << b)) | ((str_t)1 << a);  // K's s-string: J - b + a
                const uint64_t key = s == 0 ? det_key(ks_, d, L) : det_key(u, ks_, L);
                const uint64_t k = find_key(ks, nk, key);
                const int ph = stacked(own & ks_, a, b) & 1;  // <K| a+_a a_b |J> phase
                X[k * P + a * L + b] = ph ? -cj : cj;
            }
        }
    }
}

__global__ void k_count_equal_neighbours(uint64_t n, const uint64_t* __restrict__ k, unsigned int* __restrict__ dup) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i + 1 < n && k[i] == k[i + 1]) atomicAdd(dup, 1u);
}

void rdm_dets(hipStream_t stream, int L, uint64_t n, const str_t* up, const str_t* down, const double* c_host,
              double* gamma, double* Gamma) {
    if (2 * L > 64) throw Error(ERR_UNSUPPORTED, "determinant-list RDMs need L <= 32");
    int na = 0, nb = 0;
    {
        str_t u0 = 0, d0 = 0;
        CF_HIP(hipMemcpyAsync(&u0, up, sizeof(str_t), hipMemcpyDeviceToHost, stream));
        CF_HIP(hipMemcpyAsync(&d0, down, sizeof(str_t), hipMemcpyDeviceToHost, stream));
        CF_HIP(hipStreamSynchronize(stream));
        na = __builtin_popcountll(u0);
        nb = __builtin_popcountll(d0);
    }
    const int E = 1 + na * (L - na) + nb * (L - nb);
    const uint64_t nkeys = n * (uint64_t)E;
    const unsigned grid = (unsigned)((n + 127) / 128);
    DBuf<uint64_t> keys(nkeys), sorted(nkeys), uniq(nkeys);
    DBuf<int> nsel(1);
    k_det_closure_keys<<<grid, 128, 0, stream>>>(n, L, E, up, down, keys.p);
    CF_HIP(hipGetLastError());
    // the list itself must not repeat a determinant (its key sits at keys[J*E])
    {
        DBuf<uint64_t> lk(n), lks(n);
        DBuf<unsigned int> dup(1);
        CF_HIP(hipMemcpy2DAsync(lk.p, sizeof(uint64_t), keys.p, E * sizeof(uint64_t), sizeof(uint64_t), n,
                                hipMemcpyDeviceToDevice, stream));
        size_t tb = 0;
        CF_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, lk.p, lks.p, (int)n, 0, 2 * L, stream));
        DBuf<char> tmp(tb ? tb : 1);
        CF_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, lk.p, lks.p, (int)n, 0, 2 * L, stream));
        CF_HIP(hipMemsetAsync(dup.p, 0, sizeof(unsigned int), stream));
        k_count_equal_neighbours<<<grid, 128, 0, stream>>>(n, lks.p, dup.p);
        CF_HIP(hipGetLastError());
        unsigned int hd = 0;
        CF_HIP(hipMemcpyAsync(&hd, dup.p, sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
        CF_HIP(hipStreamSynchronize(stream));
        if (hd) throw Error(ERR_ARG, "determinant list repeats a determinant; RDMs need distinct determinants");
    }
    if (nkeys > 0x7fffffffull) throw Error(ERR_UNSUPPORTED, "determinant list too long for the RDM closure");
    size_t tb = 0, tb2 = 0;
    CF_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys.p, sorted.p, (int)nkeys, 0, 2 * L, stream));
    CF_HIP(hipcub::DeviceSelect::Unique(nullptr, tb2, sorted.p, uniq.p, nsel.p, (int)nkeys, stream));
    DBuf<char> tmp(std::max(tb, tb2) ? std::max(tb, tb2) : 1);
    CF_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys.p, sorted.p, (int)nkeys, 0, 2 * L, stream));
    CF_HIP(hipcub::DeviceSelect::Unique(tmp.p, tb2, sorted.p, uniq.p, nsel.p, (int)nkeys, stream));
    int hk = 0;
    CF_HIP(hipMemcpyAsync(&hk, nsel.p, sizeof(int), hipMemcpyDeviceToHost, stream));
    CF_HIP(hipStreamSynchronize(stream));
    const uint64_t nk = (uint64_t)hk;
    const size_t P = (size_t)L * L;
    DBuf<double> c(n), cK(nk), Xa(nk * P), Xb(nk * P);
    CF_HIP(hipMemcpyAsync(c.p, c_host, n * sizeof(double), hipMemcpyHostToDevice, stream));
    CF_HIP(hipMemsetAsync(cK.p, 0, cK.bytes(), stream));
    CF_HIP(hipMemsetAsync(Xa.p, 0, Xa.bytes(), stream));
    CF_HIP(hipMemsetAsync(Xb.p, 0, Xb.bytes(), stream));
    k_det_scatter_x<<<grid, 128, 0, stream>>>(n, L, up, down, c.p, uniq.p, nk, cK.p, Xa.p, Xb.p);
    CF_HIP(hipGetLastError());
    rdm_from_x(stream, L, nk, cK.p, Xa.p, Xb.p, gamma, Gamma);
}

void pack_rdm(int L, const double* gamma, const double* Gamma, double* one, double* two) {
    const size_t osz = (size_t)L * (L + 1) / 2;
    for (int s = 0; s < 2; ++s)
        for (int p = 0; p < L; ++p)
            for (int q = p; q < L; ++q) one[s * osz + (size_t)p * L + q - (size_t)p * (p + 1) / 2] = gamma[s * L * L + p * L + q];
    if (!two || !Gamma) return;
    const size_t C = (size_t)L * L;
    const size_t tsz = C * (C + 1) / 2;
    const size_t L4 = C * C;
    auto idx = [&](size_t r, size_t c) { return C * r + c - r * (r + 1) / 2; };
    for (size_t i = 0; i < 3 * tsz; ++i) two[i] = 0.0;
    for (int i = 0; i < L; ++i)
        for (int j = 0; j < L; ++j)
            for (int k = 0; k < L; ++k)
                for (int l = 0; l < L; ++l) {
                    const size_t P1 = (size_t)i * L + j, Q1 = (size_t)k * L + l;
                    if (P1 <= Q1) two[tsz + idx(P1, Q1)] = Gamma[L4 + (((size_t)i * L + j) * L + k) * L + l];
                    if (i < j && k < l) {
                        const size_t P = (size_t)i * L + j - (size_t)(i + 1) * (i + 2) / 2;
                        const size_t Q = (size_t)k * L + l - (size_t)(k + 1) * (k + 2) / 2;
                        if (P <= Q) {
                            two[idx(P, Q)] = Gamma[(((size_t)i * L + j) * L + k) * L + l];
                            two[2 * tsz + idx(P, Q)] = Gamma[2 * L4 + (((size_t)i * L + j) * L + k) * L + l];
                        }
                    }
                }
}

}  // namespace cf
