// Batched Davidson step kernels: the served stream's FCI and larger DOCI
// requests of one shape (same n) advanced in lock step, so that a step is one
// launch for every request of the batch instead of one per request (solved
// one at a time, each request's step is a chain of short launches bound by
// launch latency, not by HBM).  The reference solves these one after another
// (OrbitalOptimization::update_cim -> set_ham -> build_ci -> solve,
// SimulatedAnnealing.cpp:83-87).
//
// Lane l is grid row blockIdx.y (blockIdx.z for the dot products); its grid-x
// decomposition, and so every reduction order, depends on n alone, so a
// lane's results are the same in any batch and alone.  The arithmetic of
// each step is that of the single-request Davidson step kernels (krylov.hip):
// x = V y and S y in one pass, r = S y - theta x, the correction
// t = r / (theta - diag) (|theta - diag| >= shift), two selective Gram-Schmidt
// passes, v_m = t / |t|, then after the product the m + 1 dots V^T S_m.
#include "kernels.h"
#include "krylov_dev.h"

#include <algorithm>

namespace cf {

namespace {

constexpr int kChunk = 8;  // basis vectors per dot-product block row

__device__ __forceinline__ double* lane_parts(const BdView& b, int lane) {
    return b.parts + (size_t)lane * kBdParts;
}

// a = sum_i y[i] V[i n + q], c = sum_i y[i] S[i n + q] over i < k, summed in
// order i = 0, 1, ...; the basis size k varies per lane at run time, so the
// loads are issued eight at a time before they are used (one at a time, the
// pass ran at a third of the HBM rate).
__device__ __forceinline__ void combine2(const double* __restrict__ V, const double* __restrict__ S,
                                         const double* y, int k, uint64_t n, uint64_t q, double& a, double& c) {
    int i = 0;
    for (; i + 8 <= k; i += 8) {
        double v[8], w[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            v[j] = V[(uint64_t)(i + j) * n + q];
            w[j] = S[(uint64_t)(i + j) * n + q];
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            a += y[i + j] * v[j];
            c += y[i + j] * w[j];
        }
    }
    for (; i < k; ++i) {
        a += y[i] * V[(uint64_t)i * n + q];
        c += y[i] * S[(uint64_t)i * n + q];
    }
}

// s = sum_i c[i] V[i n + q] over i < k, in order, loads eight at a time.
__device__ __forceinline__ double combine1(const double* __restrict__ V, const double* c, int k, uint64_t n,
                                           uint64_t q) {
    double s = 0.0;
    int i = 0;
    for (; i + 8 <= k; i += 8) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = V[(uint64_t)(i + j) * n + q];
#pragma unroll
        for (int j = 0; j < 8; ++j) s += c[i + j] * v[j];
    }
    for (; i < k; ++i) s += c[i] * V[(uint64_t)i * n + q];
    return s;
}

// Restart: V[:, 0:keep] = V[:, 0:mr] Y[:, 0:keep] in place on FP64 MFMA
// (the tiling of k_rotate_mfma); blockIdx.z selects V or S.
typedef double v4d __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_bd_rotate(BdView b) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    if (st.mode == BD_IDLE || !st.restart) return;
    __shared__ double ys[kDavMaxBasis * kDavMaxBasis];
    const int m = st.mr, k = st.keep;
    const uint64_t n = b.n;
    double* Vb = (blockIdx.z ? b.S : b.V) + (uint64_t)lane * b.M * n;
    const double* Y = b.Y + (size_t)lane * kDavMaxBasis * kDavMaxBasis;
    for (int i = threadIdx.x; i < m * m; i += blockDim.x) ys[i] = Y[i];
    __syncthreads();
    const int l = threadIdx.x & 63, r = l & 15, kq = l >> 4;
    const uint64_t q0 = ((uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (q0 >= n) return;
    const uint64_t q = q0 + r;
    const bool inq = q < n;
    v4d acc[2];
    acc[0] = v4d{0.0, 0.0, 0.0, 0.0};
    acc[1] = v4d{0.0, 0.0, 0.0, 0.0};
    const int ct = (k + 15) / 16;
    for (int i0 = 0; i0 < m; i0 += 4) {
        const int i = i0 + kq;
        const double bv = (i < m && inq) ? Vb[(uint64_t)i * n + q] : 0.0;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            if (t < ct) {
                const int c = t * 16 + r;
                const double av = (i < m && c < k) ? ys[i * m + c] : 0.0;
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        if (t < ct) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = t * 16 + 4 * v + kq;
                if (c < k && inq) Vb[(uint64_t)c * n + q] = acc[t][v];
            }
        }
    }
}

// BD_RITZ: x = V y, r = S y - theta x, t = r / (theta - diag) (t = r when
// shift is 0), parts of |r|^2 and |t|^2.  BD_START_X / BD_START_SEED: t
// = x or the seeded start vector, parts of |t|^2 (|r|^2 parts zero).
__global__ void __launch_bounds__(256) k_bd_combine(BdView b) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    const int mode = st.mode;
    if (mode == BD_IDLE) return;
    __shared__ double ys[kDavMaxBasis];
    __shared__ double red[2][4];
    const uint64_t n = b.n;
    const int k = mode == BD_RITZ ? st.m : 0;
    if (threadIdx.x < kDavMaxBasis) ys[threadIdx.x] = (int)threadIdx.x < k ? st.ys[threadIdx.x] : 0.0;
    __syncthreads();
    const double theta = st.theta, shift_min = st.shift;
    const uint64_t seed = st.seed;
    const double* V = b.V + (uint64_t)lane * b.M * n;
    const double* S = b.S + (uint64_t)lane * b.M * n;
    double* x = b.x + (uint64_t)lane * n;
    double* t = b.t + (uint64_t)lane * n;
    const double* diag = b.op[lane].diag;
    double acc = 0.0, acct = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        double tq;
        if (mode == BD_RITZ) {
            double a = 0.0, c = 0.0;
            combine2(V, S, ys, k, n, q, a, c);
            x[q] = a;
            const double r = c - theta * a;
            acc += r * r;
            tq = r;
            if (shift_min > 0.0) {
                double den = theta - diag[q];
                if (fabs(den) < shift_min) den = den < 0.0 ? -shift_min : shift_min;
                tq = r / den;
            }
        } else {
            tq = mode == BD_START_X ? x[q] : start_element(seed, q);
        }
        t[q] = tq;
        acct += tq * tq;
    }
    acc = wave_sum(acc);
    acct = wave_sum(acct);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = acc;
        red[1][threadIdx.x >> 6] = acct;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double* P = lane_parts(b, lane);
        P[512 + blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        P[768 + blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

// Dot products of basis vectors with w, partial[i * Gd + block]: pass 0
// V[0:m] . t, pass 1 the same after the first Gram-Schmidt pass (skipped
// when that pass kept more than eta of t), pass 2 V[0:m+1] . S_m.
__global__ void __launch_bounds__(1024) k_bd_dot(BdView b, int pass, double eta2) {
    const int lane = blockIdx.z;
    const BdStep& st = b.step[lane];
    if (st.mode == BD_IDLE) return;
    const int m = st.m;
    const int k = pass == 2 ? m + 1 : m;
    const int i0 = blockIdx.y * kChunk;
    if (i0 >= k) return;
    __shared__ double red[kChunk][16];
    const uint64_t n = b.n;
    const double* P = lane_parts(b, lane);
    if (pass == 1 && dgks_skip(P + 768, P, b.Ga, eta2)) return;
    const double* Vb = b.V + (uint64_t)lane * b.M * n;
    const double* w = pass == 2 ? b.S + ((uint64_t)lane * b.M + m) * n : b.t + (uint64_t)lane * n;
    double* partial = b.partial + (size_t)lane * kBdPartial;
    double acc[kChunk];
#pragma unroll
    for (int j = 0; j < kChunk; ++j) acc[j] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * 1024;
    for (uint64_t q = (uint64_t)blockIdx.x * 1024 + threadIdx.x; q < n; q += stride) {
        const double wq = w[q];
#pragma unroll
        for (int j = 0; j < kChunk; ++j)
            if (i0 + j < k) acc[j] += Vb[(uint64_t)(i0 + j) * n + q] * wq;
    }
    const int l = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
        const double s = wave_sum(acc[j]);
        if (l == 0) red[j][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < kChunk && i0 + (int)threadIdx.x < k) {
        double s = 0.0;
        for (int q = 0; q < 16; ++q) s += red[threadIdx.x][q];
        partial[(uint64_t)(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = s;
    }
}

// t -= V[0:m] c (c from the pass's dot partials) and parts of |t|^2 after:
// pass 0 into pnorm, pass 1 (the gated second pass) into pnorm2.
__global__ void __launch_bounds__(256) k_bd_axpy(BdView b, int pass, double eta2) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    if (st.mode == BD_IDLE) return;
    __shared__ double cs[64];
    __shared__ double red[4];
    const uint64_t n = b.n;
    double* P = lane_parts(b, lane);
    if (pass == 1 && dgks_skip(P + 768, P, b.Ga, eta2)) return;
    const int k = st.m;
    reduce_coeffs(k, b.Gd, b.partial + (size_t)lane * kBdPartial, cs);
    __syncthreads();
    const double* Vb = b.V + (uint64_t)lane * b.M * n;
    double* w = b.t + (uint64_t)lane * n;
    double nrm = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const double s = combine1(Vb, cs, k, n, q);
        const double wn = w[q] - s;
        w[q] = wn;
        nrm += wn * wn;
    }
    nrm = wave_sum(nrm);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) P[(pass ? 256 : 0) + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// v_m = t / |t| (|t|^2 from the last pass that ran), its FP32 copy for FP32
// products, dg[62] = |r|^2, dg[63] = |t|^2.
__global__ void __launch_bounds__(256) k_bd_scale(BdView b, double eta2) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    if (st.mode == BD_IDLE) return;
    __shared__ double red[2][4];
    const uint64_t n = b.n;
    const double* P = lane_parts(b, lane);
    const int np = b.Ga;
    const double* pnorm = dgks_skip(P + 768, P, np, eta2) ? P : P + 256;
    double a = (int)threadIdx.x < np ? pnorm[threadIdx.x] : 0.0;
    double r = (int)threadIdx.x < np ? P[512 + threadIdx.x] : 0.0;
    a = wave_sum(a);
    r = wave_sum(r);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = r;
    }
    __syncthreads();
    const double q2 = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double s = q2 > 1e-28 ? 1.0 / sqrt(q2) : 0.0;
    const double* src = b.t + (uint64_t)lane * n;
    double* dst = b.V + ((uint64_t)lane * b.M + st.m) * n;
    float* d32 = (b.x32 && st.f32) ? b.x32 + (uint64_t)lane * n : nullptr;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const double v = src[q] * s;
        dst[q] = v;
        if (d32) d32[q] = (float)v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double* dg = b.dg + (size_t)lane * 64;
        dg[62] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        dg[63] = q2;
    }
}

// dg[i] = sum_b partial[i * Gd + b] for i <= m (fixed order).
__global__ void k_bd_reduce(BdView b) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    const int i = blockIdx.x;
    if (st.mode == BD_IDLE || i > st.m) return;
    const double* partial = b.partial + (size_t)lane * kBdPartial;
    double s = 0.0;
    for (int q = threadIdx.x; q < b.Gd; q += 64) s += partial[(uint64_t)i * b.Gd + q];
    s = wave_sum(s);
    if (threadIdx.x == 0) b.dg[(size_t)lane * 64 + i] = s;
}

// Parts of |S y1 - theta1 V y1|^2 over the first mr vectors (the basis the
// Ritz pair belongs to) for lanes with r2 set.
__global__ void __launch_bounds__(256) k_bd_resid2(BdView b) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    if (st.mode != BD_RITZ || !st.r2) return;
    __shared__ double ys[kDavMaxBasis];
    __shared__ double red[4];
    const uint64_t n = b.n;
    const int k = st.m;
    if (threadIdx.x < kDavMaxBasis) ys[threadIdx.x] = (int)threadIdx.x < k ? st.y1[threadIdx.x] : 0.0;
    __syncthreads();
    const double theta = st.theta1;
    const double* V = b.V + (uint64_t)lane * b.M * n;
    const double* S = b.S + (uint64_t)lane * b.M * n;
    double acc = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        double a = 0.0, c = 0.0;
        combine2(V, S, ys, k, n, q, a, c);
        const double r = c - theta * a;
        acc += r * r;
    }
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) lane_parts(b, lane)[1024 + blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// S_m[Ia][Ib] = sum_Ja Ha[Ia][Ja] C[Ja][Ib] + sum_Jb C[Ia][Jb] Hb[Jb][Ib],
// C = V_m (row-major da x db), on FP64 MFMA.  A workgroup of four wavefronts
// owns a 64 x 64 output block (wave w: the 32 x 32 quadrant (w >> 1, w & 1),
// four 16 x 16 tiles); the operands of 16 k at a time are staged in LDS
// (double-buffered), so every operand element read through L2 serves 64
// outputs (one wavefront per 32 x 32 block reading its operands directly ran
// 298 against 231 us per launch at FCI L=11, 32 requests).  Ha and Hb are
// symmetric, so the A operand of the first product is read along rows of Ha
// (coalesced).  Each output accumulates its k-steps of four in k order, the
// first product's then the second's, zero-padded past da / db.  Operand lanes:
// A[i][k] at lane i + 16 k, B[k][j] at lane j + 16 k; result D[4 e + k][j] in
// element e of lane j + 16 k.
constexpr int kSsK = 16;   // k per staged block
constexpr int kSsLd = 80;  // LDS row stride in doubles: the two k rows read by a half-wavefront use disjoint banks
__global__ void __launch_bounds__(256) k_bd_same_spin(BdView b, uint64_t da, uint64_t db, uint32_t tiles_b) {
    const int lane = blockIdx.y;
    const BdStep& st = b.step[lane];
    if (st.mode == BD_IDLE) return;  // uniform over the workgroup
    __shared__ double sa[2][kSsK][kSsLd];
    __shared__ double sb[2][kSsK][kSsLd];
    const uint64_t off = ((uint64_t)lane * b.M + st.m) * b.n;
    const double* C = b.V + off;
    double* out = b.S + off;
    const double* Ha = b.op[lane].Ha;
    const double* Hb = b.op[lane].Hb;
    const uint64_t Ia0 = (uint64_t)(blockIdx.x / tiles_b) * 64, Ib0 = (uint64_t)(blockIdx.x % tiles_b) * 64;
    const int t = threadIdx.x;
    const int w = t >> 6, l = t & 63, r = l & 15, kq = l >> 4;
    const int qi = (w >> 1) * 32, qj = (w & 1) * 32;
    const uint64_t nb1 = (da + kSsK - 1) / kSsK, nb2 = (db + kSsK - 1) / kSsK, nbt = nb1 + nb2;
    // thread t stages four A and four B operand elements of a block:
    //   H_alpha C: A[i][k] = Ha[k][i] (Ha symmetric), B[k][j] = C[k][j];
    //   C H_beta:  A[i][k] = C[i][k], B[k][j] = Hb[k][j].
    // Row-major staging: k row ka, columns cl + 16 e, so 16 lanes read and
    // store 16 consecutive doubles; the C H_beta A operand: row it of C, k
    // rows kr + e, so a wavefront stores 64 consecutive doubles of one k row.
    const int ka = t >> 4, cl = t & 15;
    const int it = t & 63, kr = 4 * (t >> 6);
    double ra[4], rb[4];
    auto fetch = [&](uint64_t kb) {
        if (kb < nb1) {
            const uint64_t kk = kb * kSsK + ka;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint64_t i = Ia0 + cl + 16 * e, j = Ib0 + cl + 16 * e;
                ra[e] = kk < da && i < da ? Ha[kk * da + i] : 0.0;
                rb[e] = kk < da && j < db ? C[kk * db + j] : 0.0;
            }
        } else {
            const uint64_t k0 = (kb - nb1) * kSsK;
            const uint64_t i = Ia0 + it, kk = k0 + ka;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint64_t ke = k0 + kr + e, j = Ib0 + cl + 16 * e;
                ra[e] = i < da && ke < db ? C[i * db + ke] : 0.0;
                rb[e] = kk < db && j < db ? Hb[kk * db + j] : 0.0;
            }
        }
    };
    auto stash = [&](uint64_t kb, int buf) {
        if (kb < nb1) {
#pragma unroll
            for (int e = 0; e < 4; ++e) sa[buf][ka][cl + 16 * e] = ra[e];
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) sa[buf][kr + e][it] = ra[e];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) sb[buf][ka][cl + 16 * e] = rb[e];
    };
    v4d acc00 = {0.0, 0.0, 0.0, 0.0}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
    // a quadrant wholly past da or db computes nothing (it still stages operands)
    const bool live = Ia0 + qi < da && Ib0 + qj < db;
    fetch(0);
    stash(0, 0);
    __syncthreads();
    for (uint64_t kb = 0; kb < nbt; ++kb) {
        const int buf = (int)(kb & 1);
        if (kb + 1 < nbt) fetch(kb + 1);  // the next block's loads run under this block's MFMAs
        const uint64_t K = kb < nb1 ? da : db;
        const uint64_t kbase = (kb < nb1 ? kb : kb - nb1) * kSsK;
#pragma unroll
        for (int s4 = 0; s4 < kSsK / 4; ++s4) {
            if (live && kbase + 4 * s4 < K) {  // k-steps of four up to da / db, no more
                const int k = 4 * s4 + kq;
                const double x0 = sa[buf][k][qi + r], x1 = sa[buf][k][qi + 16 + r];
                const double y0 = sb[buf][k][qj + r], y1 = sb[buf][k][qj + 16 + r];
                acc00 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y0, acc00, 0, 0, 0);
                acc01 = __builtin_amdgcn_mfma_f64_16x16x4f64(x0, y1, acc01, 0, 0, 0);
                acc10 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y0, acc10, 0, 0, 0);
                acc11 = __builtin_amdgcn_mfma_f64_16x16x4f64(x1, y1, acc11, 0, 0, 0);
            }
        }
        if (kb + 1 < nbt) stash(kb + 1, buf ^ 1);  // buf ^ 1 was last read before the previous barrier
        __syncthreads();
    }
    const uint64_t ib0 = Ib0 + qj + r, ib1 = ib0 + 16;
    const bool rb0 = ib0 < db, rb1 = ib1 < db;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint64_t i0 = Ia0 + qi + 4 * e + kq, i1 = i0 + 16;
        if (i0 < da) {
            if (rb0) out[i0 * db + ib0] = acc00[e];
            if (rb1) out[i0 * db + ib1] = acc01[e];
        }
        if (i1 < da) {
            if (rb0) out[i1 * db + ib0] = acc10[e];
            if (rb1) out[i1 * db + ib1] = acc11[e];
        }
    }
}

constexpr double kEta2 = 0.717 * 0.717;  // the single-request Davidson's DGKS threshold

void check_view(const BdView& b, int lanes) {
    if (lanes < 1 || lanes > kBdMaxLanes) throw Error(ERR_ARG, "batched Davidson: 1..32 lanes");
    if (!b.n || b.M < 1 || b.M > kDavMaxBasis) throw Error(ERR_ARG, "batched Davidson: bad n or basis size");
    if (b.Ga != fused_axpy_blocks(b.n) || b.Gd != fused_dot_blocks(b.n))
        throw Error(ERR_ARG, "batched Davidson: block counts do not match n");
}

}  // namespace

void bd_form_vectors(const BdView& b, int lanes, bool any_restart, hipStream_t s) {
    check_view(b, lanes);
    if (any_restart) {
        const uint64_t tiles = (b.n + 63) / 64;
        if (tiles > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "batched Davidson: n too large");
        k_bd_rotate<<<dim3((unsigned)tiles, lanes, 2), 256, 0, s>>>(b);
    }
    const dim3 ga((unsigned)b.Ga, lanes), gd((unsigned)b.Gd, (kDavMaxBasis + kChunk - 1) / kChunk, lanes);
    k_bd_combine<<<ga, 256, 0, s>>>(b);
    k_bd_dot<<<gd, 1024, 0, s>>>(b, 0, kEta2);
    k_bd_axpy<<<ga, 256, 0, s>>>(b, 0, kEta2);
    k_bd_dot<<<gd, 1024, 0, s>>>(b, 1, kEta2);
    k_bd_axpy<<<ga, 256, 0, s>>>(b, 1, kEta2);
    k_bd_scale<<<ga, 256, 0, s>>>(b, kEta2);
    CF_HIP(hipGetLastError());
}

void bd_new_column(const BdView& b, int lanes, hipStream_t s) {
    check_view(b, lanes);
    const dim3 gd((unsigned)b.Gd, (kDavMaxBasis + kChunk - 1) / kChunk, lanes);
    k_bd_dot<<<gd, 1024, 0, s>>>(b, 2, kEta2);
    k_bd_reduce<<<dim3(kDavMaxBasis, lanes), 64, 0, s>>>(b);
    CF_HIP(hipGetLastError());
}

void bd_resid2(const BdView& b, int lanes, hipStream_t s) {
    check_view(b, lanes);
    k_bd_resid2<<<dim3((unsigned)b.Ga, lanes), 256, 0, s>>>(b);
    CF_HIP(hipGetLastError());
}

void bd_same_spin(const BdView& b, int lanes, uint64_t da, uint64_t db, hipStream_t s) {
    check_view(b, lanes);
    if (da * db != b.n) throw Error(ERR_ARG, "batched same-spin sigma: da * db != n");
    const uint64_t ta = (da + 63) / 64, tb = (db + 63) / 64;
    if (ta * tb > 0x7fffffffull) throw Error(ERR_UNSUPPORTED, "batched same-spin sigma: too many tiles");
    k_bd_same_spin<<<dim3((unsigned)(ta * tb), lanes), 256, 0, s>>>(b, da, db, (uint32_t)tb);
    CF_HIP(hipGetLastError());
}

}  // namespace cf
