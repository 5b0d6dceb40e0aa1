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

// ---------------------------------------------------------------------------
// Block-staged product: no stored columns.  Rows sharing the high part hi of
// their string form the contiguous rank range base(hi, c) + rank_lo(lo), with
// rank(s) = sum_m C(k_m, m) split at the w-th orbital:
//   base(hi, c) = sum over the high bits (m = c+1, c+2, ...) of C(w + q, m).
// A workgroup owns one block (R = C(w, c) rows, every row with the same c, so
// every loop below has the same trip count in every lane):
//   low -> low hops:   target in the own block, x staged in LDS once;
//   high -> high hops: block (hi ^ jh ^ kh, c), same low pattern, so the same
//                      row index: one coalesced read of the neighbour block;
//   high -> low hops:  block (hi ^ jh, c + 1), staged in LDS per jh;
//   low -> high hops:  block (hi | kh, c - 1), staged in LDS per kh.
// Per product the x blocks pass through L2 about 1 + ch (L - w - ch) times
// (ch high pairs) plus the staged blocks; HBM sees x, y and the diagonal once.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t doci_block_base(uint32_t hi, int c, int w, const uint64_t* __restrict__ binom) {
    uint32_t b = 0;
    int m = c + 1;
    for (uint32_t t = hi; t; t &= t - 1) {
        b += (uint32_t)binom[(w + __builtin_ctz(t)) * kBinomCols + m];
        ++m;
    }
    return b;
}

constexpr int kBlkRows = 1024;  // LDS x block capacity (>= C(12, 6))
constexpr int kBlkHops = 128;   // high -> high hops per block: ch (nh - ch) <= 100 for nh <= 20
constexpr int kBlkStages = 32;  // staged blocks per block: nh <= 20

__device__ __forceinline__ int nth_bit(uint32_t v, int k) {  // position of the k-th (0-based) set bit
    for (int i = 0; i < k; ++i) v &= v - 1;
    return __builtin_ctz(v);
}

template <typename VT, int ROWS, int T, bool BATCH>
__global__ void __launch_bounds__(T) k_spmv_doci_blk(DociBlockArgs a, const VT* __restrict__ vk0,
                                                      const double* __restrict__ x0, double* __restrict__ y0,
                                                      DociBatch bt) {
    extern __shared__ double smem_blk[];
    const VT* __restrict__ vk = vk0;
    const double* __restrict__ x = x0;
    double* __restrict__ y = y0;
    const double* __restrict__ diag = a.diag;
    if (BATCH) {  // lane blockIdx.y of a batched Davidson product
        const int lane = blockIdx.y;
        const BdStep& st = bt.step[lane];
        if (st.mode == BD_IDLE || st.f32 != (sizeof(VT) == 4 ? 1 : 0)) return;
        const BdOpLane& o = bt.op[lane];
        vk = sizeof(VT) == 4 ? reinterpret_cast<const VT*>(o.vk32) : reinterpret_cast<const VT*>(o.vk64);
        diag = o.diag;
        const uint64_t off = ((uint64_t)lane * bt.M + st.m) * bt.n;
        x = bt.V + off;
        y = bt.S + off;
    }
    __shared__ uint32_t hh_base[kBlkHops];
    __shared__ double hh_val[kBlkHops];
    __shared__ uint32_t sg_base[kBlkStages], sg_orb[kBlkStages];
    const int L = a.L, w = a.w, nh = L - w;
    const uint32_t W = 1u << w, wmask = W - 1;
    // one x buffer: the own block, then each staged block in turn; a barrier
    // before every stage's write lets all threads finish reading the previous
    // contents (one buffer instead of three leaves room for more resident
    // workgroups per CU, the kernel's limit)
    double* xo = smem_blk;
    double* sbuf = smem_blk;
    VT* tab = reinterpret_cast<VT*>(smem_blk + kBlkRows);
    uint16_t* rlo = reinterpret_cast<uint16_t*>(tab + L * L);
    const int tid = threadIdx.x;
    const DociBlock blk = a.blocks[blockIdx.x];
    const uint32_t R = blk.R, base = blk.base, hi = blk.hi;
    const int c = (int)blk.c;
    const uint32_t hv = ~hi & ((1u << nh) - 1);
    const int ch = __builtin_popcount(hi), cv = nh - ch;
    const int nhh = ch * cv;
    // stages: high -> low (block (hi ^ jh, c + 1)) for every occupied high
    // orbital, then low -> high (block (hi | kh, c - 1)) for every empty one
    const int n_hl = c < w ? ch : 0, n_lh = c > 0 ? cv : 0, nst = n_hl + n_lh;
    const uint32_t R_hl = c < w ? a.lo_off[c + 2] - a.lo_off[c + 1] : 0u;
    const uint32_t R_lh = c > 0 ? a.lo_off[c] - a.lo_off[c - 1] : 0u;
    for (int t = tid; t < L * L; t += T) tab[t] = vk[t];
    for (uint32_t t = tid; t < W; t += T) rlo[t] = a.rank_lo[t];
    for (uint32_t t = tid; t < R; t += T) xo[t] = x[base + t];
    if (tid < nhh) {  // neighbour blocks of the high -> high hops, one per thread
        const int jh = nth_bit(hi, tid / cv), kh = nth_bit(hv, tid % cv);
        hh_base[tid] = doci_block_base(hi ^ (1u << jh) ^ (1u << kh), c, w, a.binom);
        hh_val[tid] = (double)vk[(w + jh) * L + (w + kh)];
    }
    if (tid < nst) {
        if (tid < n_hl) {
            const int jh = nth_bit(hi, tid);
            sg_base[tid] = doci_block_base(hi ^ (1u << jh), c + 1, w, a.binom);
            sg_orb[tid] = (uint32_t)(w + jh);
        } else {
            const int kh = nth_bit(hv, tid - n_hl);
            sg_base[tid] = doci_block_base(hi | (1u << kh), c - 1, w, a.binom);
            sg_orb[tid] = (uint32_t)(w + kh);
        }
    }
    const uint16_t* lof = a.lo_of + a.lo_off[c];
    uint32_t lo[ROWS];
    double acc[ROWS];
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const uint32_t r = tid + q * T;
        lo[q] = r < R ? lof[r] : 0u;
        acc[q] = 0.0;
    }
    __syncthreads();
    // first staged block into registers while the own block is worked on
    double pre[ROWS];
    auto fetch = [&](int sidx) {
        const uint32_t Rs = sidx < n_hl ? R_hl : R_lh, b = sg_base[sidx];
#pragma unroll
        for (int i = 0; i < ROWS; ++i) {
            const uint32_t t = tid + i * T;
            pre[i] = t < Rs ? x[b + t] : 0.0;
        }
    };
    if (nst) fetch(0);
    // diagonal and low -> low hops
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const uint32_t r = tid + q * T;
        if (r < R) {
            double d = diag[base + r];
            if (!(fabs(d) > a.zero)) d = 0.0;
            double s = (double)(VT)d * xo[r];
            const uint32_t l = lo[q], vl = ~l & wmask;
            for (uint32_t o = l; o; o &= o - 1) {
                const int j = __builtin_ctz(o);
                const uint32_t lj = l ^ (1u << j);
                const VT* trow = tab + j * L;
                for (uint32_t v = vl; v; v &= v - 1) {
                    const int k = __builtin_ctz(v);
                    s += (double)trow[k] * xo[rlo[lj | (1u << k)]];
                }
            }
            acc[q] = s;
        }
    }
    // high -> high hops: the neighbour block at the same row index, four
    // hops' loads in flight per row
    {
        int t = 0;
        for (; t + 4 <= nhh; t += 4) {
            const uint32_t b0 = hh_base[t], b1 = hh_base[t + 1], b2 = hh_base[t + 2], b3 = hh_base[t + 3];
            const double e0 = hh_val[t], e1 = hh_val[t + 1], e2 = hh_val[t + 2], e3 = hh_val[t + 3];
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const uint32_t r = tid + q * T;
                if (r < R) {
                    const double x0 = x[b0 + r], x1 = x[b1 + r], x2 = x[b2 + r], x3 = x[b3 + r];
                    acc[q] += e0 * x0 + e1 * x1 + e2 * x2 + e3 * x3;
                }
            }
        }
        for (; t < nhh; ++t) {
            const uint32_t b0 = hh_base[t];
            const double e0 = hh_val[t];
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const uint32_t r = tid + q * T;
                if (r < R) acc[q] += e0 * x[b0 + r];
            }
        }
    }
    // staged hops across the split: stage s+1 loads into registers while
    // stage s is read from LDS; two barriers per stage
    for (int sidx = 0; sidx < nst; ++sidx) {
        double* buf = sbuf;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < ROWS; ++i) buf[tid + i * T] = pre[i];
        __syncthreads();
        if (sidx + 1 < nst) fetch(sidx + 1);
        const int orb = (int)sg_orb[sidx];
        const VT* trow = tab + orb * L;  // pair elements with the high orbital (symmetric table)
        if (sidx < n_hl) {  // high orbital orb -> empty low orbital k
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const uint32_t r = tid + q * T;
                if (r < R) {
                    const uint32_t l = lo[q];
                    double s = 0.0;
                    for (uint32_t v = ~l & wmask; v; v &= v - 1) {
                        const int k = __builtin_ctz(v);
                        s += (double)trow[k] * buf[rlo[l | (1u << k)]];
                    }
                    acc[q] += s;
                }
            }
        } else {  // occupied low orbital j -> empty high orbital orb
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const uint32_t r = tid + q * T;
                if (r < R) {
                    const uint32_t l = lo[q];
                    double s = 0.0;
                    for (uint32_t o = l; o; o &= o - 1) {
                        const int j = __builtin_ctz(o);
                        s += (double)trow[j] * buf[rlo[l ^ (1u << j)]];
                    }
                    acc[q] += s;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < ROWS; ++q) {
        const uint32_t r = tid + q * T;
        if (r < R) y[base + r] = acc[q];
    }
}

int doci_block_w(int L) { return L / 2 < kDociMaxW ? L / 2 : kDociMaxW; }

template <typename VT, bool BATCH>
static void launch_doci_blk_t(const DociBlockArgs& a, const VT* vk, const double* x, double* y, hipStream_t s,
                              const DociBatch& bt, int lanes) {
    if (!a.nblocks) return;
    CF_REQUIRE(a.w >= 1 && a.w <= kDociMaxW && a.L - a.w >= 1 && a.L - a.w <= 20,
               "block-staged DOCI SpMV: 1 <= w <= 12 low orbitals and 1..20 high orbitals");
    const size_t shm = kBlkRows * sizeof(double) + (size_t)a.L * a.L * sizeof(VT) +
                       ((size_t)1 << a.w) * sizeof(uint16_t);
    // rows of the largest block over the workgroup (T * ROWS >= C(w, w/2)); the
    // neighbour-block bases are computed one per thread (T >= ch (nh - ch))
    const uint64_t maxR = binom_host(a.w, a.w / 2);
    const int nh = a.L - a.w, maxhh = (nh / 2) * ((nh + 1) / 2);
    const dim3 g(a.nblocks, (unsigned)lanes);
    if (maxR <= 64 && maxhh <= 64) k_spmv_doci_blk<VT, 1, 64, BATCH><<<g, 64, shm, s>>>(a, vk, x, y, bt);
    else if (maxR <= 128 && maxhh <= 128) k_spmv_doci_blk<VT, 1, 128, BATCH><<<g, 128, shm, s>>>(a, vk, x, y, bt);
    else if (maxR <= 256) k_spmv_doci_blk<VT, 1, 256, BATCH><<<g, 256, shm, s>>>(a, vk, x, y, bt);
    else if (maxR <= 512) k_spmv_doci_blk<VT, 2, 256, BATCH><<<g, 256, shm, s>>>(a, vk, x, y, bt);
    else k_spmv_doci_blk<VT, 4, 256, BATCH><<<g, 256, shm, s>>>(a, vk, x, y, bt);
    CF_HIP(hipGetLastError());
}

template <typename VT>
static void launch_doci_blk(const DociBlockArgs& a, const VT* vk, const double* x, double* y, hipStream_t s) {
    launch_doci_blk_t<VT, false>(a, vk, x, y, s, DociBatch{}, 1);
}

void launch_spmv_doci_blk_batch(const DociBlockArgs& a, bool f32, const DociBatch& bt, int lanes, hipStream_t s) {
    CF_REQUIRE(lanes >= 1 && lanes <= kBdMaxLanes, "batched DOCI product: 1..32 lanes");
    if (f32) launch_doci_blk_t<float, true>(a, nullptr, nullptr, nullptr, s, bt, lanes);
    else launch_doci_blk_t<double, true>(a, nullptr, nullptr, nullptr, s, bt, lanes);
}

void launch_spmv_doci_blk_f64(const DociBlockArgs& a, const double* vk, const double* x, double* y, hipStream_t s) {
    launch_doci_blk<double>(a, vk, x, y, s);
}

void launch_spmv_doci_blk_f32(const DociBlockArgs& a, const float* vk, const double* x, double* y, hipStream_t s) {
    launch_doci_blk<float>(a, vk, x, y, s);
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
