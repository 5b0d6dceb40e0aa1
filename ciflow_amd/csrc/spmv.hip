// CRS SpMV y = A x on full rows (replaces SparseMatrix_CRS_Sym::mvprod,
// SparseMatrix_CRS.cpp:700-723).
//
// The reference stores the upper triangle and scatters y[col] += a*x[row]
// (serial on the g++ build, MKL dcsrsymv on the Intel build).  Here every row
// is stored in full, so a group of lanes owns one row and only gathers:
// no atomics, no races, y overwritten.
//
// HBM-bound.  Algorithmic bytes per call:
//   B = nnz*(s_val + 4) + (n+1)*8 + 8n (x once) + 8n (y once),  s_val = 8 | 4.
// The matrix stream is read with non-temporal loads so it does not evict x
// from L2/MALL; x gathers hit cache.  Workgroups are remapped so that each of
// the 8 XCDs walks one contiguous slice of rows (hardware dispatches block b
// to XCD b % 8), which keeps the x segments a slice touches in that XCD's L2.
#include <atomic>
#include <cmath>

#include "kernels.h"

namespace cf {

template <typename T>
__device__ __forceinline__ T ntload(const T* p) {
    return __builtin_nontemporal_load(p);
}

template <typename VT, int LPR>
__global__ void __launch_bounds__(256) k_spmv(uint64_t n, uint32_t nblk, const int64_t* __restrict__ rowptr,
                                              const uint32_t* __restrict__ col, const VT* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y) {
    constexpr int RPB = 256 / LPR;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t b = blockIdx.x;
    const uint32_t lb = (b & 7u) * per + (b >> 3);
    if (lb >= nblk) return;
    const int sub = threadIdx.x % LPR;
    const uint64_t r = (uint64_t)lb * RPB + threadIdx.x / LPR;
    double acc = 0.0;
    if (r < n) {
        const int64_t s = rowptr[r], e = rowptr[r + 1];
        int64_t k = s + sub;
        for (; k + 3 * LPR < e; k += 4 * LPR) {
            const uint32_t c0 = ntload(col + k), c1 = ntload(col + k + LPR), c2 = ntload(col + k + 2 * LPR),
                           c3 = ntload(col + k + 3 * LPR);
            const double v0 = (double)ntload(val + k), v1 = (double)ntload(val + k + LPR),
                         v2 = (double)ntload(val + k + 2 * LPR), v3 = (double)ntload(val + k + 3 * LPR);
            acc += v0 * x[c0];
            acc += v1 * x[c1];
            acc += v2 * x[c2];
            acc += v3 * x[c3];
        }
        for (; k < e; k += LPR) acc += (double)ntload(val + k) * x[ntload(col + k)];
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (r < n && sub == 0) y[r] = acc;
}

template <typename VT>
static void launch_spmv_t(uint64_t n, const int64_t* rowptr, const uint32_t* col, const VT* val, const double* x,
                          double* y, int lpr, hipStream_t stream) {
    if (!n) return;
    const int rpb = 256 / lpr;
    uint64_t nblk = (n + rpb - 1) / rpb;
    uint64_t grid = (nblk + 7) / 8 * 8;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "spmv grid too large");
    switch (lpr) {
        case 64: k_spmv<VT, 64><<<(unsigned)grid, 256, 0, stream>>>(n, (uint32_t)nblk, rowptr, col, val, x, y); break;
        case 32: k_spmv<VT, 32><<<(unsigned)grid, 256, 0, stream>>>(n, (uint32_t)nblk, rowptr, col, val, x, y); break;
        case 16: k_spmv<VT, 16><<<(unsigned)grid, 256, 0, stream>>>(n, (uint32_t)nblk, rowptr, col, val, x, y); break;
        case 8: k_spmv<VT, 8><<<(unsigned)grid, 256, 0, stream>>>(n, (uint32_t)nblk, rowptr, col, val, x, y); break;
        default: throw Error(ERR_ARG, "lanes_per_row must be 8, 16, 32 or 64");
    }
    CF_HIP(hipGetLastError());
}

void launch_spmv_f64(uint64_t n, const int64_t* rowptr, const uint32_t* col, const double* val, const double* x,
                     double* y, int lanes_per_row, hipStream_t stream) {
    launch_spmv_t<double>(n, rowptr, col, val, x, y, lanes_per_row, stream);
}

void launch_spmv_f32(uint64_t n, const int64_t* rowptr, const uint32_t* col, const float* val, const double* x,
                     double* y, int lanes_per_row, hipStream_t stream) {
    launch_spmv_t<float>(n, rowptr, col, val, x, y, lanes_per_row, stream);
}

// Vectorised-stream variant: each lane reads 4 consecutive columns (16 B) and
// 4 consecutive values (32 B FP64 / 16 B FP32) per load, UNR groups in flight;
// the unaligned head and tail of a row (< 4 entries each) are read per entry.
// Same gathers and FP64 accumulation; per-row summation order differs from
// k_spmv (both are fixed, so each is deterministic).  On MI355X it is faster
// for FP32 values (FCI L=12: 3.47 ms against 4.30 ms) and for long FP64 rows.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
template <typename VT> struct Vec4;
template <> struct Vec4<double> { typedef f64x4 T; };
template <> struct Vec4<float> { typedef f32x4 T; };

template <typename VT, int LPR, int UNR>
__global__ void __launch_bounds__(256) k_spmv_v4(uint64_t n, uint32_t nblk, const int64_t* __restrict__ rowptr,
                                                 const uint32_t* __restrict__ col, const VT* __restrict__ val,
                                                 const double* __restrict__ x, double* __restrict__ y) {
    typedef typename Vec4<VT>::T V4;
    constexpr int RPB = 256 / LPR;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t b = blockIdx.x;
    const uint32_t lb = (b & 7u) * per + (b >> 3);
    if (lb >= nblk) return;
    const int sub = threadIdx.x % LPR;
    const uint64_t r = (uint64_t)lb * RPB + threadIdx.x / LPR;
    double acc = 0.0;
    if (r < n) {
        const int64_t s = rowptr[r], e = rowptr[r + 1];
        int64_t a0 = (s + 3) & ~(int64_t)3, a1 = e & ~(int64_t)3;
        if (a0 > e) a0 = e;
        if (a1 < a0) a1 = a0;
        if (s + sub < a0) acc += (double)val[s + sub] * x[col[s + sub]];
        if (a1 + sub < e) acc += (double)val[a1 + sub] * x[col[a1 + sub]];
        int64_t k = a0 + 4 * sub;
        for (; k + 4 * LPR * (UNR - 1) + 4 <= a1; k += 4 * LPR * UNR) {
            u32x4 c[UNR];
            V4 v[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                c[u] = __builtin_nontemporal_load((const u32x4*)(col + k + 4 * LPR * u));
                v[u] = __builtin_nontemporal_load((const V4*)(val + k + 4 * LPR * u));
            }
#pragma unroll
            for (int u = 0; u < UNR; ++u) {
                acc += (double)v[u].x * x[c[u].x];
                acc += (double)v[u].y * x[c[u].y];
                acc += (double)v[u].z * x[c[u].z];
                acc += (double)v[u].w * x[c[u].w];
            }
        }
        for (; k + 4 <= a1; k += 4 * LPR) {
            const u32x4 c = __builtin_nontemporal_load((const u32x4*)(col + k));
            const V4 v = __builtin_nontemporal_load((const V4*)(val + k));
            acc += (double)v.x * x[c.x];
            acc += (double)v.y * x[c.y];
            acc += (double)v.z * x[c.z];
            acc += (double)v.w * x[c.w];
        }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, LPR);
    if (r < n && sub == 0) y[r] = acc;
}

template <typename VT, int LPR, int UNR>
static void launch_v4_t(uint64_t n, const int64_t* rowptr, const uint32_t* col, const VT* val, const double* x,
                        double* y, hipStream_t stream) {
    if (!n) return;
    constexpr int rpb = 256 / LPR;
    const uint64_t nblk = (n + rpb - 1) / rpb;
    const uint64_t grid = (nblk + 7) / 8 * 8;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "spmv grid too large");
    k_spmv_v4<VT, LPR, UNR><<<(unsigned)grid, 256, 0, stream>>>(n, (uint32_t)nblk, rowptr, col, val, x, y);
    CF_HIP(hipGetLastError());
}

template <typename VT>
static void launch_v4(uint64_t n, const int64_t* rowptr, const uint32_t* col, const VT* val, const double* x,
                      double* y, int lpr, int unroll, hipStream_t stream) {
#define CF_V4(L, U) \
    if (lpr == L && unroll == U) return launch_v4_t<VT, L, U>(n, rowptr, col, val, x, y, stream)
    CF_V4(16, 2);
    CF_V4(8, 2);
    CF_V4(16, 1);
    CF_V4(4, 2);
    CF_V4(4, 1);
    CF_V4(8, 1);
    CF_V4(32, 1);
    CF_V4(32, 2);
#undef CF_V4
    throw Error(ERR_ARG, "vectorised spmv: lanes must be 4, 8, 16 or 32 and unroll 1 or 2");
}

void launch_spmv_v4_f64(uint64_t n, const int64_t* rowptr, const uint32_t* col, const double* val, const double* x,
                        double* y, int lanes_per_row, int unroll, hipStream_t stream) {
    launch_v4<double>(n, rowptr, col, val, x, y, lanes_per_row, unroll, stream);
}

void launch_spmv_v4_f32(uint64_t n, const int64_t* rowptr, const uint32_t* col, const float* val, const double* x,
                        double* y, int lanes_per_row, int unroll, hipStream_t stream) {
    launch_v4<float>(n, rowptr, col, val, x, y, lanes_per_row, unroll, stream);
}


// Sliced row gather (SELL-64): a copy of the rows in chunks of 64 consecutive
// rows, each chunk padded to its longest row rounded up to 4 entries (padding
// entries: column = the row itself, value 0).  Entry j of row 64c + l lies at
//   256 * qoff[c] + 256 * (j / 4) + 4 * l + j % 4,
// so quad q of a chunk is one contiguous stream of 64 x 16 B of columns and
// 64 x 32 B (FP64) / 16 B (FP32) of values, and the 64 gathers of one load
// instruction read the j-th columns of 64 neighbouring rows.  In a
// determinant list sorted by (up, down) those columns are neighbours too:
// measured on the CPU for a CAS-like list (the first 110 alpha strings of L=12
// times all 924 beta strings), 15.4 entries per 128-byte line of x against 4.1
// for 16 lanes of one row of the CRS layout (12.2 against 3.9 at L=16).
// A workgroup of W waves owns one chunk, or one of `split` consecutive quad
// ranges of it; wave w reads quads w, w + W, ... of its range (U quads in
// flight per lane) and the W partial sums are added in wave order through
// LDS.  With split > 1 the ranges' sums go to part[h * n + row] and
// k_sell_combine adds them in range order, so every row's summation order is
// fixed.  Splitting serves launches whose chunks would leave the last round of
// resident workgroups mostly idle (101,640 rows at 16 waves per chunk: 1,589
// workgroups on 512 slots).
template <typename VT, int W, int U>
__global__ void __launch_bounds__(64 * W) k_spmv_sell(uint64_t n, uint32_t nchunk, uint32_t split,
                                                      const int64_t* __restrict__ qoff,
                                                      const uint32_t* __restrict__ col, const VT* __restrict__ val,
                                                      const double* __restrict__ x, double* __restrict__ y,
                                                      double* __restrict__ part) {
    typedef typename Vec4<VT>::T V4;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t b = blockIdx.x;
    const uint32_t l = (b & 7u) * per + (b >> 3);
    const uint32_t c = l / split, h = l - c * split;
    if (c >= nchunk) return;  // uniform over the workgroup
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t q0 = qoff[c], nq_all = qoff[c + 1] - q0;
    const int64_t qs = (nq_all + split - 1) / split;
    const int64_t qb = (int64_t)h * qs;
    const int64_t nq = nq_all < qb + qs ? nq_all : qb + qs;
    const u32x4* cp = (const u32x4*)col + q0 * 64 + lane;
    const V4* vp = (const V4*)val + q0 * 64 + lane;
    double acc = 0.0;
    int64_t q = qb + w;
    for (; q + (int64_t)(U - 1) * W < nq; q += (int64_t)U * W) {
        u32x4 cc[U];
        V4 vv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            cc[u] = __builtin_nontemporal_load(cp + (q + u * W) * 64);
            vv[u] = __builtin_nontemporal_load(vp + (q + u * W) * 64);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc += (double)vv[u].x * x[cc[u].x];
            acc += (double)vv[u].y * x[cc[u].y];
            acc += (double)vv[u].z * x[cc[u].z];
            acc += (double)vv[u].w * x[cc[u].w];
        }
    }
    for (; q < nq; q += W) {
        const u32x4 cc = __builtin_nontemporal_load(cp + q * 64);
        const V4 vv = __builtin_nontemporal_load(vp + q * 64);
        acc += (double)vv.x * x[cc.x];
        acc += (double)vv.y * x[cc.y];
        acc += (double)vv.z * x[cc.z];
        acc += (double)vv.w * x[cc.w];
    }
    if constexpr (W > 1) {
        __shared__ double part[W][64];
        part[w][lane] = acc;
        __syncthreads();
        if (w == 0) {
#pragma unroll
            for (int i = 1; i < W; ++i) acc += part[i][lane];
        }
    }
    const uint64_t r = (uint64_t)c * 64 + lane;
    if (w == 0 && r < n) {
        if (split == 1) y[r] = acc;
        else part[(uint64_t)h * n + r] = acc;
    }
}

__global__ void __launch_bounds__(256) k_sell_combine(uint64_t n, uint32_t split, const double* __restrict__ part,
                                                      double* __restrict__ y) {
    const uint64_t r = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= n) return;
    double s = part[r];
    for (uint32_t h = 1; h < split; ++h) s += part[(uint64_t)h * n + r];
    y[r] = s;
}

// Ranges per chunk: the smallest of 1..kSellMaxSplit whose last round of
// resident workgroups (occupancy from the runtime, all CUs) is at least 90%
// full, else the fullest; 1 when there are eight or more rounds.
// (The occupancy is read once per kernel instance: every device a process
// drives here is the same model.)
template <typename VT, int W, int U>
static uint32_t sell_split_t(uint64_t nchunk) {
    static std::atomic<int> cached{0};  // resident workgroups of this instance on the device
    int slots = cached.load(std::memory_order_relaxed);
    if (!slots) {
        int dev = 0, cus = 0, per_cu = 0;
        CF_HIP(hipGetDevice(&dev));
        CF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        CF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_spmv_sell<VT, W, U>, 64 * W, 0));
        slots = per_cu > 0 && cus > 0 ? per_cu * cus : 1;
        cached.store(slots, std::memory_order_relaxed);
    }
    if (nchunk >= 8ull * (uint64_t)slots) return 1;
    uint32_t best = 1;
    double best_fill = 0.0;
    for (uint32_t s = 1; s <= kSellMaxSplit; ++s) {
        const double rounds = (double)(s * nchunk) / slots;
        const double fill = rounds / std::ceil(rounds);
        if (fill >= 0.9) return s;
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = s;
        }
    }
    return best;
}

template <typename VT, int W, int U>
static void launch_sell_t(uint64_t n, const int64_t* qoff, const uint32_t* col, const VT* val, const double* x,
                          double* y, double* part, uint64_t part_cap, hipStream_t stream) {
    const uint64_t nchunk = (n + 63) / 64;
    uint32_t split = sell_split_t<VT, W, U>(nchunk);
    while (split > 1 && (uint64_t)split * n > part_cap) --split;
    const uint64_t units = nchunk * split;
    const uint64_t grid = (units + 7) / 8 * 8;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "spmv grid too large");
    k_spmv_sell<VT, W, U><<<(unsigned)grid, 64 * W, 0, stream>>>(n, (uint32_t)nchunk, split, qoff, col, val, x, y,
                                                                 part);
    CF_HIP(hipGetLastError());
    if (split > 1) {
        k_sell_combine<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(n, split, part, y);
        CF_HIP(hipGetLastError());
    }
}

template <typename VT>
static void launch_sell(uint64_t n, const int64_t* qoff, const uint32_t* col, const VT* val, const double* x,
                        double* y, double* part, uint64_t part_cap, int waves, int unroll, hipStream_t stream) {
    if (!n) return;
#define CF_SELL(W_, U_) \
    if (waves == W_ && unroll == U_)   \
    return launch_sell_t<VT, W_, U_>(n, qoff, col, val, x, y, part, part_cap, stream)
    CF_SELL(1, 1);
    CF_SELL(1, 2);
    CF_SELL(2, 1);
    CF_SELL(2, 2);
    CF_SELL(4, 1);
    CF_SELL(4, 2);
    CF_SELL(8, 1);
    CF_SELL(8, 2);
    CF_SELL(16, 1);
    CF_SELL(16, 2);
    CF_SELL(4, 4);
    CF_SELL(8, 4);
    CF_SELL(16, 4);
#undef CF_SELL
    throw Error(ERR_ARG, "sliced spmv: waves per chunk must be 1, 2, 4, 8 or 16 and unroll 1 or 2 (4 with 4-16 waves)");
}

void launch_spmv_sell_f64(uint64_t n, const int64_t* qoff, const uint32_t* col, const double* val, const double* x,
                          double* y, double* part, uint64_t part_cap, int waves, int unroll, hipStream_t stream) {
    launch_sell<double>(n, qoff, col, val, x, y, part, part_cap, waves, unroll, stream);
}

void launch_spmv_sell_f32(uint64_t n, const int64_t* qoff, const uint32_t* col, const float* val, const double* x,
                          double* y, double* part, uint64_t part_cap, int waves, int unroll, hipStream_t stream) {
    launch_sell<float>(n, qoff, col, val, x, y, part, part_cap, waves, unroll, stream);
}

// Quads (4 entries) of each chunk's longest row: one wave per chunk.
__global__ void __launch_bounds__(256) k_sell_quads(uint64_t n, uint32_t nchunk, const int64_t* __restrict__ rowptr,
                                                    uint32_t* __restrict__ quads) {
    const uint32_t c = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (c >= nchunk) return;  // uniform over the wave
    const int lane = threadIdx.x & 63;
    const uint64_t r = (uint64_t)c * 64 + lane;
    uint32_t len = r < n ? (uint32_t)(rowptr[r + 1] - rowptr[r]) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) len = max(len, (uint32_t)__shfl_xor((int)len, off));
    if (lane == 0) quads[c] = (len + 3) / 4;
}

// Copies the CRS rows into the sliced layout: one workgroup per chunk, wave
// w writes quads w, w + 4, ... of the chunk's 64 rows.
__global__ void __launch_bounds__(256) k_sell_fill(uint64_t n, uint32_t nchunk, const int64_t* __restrict__ rowptr,
                                                   const uint32_t* __restrict__ col, const double* __restrict__ val,
                                                   const float* __restrict__ val32, const int64_t* __restrict__ qoff,
                                                   uint32_t* __restrict__ scol, double* __restrict__ sval,
                                                   float* __restrict__ sval32) {
    const uint32_t c = blockIdx.x;
    if (c >= nchunk) return;
    const int lane = threadIdx.x & 63;
    const uint64_t r = (uint64_t)c * 64 + lane;
    const int64_t s = r < n ? rowptr[r] : 0;
    const int64_t len = r < n ? rowptr[r + 1] - s : 0;
    const uint32_t pad = r < n ? (uint32_t)r : 0u;
    const int64_t q0 = qoff[c], nq = qoff[c + 1] - q0;
    for (int64_t q = threadIdx.x >> 6; q < nq; q += 4) {
        u32x4 cc;
        f64x4 vv;
        f32x4 ww;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t k = 4 * q + j;
            const bool in = k < len;
            cc[j] = in ? col[s + k] : pad;
            vv[j] = in ? val[s + k] : 0.0;
            ww[j] = in && val32 ? val32[s + k] : 0.0f;
        }
        const uint64_t o = (uint64_t)(q0 + q) * 64 + lane;
        ((u32x4*)scol)[o] = cc;
        ((f64x4*)sval)[o] = vv;
        if (sval32) ((f32x4*)sval32)[o] = ww;
    }
}

void launch_sell_quads(uint64_t n, const int64_t* rowptr, uint32_t* quads, hipStream_t stream) {
    const uint64_t nchunk = (n + 63) / 64;
    if (!nchunk) return;
    k_sell_quads<<<(unsigned)((nchunk + 3) / 4), 256, 0, stream>>>(n, (uint32_t)nchunk, rowptr, quads);
    CF_HIP(hipGetLastError());
}

void launch_sell_fill(uint64_t n, const int64_t* rowptr, const uint32_t* col, const double* val, const float* val32,
                      const int64_t* qoff, uint32_t* scol, double* sval, float* sval32, hipStream_t stream) {
    const uint64_t nchunk = (n + 63) / 64;
    if (!nchunk) return;
    k_sell_fill<<<(unsigned)nchunk, 256, 0, stream>>>(n, (uint32_t)nchunk, rowptr, col, val, val32, qoff, scol, sval,
                                                      sval32);
    CF_HIP(hipGetLastError());
}

}  // namespace cf
