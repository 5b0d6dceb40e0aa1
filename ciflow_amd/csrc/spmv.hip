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

}  // namespace cf
