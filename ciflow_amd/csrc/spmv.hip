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

}  // namespace cf
