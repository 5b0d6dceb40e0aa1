// Krylov-vector kernels for the GPU-resident Lanczos (the dense work ARPACK's
// dsaitr/dsapps do on the host in the reference, SparseMatrix_CRS.cpp:441-458).
// All HBM-bound: a basis of k vectors of length n is stored column-major
// (vector i at Vb + i*n).
#include "kernels.h"
#include "krylov_dev.h"

#include <algorithm>

namespace cf {

constexpr int kDotChunk = 8;  // vectors per block row
constexpr int kDotThreads = 256;

int multi_dot_blocks(uint64_t n) {
    uint64_t nb = (n + 2047) / 2048;
    if (nb < 1) nb = 1;
    if (nb > 1024) nb = 1024;
    return (int)nb;
}

// partial[i * nblocks + bx] = sum over this block's elements of Vb[i] . w
__global__ void __launch_bounds__(kDotThreads) k_multi_dot(uint64_t n, int k, const double* __restrict__ Vb,
                                                           const double* __restrict__ w, double* __restrict__ partial) {
    __shared__ double red[kDotChunk][kDotThreads / 64];
    const int i0 = blockIdx.y * kDotChunk;
    double acc[kDotChunk];
#pragma unroll
    for (int j = 0; j < kDotChunk; ++j) acc[j] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kDotThreads;
    for (uint64_t q = (uint64_t)blockIdx.x * kDotThreads + threadIdx.x; q < n; q += stride) {
        const double wq = w[q];
#pragma unroll
        for (int j = 0; j < kDotChunk; ++j)
            if (i0 + j < k) acc[j] += Vb[(uint64_t)(i0 + j) * n + q] * wq;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kDotChunk; ++j) {
        double s = wave_sum(acc[j]);
        if (lane == 0) red[j][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < kDotChunk && i0 + (int)threadIdx.x < k) {
        double s = 0.0;
        for (int q = 0; q < kDotThreads / 64; ++q) s += red[threadIdx.x][q];
        partial[(uint64_t)(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = s;
    }
}

void launch_multi_dot(uint64_t n, int k, const double* Vb, const double* w, double* partial, int nblocks,
                      hipStream_t stream) {
    if (k <= 0) return;
    dim3 grid((unsigned)nblocks, (unsigned)((k + kDotChunk - 1) / kDotChunk));
    k_multi_dot<<<grid, kDotThreads, 0, stream>>>(n, k, Vb, w, partial);
    CF_HIP(hipGetLastError());
}

// out[i] = sum_b partial[i*nblocks + b], fixed order (deterministic).
__global__ void k_reduce_partials(int k, int nblocks, const double* __restrict__ partial, double* __restrict__ out) {
    const int i = blockIdx.x;
    if (i >= k) return;
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += 64) s += partial[(uint64_t)i * nblocks + b];
    s = wave_sum(s);
    if (threadIdx.x == 0) out[i] = s;
}

void launch_reduce_partials(int k, int nblocks, const double* partial, double* out, hipStream_t stream) {
    if (k <= 0) return;
    k_reduce_partials<<<k, 64, 0, stream>>>(k, nblocks, partial, out);
    CF_HIP(hipGetLastError());
}

// w -= Vb[:, 0:k] c; with gate = {before, after} (two |w|^2 sums in device
// memory), nothing is done when after > eta2 before (the selective second
// Gram-Schmidt pass; every block reads the same two values)
__global__ void __launch_bounds__(256) k_multi_axpy(uint64_t n, int k, const double* __restrict__ Vb,
                                                    const double* __restrict__ c, double* __restrict__ w,
                                                    const double* __restrict__ gate_before,
                                                    const double* __restrict__ gate_after, double eta2) {
    __shared__ double cs[64];
    if (gate_before && *gate_after > eta2 * *gate_before) return;
    if (threadIdx.x < 64) cs[threadIdx.x] = (int)threadIdx.x < k ? c[threadIdx.x] : 0.0;
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += cs[i] * Vb[(uint64_t)i * n + q];
        w[q] -= s;
    }
}

void launch_multi_axpy(uint64_t n, int k, const double* Vb, const double* c, double* w, hipStream_t stream,
                       const double* gate_before, const double* gate_after, double eta) {
    if (k <= 0 || !n) return;
    if (k > 64) throw Error(ERR_ARG, "multi_axpy: k > 64");
    uint64_t grid = (n + 255) / 256;
    if (grid > 8192) grid = 8192;
    k_multi_axpy<<<(unsigned)grid, 256, 0, stream>>>(n, k, Vb, c, w, gate_before, gate_after, eta * eta);
    CF_HIP(hipGetLastError());
}

// w -= a * v1 + b * v0 with a, b read from device memory (the three-term
// Lanczos step before the re-orthogonalisation pass).
__global__ void k_local3(uint64_t n, const double* __restrict__ a, const double* __restrict__ v1,
                         const double* __restrict__ b, const double* __restrict__ v0, double* __restrict__ w) {
    const double ca = *a, cb = *b;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
        w[q] -= ca * v1[q] + cb * v0[q];
}

void launch_local3(uint64_t n, const double* a, const double* v1, const double* b, const double* v0, double* w,
                   hipStream_t stream) {
    if (!n) return;
    uint64_t grid = (n + 255) / 256;
    if (grid > 8192) grid = 8192;
    k_local3<<<(unsigned)grid, 256, 0, stream>>>(n, a, v1, b, v0, w);
    CF_HIP(hipGetLastError());
}

// dst = src * s
__global__ void k_scale_copy(uint64_t n, const double* __restrict__ src, double s, double* __restrict__ dst) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = src[q] * s;
}

void launch_scale_copy(uint64_t n, const double* src, double s, double* dst, hipStream_t stream) {
    if (!n) return;
    uint64_t grid = (n + 255) / 256;
    if (grid > 8192) grid = 8192;
    k_scale_copy<<<(unsigned)grid, 256, 0, stream>>>(n, src, s, dst);
    CF_HIP(hipGetLastError());
}

// End of a Lanczos step without a host round trip: beta = sqrt(max(0, |w|^2)),
// v_next = w * (1 / beta) (zero when beta is zero), alpha_j = c1[j] + c2[j]
// (the two Gram-Schmidt coefficients on v_j).  Same arithmetic as the host
// path (sqrt, 1/beta, then one multiply per element).
__global__ void k_finish_step(uint64_t n, int j, const double* __restrict__ c1, const double* __restrict__ c2,
                              const double* __restrict__ wnorm2, const double* __restrict__ w,
                              double* __restrict__ vnext, double* __restrict__ alpha, double* __restrict__ beta) {
    const double b = sqrt(fmax(0.0, *wnorm2));
    const double s = b > 0.0 ? 1.0 / b : 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) vnext[q] = w[q] * s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        alpha[j] = c1[j] + c2[j];
        beta[j] = b;
    }
}

void launch_finish_step(uint64_t n, int j, const double* c1, const double* c2, const double* wnorm2,
                        const double* w, double* vnext, double* alpha, double* beta, hipStream_t stream) {
    uint64_t grid = (n + 255) / 256;
    if (grid > 8192) grid = 8192;
    if (grid < 1) grid = 1;
    k_finish_step<<<(unsigned)grid, 256, 0, stream>>>(n, j, c1, c2, wnorm2, w, vnext, alpha, beta);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Step kernels with the reduction of the preceding product folded in.
// ---------------------------------------------------------------------------
int fused_dot_blocks(uint64_t n) {
    uint64_t want = (n + 4095) / 4096;  // <= 4 elements per thread
    int nb = 1;
    while ((uint64_t)nb < want && nb < 64) nb <<= 1;
    return nb;
}

int fused_axpy_blocks(uint64_t n) {
    uint64_t nb = (n + 255) / 256;
    if (nb > 256) nb = 256;
    if (nb < 1) nb = 1;
    return (int)nb;
}

__global__ void __launch_bounds__(1024) k_multi_dot_wide(uint64_t n, int k, const double* __restrict__ Vb,
                                                         const double* __restrict__ w,
                                                         double* __restrict__ partial, const double* skip_before,
                                                         const double* skip_after, int skip_np, double eta2) {
    __shared__ double red[kDotChunk][16];
    if (skip_before && dgks_skip(skip_before, skip_after, skip_np, eta2)) return;
    const int i0 = blockIdx.y * kDotChunk;
    double acc[kDotChunk];
#pragma unroll
    for (int j = 0; j < kDotChunk; ++j) acc[j] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * 1024;
    for (uint64_t q = (uint64_t)blockIdx.x * 1024 + threadIdx.x; q < n; q += stride) {
        const double wq = w[q];
#pragma unroll
        for (int j = 0; j < kDotChunk; ++j)
            if (i0 + j < k) acc[j] += Vb[(uint64_t)(i0 + j) * n + q] * wq;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < kDotChunk; ++j) {
        double s = wave_sum(acc[j]);
        if (lane == 0) red[j][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x < kDotChunk && i0 + (int)threadIdx.x < k) {
        double s = 0.0;
        for (int q = 0; q < 16; ++q) s += red[threadIdx.x][q];
        partial[(uint64_t)(i0 + threadIdx.x) * gridDim.x + blockIdx.x] = s;
    }
}

void launch_multi_dot_wide(uint64_t n, int k, const double* Vb, const double* w, double* partial, int nblocks,
                           hipStream_t stream, const double* skip_before, const double* skip_after, double eta) {
    if (k <= 0) return;
    if (nblocks < 1 || nblocks > 64 || (nblocks & (nblocks - 1)))
        throw Error(ERR_ARG, "multi_dot_wide: nblocks must be a power of two <= 64");
    dim3 grid((unsigned)nblocks, (unsigned)((k + kDotChunk - 1) / kDotChunk));
    k_multi_dot_wide<<<grid, 1024, 0, stream>>>(n, k, Vb, w, partial, skip_before, skip_after,
                                                fused_axpy_blocks(n), eta * eta);
    CF_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_local3_r(uint64_t n, int G, const double* __restrict__ partial,
                                                  const double* __restrict__ v1, const double* __restrict__ b_in,
                                                  const double* __restrict__ v0, double* __restrict__ w,
                                                  double* __restrict__ a_out) {
    __shared__ double cs[1];
    reduce_coeffs(1, G, partial, cs);
    __syncthreads();
    const double ca = cs[0], cb = *b_in;
    if (blockIdx.x == 0 && threadIdx.x == 0) *a_out = ca;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
        w[q] -= ca * v1[q] + cb * v0[q];
}

void launch_local3_r(uint64_t n, int nblocks, const double* partial, const double* v1, const double* b_in,
                     const double* v0, double* w, double* a_out, hipStream_t stream) {
    if (!n) return;
    k_local3_r<<<(unsigned)fused_axpy_blocks(n), 256, 0, stream>>>(n, nblocks, partial, v1, b_in, v0, w, a_out);
    CF_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_multi_axpy_r(uint64_t n, int k, int G, const double* __restrict__ partial,
                                                      const double* __restrict__ Vb, double* __restrict__ w,
                                                      double* __restrict__ c_out, double* __restrict__ pnorm,
                                                      const double* skip_before, const double* skip_after,
                                                      double eta2) {
    __shared__ double cs[64];
    __shared__ double red[4];
    if (skip_before && dgks_skip(skip_before, skip_after, (int)gridDim.x, eta2)) return;
    reduce_coeffs(k, G, partial, cs);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x < k) c_out[threadIdx.x] = cs[threadIdx.x];
    double nrm = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        double s = 0.0;
        for (int i = 0; i < k; ++i) s += cs[i] * Vb[(uint64_t)i * n + q];
        const double wn = w[q] - s;
        w[q] = wn;
        nrm += wn * wn;
    }
    nrm = wave_sum(nrm);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = nrm;
    __syncthreads();
    if (threadIdx.x == 0) pnorm[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_multi_axpy_r(uint64_t n, int k, int nblocks, const double* partial, const double* Vb, double* w,
                         double* c_out, double* pnorm, hipStream_t stream, const double* skip_before,
                         const double* skip_after, double eta) {
    if (!n || k <= 0) return;
    if (k > 64 || nblocks < 1 || nblocks > 64 || (nblocks & (nblocks - 1)) || k * nblocks > 16 * 256)
        throw Error(ERR_ARG, "multi_axpy_r: k <= 64 and a power-of-two nblocks <= 64 with k*nblocks <= 4096");
    k_multi_axpy_r<<<(unsigned)fused_axpy_blocks(n), 256, 0, stream>>>(n, k, nblocks, partial, Vb, w, c_out, pnorm,
                                                                       skip_before, skip_after, eta * eta);
    CF_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_finish_step_r(uint64_t n, int j, int np, const double* __restrict__ c1,
                                                       const double* __restrict__ c2,
                                                       const double* __restrict__ pnorm,
                                                       const double* __restrict__ w, double* __restrict__ vnext,
                                                       double* __restrict__ alpha, double* __restrict__ beta) {
    __shared__ double red[4];
    double x = (int)threadIdx.x < np ? pnorm[threadIdx.x] : 0.0;
    x = wave_sum(x);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
    __syncthreads();
    const double w2 = (red[0] + red[1]) + (red[2] + red[3]);
    const double b = sqrt(fmax(0.0, w2));
    const double s = b > 0.0 ? 1.0 / b : 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) vnext[q] = w[q] * s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        alpha[j] = c1[j] + c2[j];
        beta[j] = b;
    }
}

void launch_finish_step_r(uint64_t n, int j, const double* c1, const double* c2, const double* pnorm,
                          const double* w, double* vnext, double* alpha, double* beta, hipStream_t stream) {
    const int np = fused_axpy_blocks(n);
    k_finish_step_r<<<(unsigned)np, 256, 0, stream>>>(n, j, np, c1, c2, pnorm, w, vnext, alpha, beta);
    CF_HIP(hipGetLastError());
}

// In place: Vb[:, 0:k] = Vb[:, 0:m] * Y[:, 0:k]   (Y is m x m row-major), on
// FP64 MFMA as out^T = Y^T V^T per tile of 16 basis elements, so a wavefront's
// output tile is (16 new vectors) x (16 elements) and its loads and stores are
// 16 consecutive elements of one vector.  A wavefront reads all m inputs of
// its 16 elements before it writes, and tiles are disjoint, so the update is
// in place.  Reads m n 8 B, writes k n 8 B.
typedef double v4dk __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_rotate_mfma(uint64_t n, int m, int k, double* __restrict__ Vb,
                                                     const double* __restrict__ Y) {
    __shared__ double ys[64 * 64];
    for (int i = threadIdx.x; i < m * m; i += blockDim.x) ys[i] = Y[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, r = lane & 15, kq = lane >> 4;
    const uint64_t q0 = ((uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
    if (q0 >= n) return;
    const uint64_t q = q0 + r;
    const bool inq = q < n;
    v4dk acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[t] = v4dk{0.0, 0.0, 0.0, 0.0};
    const int ct = (k + 15) / 16;
    for (int i0 = 0; i0 < m; i0 += 4) {
        const int i = i0 + kq;
        const double bv = (i < m && inq) ? Vb[(uint64_t)i * n + q] : 0.0;  // V^T[i][q]
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t < ct) {
                const int c = t * 16 + r;
                const double av = (i < m && c < k) ? ys[i * m + c] : 0.0;  // Y^T[c][i]
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
    // D[c][e]: lane e + 16 (c % 4), element c / 4 -> vector t*16 + 4 v + kq, element q0 + r
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        if (t < ct) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = t * 16 + 4 * v + kq;
                if (c < k && inq) Vb[(uint64_t)c * n + q] = acc[t][v];
            }
        }
    }
}

void launch_rotate_basis(uint64_t n, int m, int k, double* Vb, const double* Y, hipStream_t stream) {
    if (m > 64 || k > m) throw Error(ERR_ARG, "rotate_basis: m > 64 or k > m");
    if (!n || k <= 0) return;
    const uint64_t grid = (n + 63) / 64;  // 4 wavefronts x 16 elements
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "rotate_basis: n too large");
    k_rotate_mfma<<<(unsigned)grid, 256, 0, stream>>>(n, m, k, Vb, Y);
    CF_HIP(hipGetLastError());
}

// Deterministic start vector: v[q] = 1 + 0.5*(u(seed, q) - 0.5), u in [0,1).
__global__ void k_fill_start(uint64_t n, uint64_t seed, double* __restrict__ v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
        v[q] = start_element(seed, q);
}

void launch_fill_start(uint64_t n, uint64_t seed, double* v, hipStream_t stream) {
    if (!n) return;
    uint64_t grid = (n + 255) / 256;
    if (grid > 8192) grid = 8192;
    k_fill_start<<<(unsigned)grid, 256, 0, stream>>>(n, seed, v);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Davidson: Ritz vector x = V y and its image S y in one pass over both
// bases, then the residual r = S y - theta x, the diagonal-preconditioned
// correction t = r / (theta - diag) (|theta - diag| clamped to >= shift_min;
// t = r when shift_min == 0: a plain Krylov expansion) and norm parts.
// ---------------------------------------------------------------------------
// Large-n Davidson step in two passes over the basis instead of four: V
// values are kept in registers so that each pass also forms the next dot
// products.  The basis size K is a template parameter so that every load of
// an element's K values is issued before the first is used.  Block parts go
// to partial[row * gridDim.x + block] (the layout of k_multi_dot, reduced by
// k_reduce_partials): rows 0..K-1 the dots, row_a the first norm, row_b the
// second (skipped when < 0).
template <int K>
__device__ __forceinline__ void block_parts(const double (&acc)[K + 2], int row_a, int row_b,
                                            double* __restrict__ partial) {
    __shared__ double red[K + 2][kDotThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int j = 0; j < K + 2; ++j) {
        const double s = wave_sum(acc[j]);
        if (lane == 0) red[j][wid] = s;
    }
    __syncthreads();
    const int i = threadIdx.x;
    const int row = i < K ? i : (i == K ? row_a : row_b);
    if (i < K + 2 && row >= 0) {
        double s = 0.0;
        for (int q = 0; q < kDotThreads / 64; ++q) s += red[i][q];
        partial[(uint64_t)row * gridDim.x + blockIdx.x] = s;
    }
}

// x = V y, r = S y - theta x, the correction t; parts of V^T t (rows
// 0..K-1), |t|^2 (row row_t) and |r|^2 (row row_r)
template <int K>
__global__ void __launch_bounds__(kDotThreads) k_combine_res_dot(uint64_t n, const double* __restrict__ V,
                                                                 const double* __restrict__ S, DavCoeffs y,
                                                                 double theta, double shift_min,
                                                                 const double* __restrict__ diag,
                                                                 double* __restrict__ x, double* __restrict__ t,
                                                                 double* __restrict__ partial, int row_t, int row_r) {
    double acc[K + 2];
#pragma unroll
    for (int j = 0; j < K + 2; ++j) acc[j] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kDotThreads;
    for (uint64_t q = (uint64_t)blockIdx.x * kDotThreads + threadIdx.x; q < n; q += stride) {
        double vq[K], sq[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            vq[i] = V[(uint64_t)i * n + q];
            sq[i] = S[(uint64_t)i * n + q];
        }
        const double dq = shift_min > 0.0 ? diag[q] : 0.0;
        double a = 0.0, b = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) {
            a += y.c[i] * vq[i];
            b += y.c[i] * sq[i];
        }
        x[q] = a;
        const double r = b - theta * a;
        double tq = r;
        if (shift_min > 0.0) {
            double den = theta - dq;
            if (fabs(den) < shift_min) den = den < 0.0 ? -shift_min : shift_min;
            tq = r / den;
        }
        t[q] = tq;
#pragma unroll
        for (int i = 0; i < K; ++i) acc[i] += vq[i] * tq;
        acc[K] += tq * tq;
        acc[K + 1] += r * r;
    }
    block_parts<K>(acc, row_t, row_r, partial);
}

// t -= V c (c: K coefficients in device memory), then parts of V^T t (rows
// 0..K-1) and |t|^2 (row row_t)
template <int K>
__global__ void __launch_bounds__(kDotThreads) k_axpy_dot(uint64_t n, const double* __restrict__ V,
                                                          const double* __restrict__ c, double* __restrict__ t,
                                                          double* __restrict__ partial, int row_t) {
    __shared__ double cs[K];
    if (threadIdx.x < K) cs[threadIdx.x] = c[threadIdx.x];
    __syncthreads();
    double acc[K + 2];
#pragma unroll
    for (int j = 0; j < K + 2; ++j) acc[j] = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * kDotThreads;
    for (uint64_t q = (uint64_t)blockIdx.x * kDotThreads + threadIdx.x; q < n; q += stride) {
        double vq[K];
#pragma unroll
        for (int i = 0; i < K; ++i) vq[i] = V[(uint64_t)i * n + q];
        double s = 0.0;
#pragma unroll
        for (int i = 0; i < K; ++i) s += cs[i] * vq[i];
        const double tq = t[q] - s;
        t[q] = tq;
#pragma unroll
        for (int i = 0; i < K; ++i) acc[i] += vq[i] * tq;
        acc[K] += tq * tq;
    }
    block_parts<K>(acc, row_t, -1, partial);
}

template <int K>
static void combine_res_dot_k(uint64_t n, const double* V, const double* S, const DavCoeffs& y, double theta,
                              double shift_min, const double* diag, double* x, double* t, double* partial,
                              int nblocks, int row_t, int row_r, hipStream_t stream) {
    k_combine_res_dot<K><<<(unsigned)nblocks, kDotThreads, 0, stream>>>(n, V, S, y, theta, shift_min, diag, x, t,
                                                                        partial, row_t, row_r);
}

template <int K>
static void axpy_dot_k(uint64_t n, const double* V, const double* c, double* t, double* partial, int nblocks,
                       int row_t, hipStream_t stream) {
    k_axpy_dot<K><<<(unsigned)nblocks, kDotThreads, 0, stream>>>(n, V, c, t, partial, row_t);
}

// one instance per basis size 1..kDavMaxBasis
#define CF_DAV_SIZES(X)                                                                                         \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) \
        X(20) X(21) X(22) X(23) X(24)
static_assert(kDavMaxBasis == 24, "CF_DAV_SIZES lists basis sizes 1..24");

void launch_combine_res_dot(uint64_t n, int k, const double* V, const double* S, const double* y, double theta,
                            double shift_min, const double* diag, double* x, double* t, double* partial, int nblocks,
                            int row_t, int row_r, hipStream_t stream) {
    if (k < 1 || k > kDavMaxBasis) throw Error(ERR_ARG, "combine_res_dot: 1 <= k <= 24");
    if (row_t < k || row_r < k || row_t == row_r || row_t >= 64 || row_r >= 64)
        throw Error(ERR_ARG, "combine_res_dot: norm rows must lie in [k, 64) and differ");
    DavCoeffs yc{};
    for (int i = 0; i < k; ++i) yc.c[i] = y[i];
    switch (k) {
#define CF_CASE(K)                                                                                       \
    case K:                                                                                              \
        combine_res_dot_k<K>(n, V, S, yc, theta, shift_min, diag, x, t, partial, nblocks, row_t, row_r, \
                             stream);                                                                    \
        break;
        CF_DAV_SIZES(CF_CASE)
#undef CF_CASE
    }
    CF_HIP(hipGetLastError());
}

void launch_axpy_dot(uint64_t n, int k, const double* V, const double* c, double* t, double* partial, int nblocks,
                     int row_t, hipStream_t stream) {
    if (k < 1 || k > kDavMaxBasis) throw Error(ERR_ARG, "axpy_dot: 1 <= k <= 24");
    if (row_t < k || row_t >= 64) throw Error(ERR_ARG, "axpy_dot: norm row must lie in [k, 64)");
    switch (k) {
#define CF_CASE(K)                                                        \
    case K:                                                               \
        axpy_dot_k<K>(n, V, c, t, partial, nblocks, row_t, stream);       \
        break;
        CF_DAV_SIZES(CF_CASE)
#undef CF_CASE
    }
    CF_HIP(hipGetLastError());
}
#undef CF_DAV_SIZES

// Fused forms for n <= kFusedMaxN (one workgroup per fused_axpy_blocks slot):
// x, hx, the correction t and this block's part of |r|^2 in one pass ...
__global__ void __launch_bounds__(256) k_combine_residual(uint64_t n, int k, const double* __restrict__ V,
                                                          const double* __restrict__ S, DavCoeffs y,
                                                          double theta, double shift_min,
                                                          const double* __restrict__ diag, double* __restrict__ x,
                                                          double* __restrict__ hx, double* __restrict__ t,
                                                          double* __restrict__ rparts, double* __restrict__ tparts) {
    __shared__ double ys[kDavMaxBasis];
    __shared__ double red[2][4];
    if (threadIdx.x < kDavMaxBasis) ys[threadIdx.x] = (int)threadIdx.x < k ? y.c[threadIdx.x] : 0.0;
    __syncthreads();
    double acc = 0.0, acct = 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < k; ++i) {
            a += ys[i] * V[(uint64_t)i * n + q];
            b += ys[i] * S[(uint64_t)i * n + q];
        }
        x[q] = a;
        hx[q] = b;
        const double r = b - theta * a;
        acc += r * r;
        double tq = r;
        if (shift_min > 0.0) {
            double den = theta - diag[q];
            if (fabs(den) < shift_min) den = den < 0.0 ? -shift_min : shift_min;
            tq = r / den;
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
        rparts[blockIdx.x] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        tparts[blockIdx.x] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}

void launch_combine_residual(uint64_t n, int k, const double* V, const double* S, const double* y, double theta,
                             double shift_min, const double* diag, double* x, double* hx, double* t, double* rparts,
                             double* tparts, hipStream_t stream) {
    if (k < 1 || k > kDavMaxBasis) throw Error(ERR_ARG, "combine_residual: 1 <= k <= 24");
    DavCoeffs yc{};
    for (int i = 0; i < k; ++i) yc.c[i] = y[i];
    k_combine_residual<<<(unsigned)fused_axpy_blocks(n), 256, 0, stream>>>(n, k, V, S, yc, theta, shift_min, diag,
                                                                           x, hx, t, rparts, tparts);
    CF_HIP(hipGetLastError());
}

// ... and dst = src / |src| from the per-block |src|^2 parts of the last
// multi_axpy_r (np parts, fixed order); block 0 also stores the sums of
// rparts (|r|^2) and pnorm (|src|^2) into out[0], out[1].
__global__ void __launch_bounds__(256) k_scale_rnorm_parts(uint64_t n, const double* __restrict__ src,
                                                           const double* __restrict__ pnorm,
                                                           const double* __restrict__ rparts, int np,
                                                           double* __restrict__ dst, double* __restrict__ out,
                                                           const double* tparts, const double* pnorm2,
                                                           double eta2) {
    __shared__ double red[2][4];
    // with a selective second pass: its parts, unless the first pass sufficed
    if (tparts && !dgks_skip(tparts, pnorm, np, eta2)) pnorm = pnorm2;
    double a = (int)threadIdx.x < np ? pnorm[threadIdx.x] : 0.0;
    double r = (int)threadIdx.x < np ? rparts[threadIdx.x] : 0.0;
    a = wave_sum(a);
    r = wave_sum(r);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = r;
    }
    __syncthreads();
    const double q2 = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
    const double s = q2 > 1e-28 ? 1.0 / sqrt(q2) : 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = src[q] * s;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out[0] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        out[1] = q2;
    }
}

void launch_scale_rnorm_parts(uint64_t n, const double* src, const double* pnorm, const double* rparts, double* dst,
                              double* out, hipStream_t stream, const double* tparts, const double* pnorm2,
                              double eta) {
    const int np = fused_axpy_blocks(n);
    k_scale_rnorm_parts<<<(unsigned)np, 256, 0, stream>>>(n, src, pnorm, rparts, np, dst, out, tparts, pnorm2,
                                                          eta * eta);
    CF_HIP(hipGetLastError());
}

// dst = src / sqrt(*norm2) (zeros when *norm2 <= 1e-28): normalisation by a
// norm that stays on the device
__global__ void k_scale_rnorm(uint64_t n, const double* __restrict__ src, const double* __restrict__ norm2,
                              double* __restrict__ dst) {
    const double q2 = *norm2;
    const double s = q2 > 1e-28 ? 1.0 / sqrt(q2) : 0.0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) dst[q] = src[q] * s;
}

void launch_scale_rnorm(uint64_t n, const double* src, const double* norm2, double* dst, hipStream_t stream) {
    const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, 2048);
    k_scale_rnorm<<<grid, 256, 0, stream>>>(n, src, norm2, dst);
    CF_HIP(hipGetLastError());
}

void launch_copy(uint64_t n, const double* src, double* dst, hipStream_t stream) {
    if (!n) return;
    CF_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

}  // namespace cf
