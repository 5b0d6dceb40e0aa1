// FCI SpMV on position-major complete rows (replaces SparseMatrix_CRS_Sym::mvprod,
// SparseMatrix_CRS.cpp:700-723, for FCI matrices).
//
// Every FCI row (ia, ib) has the same positions for a given alpha string ia:
// its alpha excitation list (sorted by target Ja) opens one segment per entry
// -- alpha self: the full beta list of ib; alpha single: the beta self+singles
// list of ib; alpha double: the single column (Ja, ib).  The values of all rows
// of ia are stored position-major, ell[(ia*R + t)*db + ib] (zeros where the CRS
// drops |v| <= 1e-14), so a thread per beta string reads its values with
// wavefront-contiguous loads.  Columns are not stored: the beta targets come
// from two small index tables shared by every alpha string (L2-resident), and
// the x block of each alpha target, x[Ja, :], is staged in LDS once per
// workgroup and read there, so the per-nonzero scattered HBM/L2 gather of the
// CRS kernel becomes an LDS read.  Alpha doubles read x[Ja, ib], contiguous
// across the wavefront.
//
// HBM-bound: algorithmic bytes per call
//   B = nnz_complete * s_val + 8n (y) + x blocks staged per workgroup,
// with nnz_complete = n * R (config 2: 1,553,018,544 * 8 B = 12.4 GB FP64).
#include "spmv_ell.h"

namespace cf {

__global__ void k_ell_index(uint64_t db, int Sb, int Eb, const Exc* __restrict__ subb,
                            const Exc* __restrict__ fullb, uint16_t* __restrict__ jsub,
                            uint16_t* __restrict__ jfull) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t ns = db * (uint64_t)Sb, nf = db * (uint64_t)Eb;
    if (t < ns) {
        const uint64_t ib = t / Sb, u = t - ib * Sb;
        jsub[u * db + ib] = (uint16_t)subb[t].J;
    }
    if (t < nf) {
        const uint64_t ib = t / Eb, u = t - ib * Eb;
        jfull[u * db + ib] = (uint16_t)fullb[t].J;
    }
}

void launch_ell_index(uint64_t db, int Sb, int Eb, const Exc* subb, const Exc* fullb, uint16_t* jsub,
                      uint16_t* jfull, hipStream_t s) {
    if (db > 65535) throw Error(ERR_UNSUPPORTED, "ELL index tables need d_beta <= 65535");
    const uint64_t tot = db * (uint64_t)(Eb > Sb ? Eb : Sb);
    if (!tot) return;
    k_ell_index<<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(db, Sb, Eb, subb, fullb, jsub, jfull);
    CF_HIP(hipGetLastError());
}

template <typename VT>
__device__ __forceinline__ double ldv(const VT* p) {
    return (double)__builtin_nontemporal_load(p);
}

template <typename VT, int NT>
__global__ void __launch_bounds__(NT) k_spmv_ell(EllArgs a, uint32_t chunks, const VT* __restrict__ ell,
                                                 const double* __restrict__ x, double* __restrict__ y) {
    extern __shared__ double xs[];
    const uint64_t db = a.db;
    const uint64_t ia = blockIdx.x / chunks;
    const uint64_t ib = (uint64_t)(blockIdx.x - ia * chunks) * NT + threadIdx.x;
    const bool act = ib < db;
    const uint32_t* so = a.segoff + ia * (uint64_t)(a.Ea + 1);
    const uint64_t R = so[a.Ea];
    const Exc* la = a.fa + ia * (uint64_t)a.Ea;
    const VT* v = ell + ia * R * db + (act ? ib : 0);
    const uint16_t* js = a.jsub + (act ? ib : 0);
    const uint16_t* jf = a.jfull + (act ? ib : 0);
    double acc = 0.0;
    for (int k = 0; k < a.Ea; ++k) {
        const Exc ea = la[k];
        const uint64_t t0 = so[k];
        if (ea.type == 2) {  // alpha double: column (Ja, ib)
            if (act) acc += ldv(v + t0 * db) * x[(uint64_t)ea.J * db + ib];
            continue;
        }
        __syncthreads();
        const double* src = x + (uint64_t)ea.J * db;
        for (uint64_t i = threadIdx.x; i < db; i += NT) xs[i] = src[i];
        __syncthreads();
        if (!act) continue;
        const VT* vk = v + t0 * db;
        if (ea.type == 1) {
            const int n = a.Sb;
            int u = 0;
            for (; u + 4 <= n; u += 4) {
                const double v0 = ldv(vk + (uint64_t)u * db), v1 = ldv(vk + (uint64_t)(u + 1) * db),
                             v2 = ldv(vk + (uint64_t)(u + 2) * db), v3 = ldv(vk + (uint64_t)(u + 3) * db);
                const uint16_t j0 = js[(uint64_t)u * db], j1 = js[(uint64_t)(u + 1) * db],
                               j2 = js[(uint64_t)(u + 2) * db], j3 = js[(uint64_t)(u + 3) * db];
                acc += v0 * xs[j0];
                acc += v1 * xs[j1];
                acc += v2 * xs[j2];
                acc += v3 * xs[j3];
            }
            for (; u < n; ++u) acc += ldv(vk + (uint64_t)u * db) * xs[js[(uint64_t)u * db]];
        } else {
            const int n = a.Eb;
            int u = 0;
            for (; u + 4 <= n; u += 4) {
                const double v0 = ldv(vk + (uint64_t)u * db), v1 = ldv(vk + (uint64_t)(u + 1) * db),
                             v2 = ldv(vk + (uint64_t)(u + 2) * db), v3 = ldv(vk + (uint64_t)(u + 3) * db);
                const uint16_t j0 = jf[(uint64_t)u * db], j1 = jf[(uint64_t)(u + 1) * db],
                               j2 = jf[(uint64_t)(u + 2) * db], j3 = jf[(uint64_t)(u + 3) * db];
                acc += v0 * xs[j0];
                acc += v1 * xs[j1];
                acc += v2 * xs[j2];
                acc += v3 * xs[j3];
            }
            for (; u < n; ++u) acc += ldv(vk + (uint64_t)u * db) * xs[jf[(uint64_t)u * db]];
        }
    }
    if (act) y[ia * db + ib] = acc;
}

template <typename VT>
static void launch_ell(const EllArgs& a, int threads, const VT* ell, const double* x, double* y, hipStream_t s) {
    if (a.db > kEllMaxDb) throw Error(ERR_UNSUPPORTED, "ELL SpMV needs d_beta <= 8192");
    const int nt = threads > 0 ? threads : 512;
    const uint64_t chunks = (a.db + nt - 1) / nt;
    const uint64_t grid = a.da * chunks;
    if (!grid) return;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "ELL SpMV grid too large");
    const size_t lds = a.db * sizeof(double);
    switch (nt) {
        case 256: k_spmv_ell<VT, 256><<<(unsigned)grid, 256, lds, s>>>(a, (uint32_t)chunks, ell, x, y); break;
        case 512: k_spmv_ell<VT, 512><<<(unsigned)grid, 512, lds, s>>>(a, (uint32_t)chunks, ell, x, y); break;
        case 1024: k_spmv_ell<VT, 1024><<<(unsigned)grid, 1024, lds, s>>>(a, (uint32_t)chunks, ell, x, y); break;
        default: throw Error(ERR_ARG, "ELL SpMV workgroup size must be 256, 512 or 1024");
    }
    CF_HIP(hipGetLastError());
}

void launch_spmv_ell_f64(const EllArgs& a, int threads, const double* ell, const double* x, double* y,
                         hipStream_t s) {
    launch_ell<double>(a, threads, ell, x, y, s);
}
void launch_spmv_ell_f32(const EllArgs& a, int threads, const float* ell, const double* x, double* y,
                         hipStream_t s) {
    launch_ell<float>(a, threads, ell, x, y, s);
}

}  // namespace cf
