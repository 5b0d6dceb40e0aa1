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

// One workgroup per (alpha string ia, beta chunk, part p of P): the part covers
// the alpha-list entries whose segments start in positions [p*R/P, (p+1)*R/P),
// so the P parts of a row split its work evenly and the grid has P times more
// workgroups (better filling of the last dispatch round).  Staged blocks are
// loaded G at a time (two barriers per group).  With P > 1 each part writes
// its partial row sum to part[p*n + row] and k_ell_sum_parts adds them in part
// order (deterministic).
template <typename VT, int NT>
__global__ void __launch_bounds__(NT) k_spmv_ell(EllArgs a, uint32_t chunks, uint32_t parts, uint32_t group,
                                                 const VT* __restrict__ ell, const double* __restrict__ x,
                                                 double* __restrict__ y, double* __restrict__ part) {
    extern __shared__ double xs[];  // group * db
    const uint64_t db = a.db;
    const uint32_t per = chunks * parts;
    const uint64_t ia = blockIdx.x / per;
    const uint32_t rem = blockIdx.x - (uint32_t)ia * per;
    const uint32_t pi = rem / chunks, ci = rem - pi * chunks;
    const uint64_t ib = (uint64_t)ci * NT + threadIdx.x;
    const bool act = ib < db;
    const uint32_t* so = a.segoff + ia * (uint64_t)(a.Ea + 1);
    const uint64_t R = so[a.Ea];
    const Exc* la = a.fa + ia * (uint64_t)a.Ea;
    // alpha-list range of this part: segments starting in [t_lo, t_hi)
    const uint64_t t_lo = R * pi / parts, t_hi = R * (pi + 1) / parts;
    int k0 = 0, k1 = 0;
    {
        int lo = 0, hi = a.Ea;  // first k with so[k] >= t_lo
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (so[mid] < t_lo) lo = mid + 1;
            else hi = mid;
        }
        k0 = lo;
        lo = k0;
        hi = a.Ea;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (so[mid] < t_hi) lo = mid + 1;
            else hi = mid;
        }
        k1 = lo;
    }
    const VT* v = ell + ia * R * db + (act ? ib : 0);
    const uint16_t* js = a.jsub + (act ? ib : 0);
    const uint16_t* jf = a.jfull + (act ? ib : 0);
    double acc = 0.0;
    int kst = k0;      // next alpha entry to consider for staging
    int staged = 0;    // blocks in the current group
    int slot = 0;      // next block of the group to use
    for (int k = k0; k < k1; ++k) {
        const Exc ea = la[k];
        const uint64_t t0 = so[k];
        if (ea.type == 2) {  // alpha double: column (Ja, ib)
            if (act) acc += ldv(v + t0 * db) * x[(uint64_t)ea.J * db + ib];
            continue;
        }
        if (slot == staged) {  // load the next group of staged blocks (uniform decision)
            __syncthreads();
            int g = 0;
            for (; kst < k1 && g < (int)group; ++kst) {
                const Exc e2 = la[kst];
                if (e2.type == 2) continue;
                const double* src = x + (uint64_t)e2.J * db;
                for (uint64_t i = threadIdx.x; i < db; i += NT) xs[(uint64_t)g * db + i] = src[i];
                ++g;
            }
            staged = g;
            slot = 0;
            __syncthreads();
        }
        const double* xb = xs + (uint64_t)slot * db;
        ++slot;
        if (!act) continue;
        const VT* vk = v + t0 * db;
        const int n = ea.type == 1 ? a.Sb : a.Eb;
        const uint16_t* jt = ea.type == 1 ? js : jf;
        int u = 0;
        for (; u + 4 <= n; u += 4) {
            const double v0 = ldv(vk + (uint64_t)u * db), v1 = ldv(vk + (uint64_t)(u + 1) * db),
                         v2 = ldv(vk + (uint64_t)(u + 2) * db), v3 = ldv(vk + (uint64_t)(u + 3) * db);
            const uint16_t j0 = jt[(uint64_t)u * db], j1 = jt[(uint64_t)(u + 1) * db],
                           j2 = jt[(uint64_t)(u + 2) * db], j3 = jt[(uint64_t)(u + 3) * db];
            acc += v0 * xb[j0];
            acc += v1 * xb[j1];
            acc += v2 * xb[j2];
            acc += v3 * xb[j3];
        }
        for (; u < n; ++u) acc += ldv(vk + (uint64_t)u * db) * xb[jt[(uint64_t)u * db]];
    }
    if (!act) return;
    if (parts == 1) y[ia * db + ib] = acc;
    else part[(uint64_t)pi * (a.da * db) + ia * db + ib] = acc;
}

__global__ void k_ell_sum_parts(uint64_t n, uint32_t parts, const double* __restrict__ part, double* __restrict__ y) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        double s = part[q];
        for (uint32_t p = 1; p < parts; ++p) s += part[(uint64_t)p * n + q];
        y[q] = s;
    }
}

template <typename VT>
static void launch_ell(const EllArgs& a, int threads, const VT* ell, const double* x, double* y, double* part,
                       int parts, hipStream_t s) {
    if (a.db > kEllMaxDb) throw Error(ERR_UNSUPPORTED, "ELL SpMV needs d_beta <= 8192");
    const int nt = threads > 0 ? threads : 512;
    if (parts < 1 || parts > kEllMaxParts) throw Error(ERR_ARG, "ELL SpMV parts must be in [1, 8]");
    if (parts > 1 && !part) throw Error(ERR_ARG, "ELL SpMV with parts > 1 needs a partial-sum buffer");
    const uint64_t chunks = (a.db + nt - 1) / nt;
    const uint64_t grid = a.da * chunks * (uint64_t)parts;
    if (!grid) return;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "ELL SpMV grid too large");
    // staged blocks per group: up to 64 KB of LDS, at most 16
    uint64_t group = 65536 / (a.db * sizeof(double));
    if (group < 1) group = 1;
    if (group > 16) group = 16;
    const size_t lds = group * a.db * sizeof(double);
    const uint32_t c = (uint32_t)chunks, P = (uint32_t)parts, G = (uint32_t)group;
    switch (nt) {
        case 256: k_spmv_ell<VT, 256><<<(unsigned)grid, 256, lds, s>>>(a, c, P, G, ell, x, y, part); break;
        case 512: k_spmv_ell<VT, 512><<<(unsigned)grid, 512, lds, s>>>(a, c, P, G, ell, x, y, part); break;
        case 1024: k_spmv_ell<VT, 1024><<<(unsigned)grid, 1024, lds, s>>>(a, c, P, G, ell, x, y, part); break;
        default: throw Error(ERR_ARG, "ELL SpMV workgroup size must be 256, 512 or 1024");
    }
    CF_HIP(hipGetLastError());
    if (parts > 1) {
        const uint64_t n = a.da * a.db;
        uint64_t g2 = (n + 255) / 256;
        if (g2 > 8192) g2 = 8192;
        k_ell_sum_parts<<<(unsigned)g2, 256, 0, s>>>(n, (uint32_t)parts, part, y);
        CF_HIP(hipGetLastError());
    }
}

void launch_spmv_ell_f64(const EllArgs& a, int threads, int parts, const double* ell, const double* x, double* y,
                         double* part, hipStream_t s) {
    launch_ell<double>(a, threads, ell, x, y, part, parts, s);
}
void launch_spmv_ell_f32(const EllArgs& a, int threads, int parts, const float* ell, const double* x, double* y,
                         double* part, hipStream_t s) {
    launch_ell<float>(a, threads, ell, x, y, part, parts, s);
}

}  // namespace cf
