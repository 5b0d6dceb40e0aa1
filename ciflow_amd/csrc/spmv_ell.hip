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

// ---------------------------------------------------------------------------
// Double-buffered kernel.  Per alpha string the entries are regrouped (doubles,
// then singles, then self): the doubles run first, eight columns (Ja, ib) in
// flight per thread, while the x block of the first staged segment is already
// loading into registers; each staged segment then loads the next segment's x
// block into registers while it reads its own block from LDS, and stores it
// into the other LDS buffer before the one barrier per segment.  With VEC = 2 a
// thread owns beta strings (ib, ib+1) and reads their values, indices and x
// entries with paired loads (VEC = 4: four strings, 16/32-byte loads).
// ---------------------------------------------------------------------------
__global__ void k_ell_entries(uint64_t da, int Ea, int nda, int nsa, const Exc* __restrict__ fa,
                              const uint32_t* __restrict__ segoff, uint2* __restrict__ ent) {
    const uint64_t ia = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ia >= da) return;
    const Exc* la = fa + ia * (uint64_t)Ea;
    const uint32_t* so = segoff + ia * (uint64_t)(Ea + 1);
    uint2* e = ent + ia * (uint64_t)Ea;
    int d = 0, sg = 0;
    for (int k = 0; k < Ea; ++k) {
        const uint2 w = make_uint2(so[k], la[k].J);
        const int ty = la[k].type;
        if (ty == 2) {
            if (d < nda) e[d] = w;
            ++d;
        } else if (ty == 1) {
            if (sg < nsa) e[nda + sg] = w;
            ++sg;
        } else {
            e[nda + nsa] = w;
        }
    }
}

void launch_ell_entries(uint64_t da, int Ea, int nda, int nsa, const Exc* fa, const uint32_t* segoff, uint2* ent,
                        hipStream_t s) {
    if (!da) return;
    if (nda + nsa + 1 != Ea) throw Error(ERR_STATE, "ELL entries: alpha list counts do not add up");
    k_ell_entries<<<(unsigned)((da + 255) / 256), 256, 0, s>>>(da, Ea, nda, nsa, fa, segoff, ent);
    CF_HIP(hipGetLastError());
}

template <typename T, int V>
struct Vec;
template <typename T>
struct Vec<T, 1> {
    using type = T;
};
template <typename T>
struct Vec<T, 2> {
    using type = T __attribute__((ext_vector_type(2)));
};
template <typename T>
struct Vec<T, 4> {
    using type = T __attribute__((ext_vector_type(4)));
};

template <typename T, int V>
__device__ __forceinline__ T comp(const typename Vec<T, V>::type& v, int w) {
    if constexpr (V == 1) return v;
    else return v[w];
}

// values of one position for the thread's V beta strings
template <typename VT, int V>
__device__ __forceinline__ typename Vec<VT, V>::type ldvv(const VT* p) {
    return __builtin_nontemporal_load(reinterpret_cast<const typename Vec<VT, V>::type*>(p));
}

// the V 16-bit beta targets of one position, packed low string first
template <int V>
__device__ __forceinline__ uint64_t ldj(const uint16_t* p) {
    if constexpr (V == 1) return *p;
    else if constexpr (V == 2) return *reinterpret_cast<const uint32_t*>(p);
    else return *reinterpret_cast<const uint64_t*>(p);
}

// acc[w] += sum_u v[u*db + w] * xb[j[u*db + w]], u < cnt, eight positions in flight
template <typename VT, int V>
__device__ __forceinline__ void ell_segment(const VT* __restrict__ vk, const uint16_t* __restrict__ jt, uint64_t db,
                                            int cnt, const double* __restrict__ xb, double* acc) {
    using VV = typename Vec<VT, V>::type;
    int u = 0;
    for (; u + 8 <= cnt; u += 8) {
        VV vv[8];
        uint64_t jj[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            vv[q] = ldvv<VT, V>(vk + (uint64_t)(u + q) * db);
            jj[q] = ldj<V>(jt + (uint64_t)(u + q) * db);
        }
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int w = 0; w < V; ++w)
                acc[w] += (double)comp<VT, V>(vv[q], w) * xb[(uint32_t)(jj[q] >> (16 * w)) & 0xffffu];
    }
    for (; u < cnt; ++u) {
        const VV vv = ldvv<VT, V>(vk + (uint64_t)u * db);
        const uint64_t jj = ldj<V>(jt + (uint64_t)u * db);
#pragma unroll
        for (int w = 0; w < V; ++w) acc[w] += (double)comp<VT, V>(vv, w) * xb[(uint32_t)(jj >> (16 * w)) & 0xffffu];
    }
}

template <int PM>
__device__ __forceinline__ void ell_fetch(const double* __restrict__ src, uint64_t db, double* pre) {
#pragma unroll
    for (int p = 0; p < PM; ++p) {
        const uint64_t i = threadIdx.x + (uint64_t)p * blockDim.x;
        pre[p] = i < db ? src[i] : 0.0;
    }
}

template <int PM>
__device__ __forceinline__ void ell_store(double* __restrict__ dst, uint64_t db, const double* pre) {
#pragma unroll
    for (int p = 0; p < PM; ++p) {
        const uint64_t i = threadIdx.x + (uint64_t)p * blockDim.x;
        if (i < db) dst[i] = pre[p];
    }
}

// V = 1, PM <= 4: held to 64 VGPRs so that two 1024-thread workgroups share a CU.
template <typename VT, int V, int PM>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(V == 1 && PM <= 4 ? 8 : 1))) k_spmv_ell2(EllArgs a, uint32_t chunks, const VT* __restrict__ ell,
                                                    const double* __restrict__ x, double* __restrict__ y) {
    using XV = typename Vec<double, V>::type;
    extern __shared__ double xs[];  // [2][db]
    const uint64_t db = a.db;
    const uint64_t ia = blockIdx.x / chunks;
    const uint64_t ib = ((uint64_t)(blockIdx.x - ia * chunks) * blockDim.x + threadIdx.x) * V;
    const bool act = ib < db;
    const uint64_t ibc = act ? ib : 0;
    const uint2* __restrict__ e = a.ent + ia * (uint64_t)a.Ea;
    const VT* __restrict__ v = ell + ia * a.R * db + ibc;
    const int nda = a.nda, nsa = a.nsa;

    double pre[PM];
    ell_fetch<PM>(x + (uint64_t)e[nda].y * db, db, pre);

    double acc[V];
#pragma unroll
    for (int w = 0; w < V; ++w) acc[w] = 0.0;
    if (act) {  // alpha doubles: column (Ja, ib), contiguous across the wavefront
        const double* xc = x + ibc;
        int d = 0;
        for (; d + 8 <= nda; d += 8) {
            typename Vec<VT, V>::type vv[8];
            XV xv[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const uint2 en = e[d + q];
                vv[q] = ldvv<VT, V>(v + (uint64_t)en.x * db);
                xv[q] = *reinterpret_cast<const XV*>(xc + (uint64_t)en.y * db);
            }
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int w = 0; w < V; ++w) acc[w] += (double)comp<VT, V>(vv[q], w) * comp<double, V>(xv[q], w);
        }
        for (; d < nda; ++d) {
            const uint2 en = e[d];
            const auto vv = ldvv<VT, V>(v + (uint64_t)en.x * db);
            const XV xv = *reinterpret_cast<const XV*>(xc + (uint64_t)en.y * db);
#pragma unroll
            for (int w = 0; w < V; ++w) acc[w] += (double)comp<VT, V>(vv, w) * comp<double, V>(xv, w);
        }
    }
    ell_store<PM>(xs, db, pre);
    __syncthreads();

    const int nseg = nsa + 1;
    int buf = 0;
    for (int sg = 0; sg < nseg; ++sg) {
        const uint2 cur = e[nda + sg];
        const bool more = sg + 1 < nseg;
        if (more) ell_fetch<PM>(x + (uint64_t)e[nda + sg + 1].y * db, db, pre);
        if (act) {
            const bool self = sg == nsa;
            ell_segment<VT, V>(v + (uint64_t)cur.x * db, (self ? a.jfull : a.jsub) + ibc, db, self ? a.Eb : a.Sb,
                               xs + (uint64_t)buf * db, acc);
        }
        if (more) ell_store<PM>(xs + (uint64_t)(buf ^ 1) * db, db, pre);
        __syncthreads();
        buf ^= 1;
    }
    if (act) {
        if constexpr (V == 1) {
            y[ia * db + ib] = acc[0];
        } else {
            XV out;
#pragma unroll
            for (int w = 0; w < V; ++w) out[w] = acc[w];
            *reinterpret_cast<XV*>(y + ia * db + ib) = out;
        }
    }
}

// Measured on MI355X (FCI L = 11, 12): 1024 beta strings per workgroup (or
// one workgroup covering d_beta), two per thread for FP64 values and four for
// FP32 where d_beta allows (L = 12: 2.04 ms FP64, 76% of HBM peak, and
// 1.21 ms FP32, against 2.26 / 1.46 ms with one string per thread).
EllShape ell_default_shape(uint64_t db, int value_bytes) {
    EllShape sh;
    sh.vec = value_bytes == 4 && db % 4 == 0 ? 4 : db % 2 == 0 ? 2 : 1;
    const uint64_t cap = 1024 / sh.vec;
    uint64_t t = (db / sh.vec + 63) / 64 * 64;
    if (t > cap) t = cap;
    if (t < 64) t = 64;
    sh.threads = (int)t;
    return sh;
}

template <typename VT, int V>
static void launch_ell2(const EllArgs& a, int nt, uint64_t grid, uint32_t chunks, const VT* ell, const double* x,
                        double* y, hipStream_t s) {
    const uint64_t pm = (a.db + nt - 1) / nt;
    const size_t lds = 2 * a.db * sizeof(double);
#define CF_ELL2(P) k_spmv_ell2<VT, V, P><<<(unsigned)grid, nt, lds, s>>>(a, chunks, ell, x, y)
    if (pm <= 1) CF_ELL2(1);
    else if (pm <= 2) CF_ELL2(2);
    else if (pm <= 4) CF_ELL2(4);
    else if (pm <= 8) CF_ELL2(8);
    else if (pm <= 16) CF_ELL2(16);
    else throw Error(ERR_ARG, "ELL SpMV: workgroup too small for d_beta");
#undef CF_ELL2
}

template <typename VT>
static void launch_ell(const EllArgs& a, EllShape sh, const VT* ell, const double* x, double* y, hipStream_t s) {
    if (a.db > kEllMaxDb) throw Error(ERR_UNSUPPORTED, "ELL SpMV needs d_beta <= 8192");
    if (!a.da || !a.db) return;
    const int nt = sh.threads > 0 ? sh.threads : ell_default_shape(a.db, sizeof(VT)).threads;
    if (sh.vec == 0) {  // single-buffered kernel
        const uint64_t chunks = (a.db + nt - 1) / nt;
        const uint64_t grid = a.da * chunks;
        if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "ELL SpMV grid too large");
        const size_t lds = a.db * sizeof(double);
        switch (nt) {
            case 256: k_spmv_ell<VT, 256><<<(unsigned)grid, 256, lds, s>>>(a, (uint32_t)chunks, ell, x, y); break;
            case 512: k_spmv_ell<VT, 512><<<(unsigned)grid, 512, lds, s>>>(a, (uint32_t)chunks, ell, x, y); break;
            case 1024: k_spmv_ell<VT, 1024><<<(unsigned)grid, 1024, lds, s>>>(a, (uint32_t)chunks, ell, x, y); break;
            default: throw Error(ERR_ARG, "ELL SpMV (vec 0) workgroup size must be 256, 512 or 1024");
        }
        CF_HIP(hipGetLastError());
        return;
    }
    if (nt < 64 || nt > 1024 || nt % 64) throw Error(ERR_ARG, "ELL SpMV workgroup size must be a multiple of 64 <= 1024");
    if (sh.vec != 1 && sh.vec != 2 && sh.vec != 4) throw Error(ERR_ARG, "ELL SpMV vec must be 0, 1, 2 or 4");
    if (a.db % sh.vec) throw Error(ERR_ARG, "ELL SpMV: d_beta must be a multiple of vec");
    if (!a.ent || a.nda + a.nsa + 1 != a.Ea) throw Error(ERR_STATE, "ELL SpMV entries table missing");
    const uint64_t per = (uint64_t)nt * sh.vec;
    const uint64_t chunks = (a.db + per - 1) / per;
    const uint64_t grid = a.da * chunks;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "ELL SpMV grid too large");
    if (sh.vec == 1) launch_ell2<VT, 1>(a, nt, grid, (uint32_t)chunks, ell, x, y, s);
    else if (sh.vec == 2) launch_ell2<VT, 2>(a, nt, grid, (uint32_t)chunks, ell, x, y, s);
    else launch_ell2<VT, 4>(a, nt, grid, (uint32_t)chunks, ell, x, y, s);
    CF_HIP(hipGetLastError());
}

void launch_spmv_ell_f64(const EllArgs& a, EllShape sh, const double* ell, const double* x, double* y,
                         hipStream_t s) {
    launch_ell<double>(a, sh, ell, x, y, s);
}
void launch_spmv_ell_f32(const EllArgs& a, EllShape sh, const float* ell, const double* x, double* y,
                         hipStream_t s) {
    launch_ell<float>(a, sh, ell, x, y, s);
}

}  // namespace cf
