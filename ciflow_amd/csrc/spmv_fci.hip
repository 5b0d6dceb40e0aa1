// FCI-structured CRS SpMV with LDS-staged vector segments.
//
// Same CRS arrays and the same y = A x as k_spmv (SparseMatrix_CRS_Sym::mvprod,
// SparseMatrix_CRS.cpp:700-723), one wavefront per row, but the x gathers are
// served from LDS.  Why: an FCI row (ia, ib) of 1,819 entries (L=12) touches
// ~1,400 distinct x cache lines -- 36 alpha-single blocks x_{ja, *} hit at 37
// scattered beta positions each, and 225 alpha-double entries x_{ja, ib} in 225
// different blocks -- so a plain gather SpMV is bound by L2->L1 line traffic
// (~128 B moved per 8 B used), not by HBM.
//
// Rows of one alpha string ia share the column-block structure (the alpha
// list of ia; every complete row has the same length R and the same segment
// offsets).  A workgroup takes a tile of W consecutive rows (ia, ib0..ib0+W-1);
// its W waves walk their rows chunk by chunk (64 entries) in lockstep:
//   * each alpha self/single segment's x block x[ja*db .. ja*db+db) is staged
//     into an NS-slot LDS ring when its first chunk comes up and is shared by
//     the W rows (safe when (NS-1) * S_beta >= 64: the NS-1 segments between
//     two users of a slot span at least one chunk);
//   * the alpha-double values x[ja*db + ib] of the W rows are contiguous in
//     ib: all of them are staged once per tile (W values per double target);
//   * a per-alpha-string descriptor table (read through L1, identical for
//     the W rows) maps each row position t to the LDS offset of its x value.
// L2 line traffic per row falls from ~1,400 to ~(37*db/16 + ndouble)/W lines.
// The matrix stream itself is read exactly as by k_spmv (coalesced 64-entry
// chunks, non-temporal).  Tiles that contain a row with dropped entries
// (|v| <= zero, rows shorter than R) fall back to plain gathers inside the
// same kernel.
#include "spmv_fci.h"

namespace cf {

// ---------------------------------------------------------------------------
// precompute, one thread per alpha string
// ---------------------------------------------------------------------------
__global__ void k_fci_spmv_tables(uint64_t da, uint64_t db, int Ea, const Exc* __restrict__ full_a,
                                  const uint32_t* __restrict__ segoff, int R, int W, int NS, int nsd_max,
                                  int nd_max, int nchunk, int32_t* __restrict__ desc, uint32_t* __restrict__ sd_J,
                                  uint16_t* __restrict__ chunk_first, uint32_t* __restrict__ dbl_J) {
    const uint64_t ia = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ia >= da) return;
    const Exc* la = full_a + ia * (uint64_t)Ea;
    const uint32_t* so = segoff + ia * (uint64_t)(Ea + 1);
    int32_t* ds = desc + ia * (uint64_t)R;
    uint32_t* sj = sd_J + ia * (uint64_t)nsd_max;
    uint16_t* cf = chunk_first + ia * (uint64_t)(nchunk + 1);
    uint32_t* dj = dbl_J + ia * (uint64_t)nd_max;
    // LDS layout (doubles): [0, NS*db) ring of NS blocks, [NS*db, NS*db + nd*W) doubles
    const int32_t xd0 = NS * (int32_t)db;
    int o = 0, q = 0;
    int next_chunk = 0;
    for (int e = 0; e < Ea; ++e) {
        const int t = la[e].type;
        const uint32_t s = so[e], len = so[e + 1] - so[e];
        const uint32_t J = la[e].J;
        if (t == 2) {
            // double: x from the doubles area, + w added per row
            ds[s] = (xd0 + q * W) * 2 + 1;
            dj[q] = J;
            ++q;
        } else {
            // self / single: x[col] lives at slot*db + (col - J*db)
            const int32_t off = (int32_t)((o % NS) * (int32_t)db) - (int32_t)((int64_t)J * (int64_t)db);
            for (uint32_t k = 0; k < len; ++k) ds[s + k] = off * 2;
            const int c = (int)(s / 64);
            while (next_chunk <= c) cf[next_chunk++] = (uint16_t)o;
            sj[o] = J;
            ++o;
        }
    }
    while (next_chunk <= nchunk) cf[next_chunk++] = (uint16_t)o;
}

// ---------------------------------------------------------------------------
// tile kernel
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T ntl(const T* p) {
    return __builtin_nontemporal_load(p);
}

// One wavefront per tile of W rows (no barriers): the wave walks the alpha
// list of ia once; for every self/single segment it stages the x block of
// the target alpha string into its own LDS (coalesced loads, prefetched one
// segment ahead in registers) and applies it to the W rows' entries of that
// segment; for every double it reads the W rows' single entries, whose x
// values x[ja*db + ib0 .. ib0+W) are contiguous.  Lane-partial sums of the W
// rows live in registers; doubles add into a per-lane (= per-row) register.
constexpr int kMaxStage = 16;  // staging registers per lane: db <= 64*16

template <typename VT, int W>
__global__ void __launch_bounds__(64) k_spmv_fci_wave(FciSpmvArgs a, const VT* __restrict__ val,
                                                     const double* __restrict__ x, double* __restrict__ y) {
    extern __shared__ double xs[];  // db
    const uint32_t per = gridDim.x >> 3;
    const uint32_t tile = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (tile >= a.ntiles) return;
    const uint64_t db = a.db;
    const uint32_t tpa = a.tiles_per_alpha;
    const uint64_t ia = tile / tpa;
    const uint64_t ib0 = (uint64_t)(tile - ia * tpa) * W;
    const int nrows = (int)((db - ib0) < (uint64_t)W ? (db - ib0) : (uint64_t)W);
    const int lane = threadIdx.x;
    const uint64_t r0 = ia * db + ib0;
    const int R = a.R;
    const int64_t rp0 = a.rowptr[r0];
    const bool complete = (a.rowptr[r0 + nrows] - rp0) == (int64_t)R * nrows;
    if (!complete) {
        // rows with dropped entries: plain gathers, one row after the other
        for (int w = 0; w < nrows; ++w) {
            const int64_t s = a.rowptr[r0 + w], e = a.rowptr[r0 + w + 1];
            double acc = 0.0;
            for (int64_t k = s + lane; k < e; k += 64) acc += (double)ntl(val + k) * x[ntl(a.col + k)];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            if (lane == 0) y[r0 + w] = acc;
        }
        return;
    }
    const Exc* __restrict__ la = a.alist + ia * (uint64_t)a.Ea;
    const uint32_t* __restrict__ so = a.segoff + ia * (uint64_t)(a.Ea + 1);
    const uint32_t* __restrict__ sdJ = a.sd_J + ia * (uint64_t)a.nsd_max;
    const int nst = (int)((db + 63) / 64);
    int64_t base[W];
#pragma unroll
    for (int w = 0; w < W; ++w) base[w] = rp0 + (int64_t)R * (w < nrows ? w : 0);
    double acc[W];
#pragma unroll
    for (int w = 0; w < W; ++w) acc[w] = 0.0;
    double accd = 0.0;  // doubles of row `lane`
    double st[kMaxStage];
    int o = 0;  // self/single ordinal
    auto prefetch = [&](int ord) {
        const double* src = x + (uint64_t)sdJ[ord] * db;
#pragma unroll
        for (int i = 0; i < kMaxStage; ++i) {
            const uint64_t q = (uint64_t)i * 64 + lane;
            if (i < nst && q < db) st[i] = src[q];
        }
    };
    prefetch(0);
    const int Ea = a.Ea, nsd = a.nsd;
    for (int e = 0; e < Ea; ++e) {
        const Exc ex = la[e];
        const uint32_t off = so[e];
        if (ex.type == 2) {
            if (lane < nrows) {
                const int64_t k = rp0 + (int64_t)R * lane + off;
                accd += (double)ntl(val + k) * x[ntl(a.col + k)];
            }
            continue;
        }
        // self or single: stage the block, then apply it to the W rows
#pragma unroll
        for (int i = 0; i < kMaxStage; ++i) {
            const uint64_t q = (uint64_t)i * 64 + lane;
            if (i < nst && q < db) xs[q] = st[i];
        }
        if (o + 1 < nsd) prefetch(o + 1);
        ++o;
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): own LDS writes done
        __builtin_amdgcn_wave_barrier();
        const uint32_t len = so[e + 1] - off;
        const int64_t cbase = (int64_t)ex.J * (int64_t)db;
        for (uint32_t t0 = 0; t0 < len; t0 += 64) {
            const uint32_t t = t0 + lane;
            const bool ok = t < len;
            uint32_t cc[W];
            VT vv[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const bool okw = ok && w < nrows;
                cc[w] = okw ? ntl(a.col + base[w] + off + t) : (uint32_t)cbase;
                vv[w] = okw ? ntl(val + base[w] + off + t) : (VT)0;
            }
#pragma unroll
            for (int w = 0; w < W; ++w) acc[w] += (double)vv[w] * xs[(int64_t)cc[w] - cbase];
        }
        __builtin_amdgcn_wave_barrier();
    }
    // row w total = sum over lanes of acc[w] + accd of lane w
#pragma unroll
    for (int w = 0; w < W; ++w) {
        double v = acc[w];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[w] = v;
    }
    if (lane < nrows) {
        double tot = accd;
#pragma unroll
        for (int w = 0; w < W; ++w)
            if (w == lane) tot += acc[w];
        y[r0 + lane] = tot;
    }
}

size_t fci_spmv_lds_bytes(const FciSpmvArgs& a, int W) {
    (void)W;
    return a.db * sizeof(double);
}

template <typename VT>
static void launch_t(const FciSpmvArgs& a, const VT* val, const double* x, double* y, hipStream_t stream) {
    if (a.db > 64 * kMaxStage) throw Error(ERR_UNSUPPORTED, "beta space too large for register staging");
    const uint64_t grid = (a.ntiles + 7) / 8 * 8;
    const size_t lds = fci_spmv_lds_bytes(a, a.W);
    switch (a.W) {
        case 8: k_spmv_fci_wave<VT, 8><<<(unsigned)grid, 64, lds, stream>>>(a, val, x, y); break;
        case 16: k_spmv_fci_wave<VT, 16><<<(unsigned)grid, 64, lds, stream>>>(a, val, x, y); break;
        case 32: k_spmv_fci_wave<VT, 32><<<(unsigned)grid, 64, lds, stream>>>(a, val, x, y); break;
        default: throw Error(ERR_ARG, "FCI SpMV tile height must be 8, 16 or 32");
    }
    CF_HIP(hipGetLastError());
}

void launch_spmv_fci_f64(const FciSpmvArgs& a, const double* val, const double* x, double* y, hipStream_t s) {
    launch_t<double>(a, val, x, y, s);
}
void launch_spmv_fci_f32(const FciSpmvArgs& a, const float* val, const double* x, double* y, hipStream_t s) {
    launch_t<float>(a, val, x, y, s);
}

void FciSpmvPlan::init(uint64_t da_, uint64_t db_, int Ea, const Exc* full_a, const uint32_t* segoff, int R_,
                       int nd_, int Sb, int W_, hipStream_t stream) {
    if (3 * Sb >= 64) a.NS = 4;
    else if (7 * Sb >= 64) a.NS = 8;
    else throw Error(ERR_UNSUPPORTED, "beta single lists too short for the staged FCI SpMV");
    a.db = db_;
    a.R = R_;
    a.Ea = Ea;
    a.alist = full_a;
    a.segoff = segoff;
    a.nsd = Ea - nd_;
    a.W = W_;
    a.nd = nd_;
    a.nd_max = nd_ > 0 ? nd_ : 1;
    a.nsd_max = Ea - nd_ > 0 ? Ea - nd_ : 1;
    a.nchunk = (R_ + 63) / 64;
    a.tiles_per_alpha = (uint32_t)((db_ + W_ - 1) / W_);
    const uint64_t nt = da_ * a.tiles_per_alpha;
    if (nt >= (1ull << 32)) throw Error(ERR_UNSUPPORTED, "too many SpMV tiles");
    a.ntiles = (uint32_t)nt;
    desc.alloc(da_ * (uint64_t)R_);
    sdJ.alloc(da_ * (uint64_t)a.nsd_max);
    cfirst.alloc(da_ * (uint64_t)(a.nchunk + 1));
    dJ.alloc(da_ * (uint64_t)a.nd_max);
    k_fci_spmv_tables<<<(unsigned)((da_ + 63) / 64), 64, 0, stream>>>(da_, db_, Ea, full_a, segoff, R_, W_, a.NS,
                                                                       a.nsd_max, a.nd_max, a.nchunk, desc.p, sdJ.p,
                                                                       cfirst.p, dJ.p);
    CF_HIP(hipGetLastError());
    a.desc = desc.p;
    a.sd_J = sdJ.p;
    a.chunk_first = cfirst.p;
    a.dbl_J = dJ.p;
}

}  // namespace cf
