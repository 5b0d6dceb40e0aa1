// On-the-fly FCI sigma = H c without a stored matrix (replaces the O(dim^2)
// on-the-fly product of CI_Big::mvprod, CIMethod.cpp:943-984, which is
// infeasible at 1.66e8 determinants).
//
// With C viewed as the d_a x d_b matrix C[Ia][Ib] (alpha rank slow, beta rank
// fast, CIMethod.cpp:771), H splits exactly into
//   sigma = Ha C + C Hb^T + sigma_ab
// Ha / Hb: the same-spin string Hamiltonians (one-string diagonal, single
//   excitations h_pq + sum_z (V_pzqz - V_pzzq) with their phase, same-spin
//   doubles) -- the entries of the sorted excitation lists, scattered into a
//   dense d x d matrix once per Hamiltonian.  The two products are FP64 GEMMs
//   (rocBLAS on MFMA) for either operand precision: 2 d^3 flop each.
// sigma_ab(Ia,Ib) = sum_{a in A(Ia)} sum_{b in A(Ib)} s_a s_b V[p_a][q_b][r_a][s_b] C(J_a, J_b),
//   A(I) = the single excitations of I (p in I -> r, J = I - p + r, phase
//   s) plus the number operators (p = r in I, J = I, s = +1): this is
//   sum_pqrs V_pqrs E^a_pr E^b_qs and carries the alpha-beta Coulomb parts of
//   the reference's diagonal (jmix) and single elements
//   (CIMethod.cpp:343-351, 409-427) and its ab doubles (:444-457).
//   Kernel: a workgroup owns one alpha string Ia and 1024 beta strings (one
//   per thread, whose singles sit in registers); for each alpha action it
//   stages the C row J_a and the V slice V[p_a][.][r_a][.] in LDS and every
//   thread gathers its |A(Ib)| terms from LDS.
// Algorithmic HBM stream per sigma (L=16 8a8b): the C rows staged per alpha
// action, d_a * |A_a| * d_b * 8 B = 12,870 * 65 * 12,870 * 8 B = 86 GB (from
// L2 when the 13 beta chunks of one Ia run together), plus the GEMM operands.
#include "sigma.h"

namespace cf {

// Dense same-spin string Hamiltonian from the excitation lists.
template <typename T>
__global__ void k_dense_string_ham(uint64_t d, int E, const Exc* __restrict__ full, T* __restrict__ H) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d * (uint64_t)E) return;
    const uint64_t I = t / E;
    const Exc x = full[t];
    double v = x.x;
    if (x.type == 1 && x.sign < 0) v = -v;
    H[I * d + x.J] = (T)v;
}

template <typename T>
void launch_dense_string_ham(uint64_t d, int E, const Exc* full, T* H, hipStream_t s) {
    CF_HIP(hipMemsetAsync(H, 0, d * d * sizeof(T), s));
    const uint64_t tot = d * (uint64_t)E;
    k_dense_string_ham<T><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(d, E, full, H);
    CF_HIP(hipGetLastError());
}
template void launch_dense_string_ham<double>(uint64_t, int, const Exc*, double*, hipStream_t);

constexpr int kMaxSingles = 64;  // n (L - n) <= 64; L <= 31 (q = L marks the zero slot)

// Staged row elements per thread held in registers while the previous action
// computes (db <= NT * kMaxPrefetch uses a prefetching kernel).
constexpr int kMaxPrefetch = 16;

// V slice of one alpha action, staged in LDS: self -> sum over p in Ia of
// V[p][q][p][s]; single p -> r -> s_a V[p][q][r][s].
template <int NT>
__device__ __forceinline__ void stage_vslice(const SigmaArgs& a, const Exc& x, str_t Sa, double* Vs) {
    const int L = a.L;
    if (threadIdx.x == 0) Vs[L * L] = 0.0;
    for (int i = threadIdx.x; i < L * L; i += NT) {
        const int q = i / L, s = i % L;
        double v = 0.0;
        if (x.type == 0) {
            for (str_t t = Sa; t; t &= t - 1) {
                const int p = __builtin_ctzll(t);
                v += a.V[(((size_t)p * L + q) * L + p) * L + s];
            }
        } else {
            v = a.V[(((size_t)x.f * L + q) * L + x.t) * L + s];
            if (x.sign < 0) v = -v;
        }
        Vs[i] = v;
    }
}

// sigma_ab += ..., CT = double or float (staged C precision); accumulation FP64.
// The next action's C row is loaded into registers (PF elements per thread)
// while the current one is processed from LDS, so the row fetch overlaps the
// gathers; PF = 0 stages each row after the previous action has finished.
template <typename CT, int PF, int NT>
__global__ void __launch_bounds__(NT) k_sigma_ab(SigmaArgs a, const CT* __restrict__ C, double* __restrict__ sig) {
    extern __shared__ unsigned char smem[];
    const uint64_t db = a.db;
    const uint32_t nchunk = a.nchunk;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (lb >= a.nblocks) return;
    const uint64_t Ia = lb / nchunk;
    const uint64_t Ib = (uint64_t)(lb - Ia * nchunk) * NT + threadIdx.x;
    const bool act = Ib < db;
    const int L = a.L;
    double* Vs = (double*)smem;             // L*L, then a zero at [L*L] (and one pad slot)
    CT* row = (CT*)(Vs + L * L + 2);        // db
    // this thread's beta actions, packed (J << 11 | q << 6 | s << 1 | sign) in
    // list order; the self entry points at the zero slot (q = L, s = 0) and the
    // number operators are summed separately from the bits of Ib.  Every index
    // is a compile-time constant, so the list stays in registers.
    uint32_t sb[kMaxSingles + 1];
    const int Sb = a.Sb;
    str_t sbits = 0;
    if (act) {
        sbits = a.sb[Ib];
        const Exc* lb_ = a.subb + Ib * (uint64_t)Sb;
#pragma unroll
        for (int e = 0; e < kMaxSingles + 1; ++e) {
            if (e < Sb) {
                const Exc x = lb_[e];
                sb[e] = x.type == 1
                            ? (x.J << 11) | ((uint32_t)x.f << 6) | ((uint32_t)x.t << 1) | (x.sign < 0 ? 1u : 0u)
                            : ((uint32_t)Ib << 11) | ((uint32_t)L << 6);
            }
        }
    }
    const Exc* la = a.suba + Ia * (uint64_t)a.Sa;
    const str_t Sa = a.sa[Ia];
    CT pf[PF > 0 ? PF : 1];
    if (PF > 0) {
        const CT* src = C + (uint64_t)la[0].J * db;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint64_t i = (uint64_t)k * NT + threadIdx.x;
            pf[k] = i < db ? src[i] : CT(0);
        }
    }
    double acc = 0.0;
    for (int e = 0; e < a.Sa; ++e) {
        const Exc x = la[e];
        __syncthreads();
        stage_vslice<NT>(a, x, Sa, Vs);
        if (PF > 0) {
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const uint64_t i = (uint64_t)k * NT + threadIdx.x;
                if (i < db) row[i] = pf[k];
            }
        } else {
            const CT* src = C + (uint64_t)x.J * db;
            for (uint64_t i = threadIdx.x; i < db; i += NT) row[i] = src[i];
        }
        __syncthreads();
        if (PF > 0 && e + 1 < a.Sa) {
            const CT* src = C + (uint64_t)la[e + 1].J * db;
#pragma unroll
            for (int k = 0; k < PF; ++k) {
                const uint64_t i = (uint64_t)k * NT + threadIdx.x;
                pf[k] = i < db ? src[i] : CT(0);
            }
        }
        if (act) {
            double t = 0.0;
            // beta number operators: J_b = Ib
            double dsum = 0.0;
            for (str_t u = sbits; u; u &= u - 1) {
                const int q = __builtin_ctzll(u);
                dsum += Vs[q * L + q];
            }
            t += dsum * (double)row[Ib];
#pragma unroll
            for (int j = 0; j < kMaxSingles + 1; ++j) {
                if (j < Sb) {
                    const uint32_t c = sb[j];
                    const int q = (c >> 6) & 31, s = (c >> 1) & 31;
                    double v = Vs[q * L + s] * (double)row[c >> 11];
                    t += (c & 1) ? -v : v;
                }
            }
            acc += t;
        }
    }
    if (act) sig[Ia * db + Ib] += acc;
}

size_t sigma_ab_lds_bytes(const SigmaArgs& a, bool f32) {
    return ((size_t)a.L * a.L + 2) * sizeof(double) + a.db * (f32 ? sizeof(float) : sizeof(double));
}

template <typename CT, int NT>
static void launch_nt(SigmaArgs a, const CT* C, double* sig, hipStream_t s) {
    a.nchunk = (uint32_t)((a.db + NT - 1) / NT);
    const uint64_t nb = a.da * (uint64_t)a.nchunk;
    if (nb > 0xFFFFFFF0ull) throw Error(ERR_UNSUPPORTED, "too many sigma workgroups");
    a.nblocks = (uint32_t)nb;
    const uint64_t grid = (nb + 7) / 8 * 8;
    const size_t lds = sigma_ab_lds_bytes(a, sizeof(CT) == 4);
    const uint64_t per_thread = (a.db + NT - 1) / NT;
    if (per_thread <= 4)
        k_sigma_ab<CT, 4, NT><<<(unsigned)grid, NT, lds, s>>>(a, C, sig);
    else if (per_thread <= kMaxPrefetch)
        k_sigma_ab<CT, kMaxPrefetch, NT><<<(unsigned)grid, NT, lds, s>>>(a, C, sig);
    else if (sizeof(CT) == 4 && per_thread <= 17)  // FP32 rows, L = 16 at 768 threads (17 per thread)
        k_sigma_ab<CT, 17, NT><<<(unsigned)grid, NT, lds, s>>>(a, C, sig);
    else
        k_sigma_ab<CT, 0, NT><<<(unsigned)grid, NT, lds, s>>>(a, C, sig);
    CF_HIP(hipGetLastError());
}

template <typename CT>
void launch_sigma_ab(const SigmaArgs& a, int threads, const CT* C, double* sig, hipStream_t s) {
    switch (threads) {
        case 128: launch_nt<CT, 128>(a, C, sig, s); break;
        case 256: launch_nt<CT, 256>(a, C, sig, s); break;
        case 512: launch_nt<CT, 512>(a, C, sig, s); break;
        case 768: launch_nt<CT, 768>(a, C, sig, s); break;
        case 1024: launch_nt<CT, 1024>(a, C, sig, s); break;
        default: throw Error(ERR_ARG, "sigma workgroup size must be 128, 256, 512, 768 or 1024");
    }
}
template void launch_sigma_ab<double>(const SigmaArgs&, int, const double*, double*, hipStream_t);
template void launch_sigma_ab<float>(const SigmaArgs&, int, const float*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// Kernel 2.  The V slice of an alpha action is a contiguous L^2 block of a
// table built once per Hamiltonian (vpair for a single p -> r, vself for the
// alpha number operators of Ia), so the next action's slice and C row are
// loaded into registers while the current action is processed from LDS; with
// a double-buffered LDS one barrier per action remains and no thread waits on
// a global load at a barrier.  Slices are held in the staged precision CT.
// ---------------------------------------------------------------------------
template <typename CT>
__global__ void k_sigma_vpair(int L, const double* __restrict__ V, CT* __restrict__ vpair) {
    const uint64_t L2 = (uint64_t)L * L, t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= L2 * L2) return;
    const uint64_t pr = t / L2, qs = t - pr * L2;
    const int p = (int)(pr / L), r = (int)(pr % L), q = (int)(qs / L), s = (int)(qs % L);
    vpair[t] = (CT)V[(((uint64_t)p * L + q) * L + r) * L + s];
}

template <typename CT>
__global__ void k_sigma_vself(uint64_t da, int L, const str_t* __restrict__ sa, const double* __restrict__ V,
                              CT* __restrict__ vself) {
    const uint64_t L2 = (uint64_t)L * L, t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= da * L2) return;
    const uint64_t Ia = t / L2, qs = t - Ia * L2;
    const int q = (int)(qs / L), s = (int)(qs % L);
    double v = 0.0;
    for (str_t u = sa[Ia]; u; u &= u - 1) {
        const int p = __builtin_ctzll(u);
        v += V[(((uint64_t)p * L + q) * L + p) * L + s];
    }
    vself[t] = (CT)v;
}

template <typename CT>
void launch_sigma_tables(const SigmaArgs& a, CT* vpair, CT* vself, hipStream_t s) {
    const uint64_t L2 = (uint64_t)a.L * a.L;
    k_sigma_vpair<CT><<<(unsigned)((L2 * L2 + 255) / 256), 256, 0, s>>>(a.L, a.V, vpair);
    CF_HIP(hipGetLastError());
    const uint64_t tot = a.da * L2;
    if (tot) {
        k_sigma_vself<CT><<<(unsigned)((tot + 255) / 256), 256, 0, s>>>(a.da, a.L, a.sa, a.V, vself);
        CF_HIP(hipGetLastError());
    }
}
template void launch_sigma_tables<double>(const SigmaArgs&, double*, double*, hipStream_t);
template void launch_sigma_tables<float>(const SigmaArgs&, float*, float*, hipStream_t);

template <typename CT>
__device__ __forceinline__ const CT* vpair_tab(const SigmaArgs& a);
template <>
__device__ __forceinline__ const double* vpair_tab<double>(const SigmaArgs& a) { return a.vpair64; }
template <>
__device__ __forceinline__ const float* vpair_tab<float>(const SigmaArgs& a) { return a.vpair32; }
template <typename CT>
__device__ __forceinline__ const CT* vself_tab(const SigmaArgs& a);
template <>
__device__ __forceinline__ const double* vself_tab<double>(const SigmaArgs& a) { return a.vself64; }
template <>
__device__ __forceinline__ const float* vself_tab<float>(const SigmaArgs& a) { return a.vself32; }

// PF * NT >= d_beta, PV * NT >= L^2 and NS >= the beta list length (checked by
// the launcher); NS sizes the register copy of the thread's beta list.
template <typename CT, int PF, int PV, int NT, bool DB, int NS>
__global__ void __launch_bounds__(NT) k_sigma_ab2(SigmaArgs a, const CT* __restrict__ C, double* __restrict__ sig) {
    extern __shared__ unsigned char smem[];
    const uint64_t db = a.db;
    const uint32_t nchunk = a.nchunk;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (lb >= a.nblocks) return;
    const uint64_t Ia = lb / nchunk;
    const uint64_t Ib = (uint64_t)(lb - Ia * nchunk) * NT + threadIdx.x;
    const bool act = Ib < db;
    const int L = a.L, L2 = L * L;
    // buffer b: V slice at base + b*bsz (L^2 entries, a zero at [L^2]), C row after it
    const uint64_t bsz = (uint64_t)L2 + 2 + db;
    CT* base = (CT*)smem;
    if (threadIdx.x == 0) {
        base[L2] = CT(0);
        if (DB) base[bsz + L2] = CT(0);
    }
    uint32_t sb[NS];
    const int Sb = a.Sb;
    str_t sbits = 0;
    if (act) {
        sbits = a.sb[Ib];
        const Exc* lb_ = a.subb + Ib * (uint64_t)Sb;
#pragma unroll
        for (int e = 0; e < NS; ++e) {
            if (e < Sb) {
                const Exc x = lb_[e];
                sb[e] = x.type == 1
                            ? (x.J << 11) | ((uint32_t)x.f << 6) | ((uint32_t)x.t << 1) | (x.sign < 0 ? 1u : 0u)
                            : ((uint32_t)Ib << 11) | ((uint32_t)L << 6);
            }
        }
    }
    const Exc* la = a.suba + Ia * (uint64_t)a.Sa;
    const CT* vpair = vpair_tab<CT>(a);
    const CT* vself = vself_tab<CT>(a) + Ia * (uint64_t)L2;
    CT pv[PV], pf[PF];
    bool neg_next = false;
    auto fetch = [&](int e) {
        const Exc x = la[e];
        neg_next = x.type == 1 && x.sign < 0;
        const CT* vsrc = x.type == 0 ? vself : vpair + (uint64_t)(x.f * L + x.t) * L2;
#pragma unroll
        for (int k = 0; k < PV; ++k) {
            const int i = k * NT + threadIdx.x;
            pv[k] = i < L2 ? vsrc[i] : CT(0);
        }
        const CT* src = C + (uint64_t)x.J * db;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint64_t i = (uint64_t)k * NT + threadIdx.x;
            pf[k] = i < db ? src[i] : CT(0);
        }
    };
    auto store = [&](int b) {
        CT* Vs = base + b * bsz;
        CT* row = Vs + L2 + 2;
#pragma unroll
        for (int k = 0; k < PV; ++k) {
            const int i = k * NT + threadIdx.x;
            if (i < L2) Vs[i] = pv[k];
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint64_t i = (uint64_t)k * NT + threadIdx.x;
            if (i < db) row[i] = pf[k];
        }
    };
    fetch(0);
    store(0);
    bool neg = neg_next;
    __syncthreads();
    double acc = 0.0;
    const int Sa = a.Sa;
    for (int e = 0; e < Sa; ++e) {
        const int cur = DB ? (e & 1) : 0;
        const bool more = e + 1 < Sa;
        if (more) fetch(e + 1);
        if (act) {
            const CT* Vs = base + cur * bsz;
            const CT* row = Vs + L2 + 2;
            double dsum = 0.0;  // beta number operators: J_b = Ib
            for (str_t u = sbits; u; u &= u - 1) {
                const int q = __builtin_ctzll(u);
                dsum += (double)Vs[q * L + q];
            }
            double t = dsum * (double)row[Ib];
#pragma unroll
            for (int j = 0; j < NS; ++j) {
                if (j < Sb) {
                    const uint32_t c = sb[j];
                    const int q = (c >> 6) & 31, s = (c >> 1) & 31;
                    const double v = (double)Vs[q * L + s] * (double)row[c >> 11];
                    t += (c & 1) ? -v : v;
                }
            }
            acc += neg ? -t : t;
        }
        if (more) {
            if (!DB) __syncthreads();  // the single buffer is still being read
            store(DB ? cur ^ 1 : 0);
            neg = neg_next;
        }
        __syncthreads();
    }
    if (act) sig[Ia * db + Ib] += acc;
}

size_t sigma_ab2_lds_bytes(const SigmaArgs& a, bool f32, bool dbuf) {
    return (dbuf ? 2 : 1) * ((size_t)a.L * a.L + 2 + a.db) * (f32 ? sizeof(float) : sizeof(double));
}

constexpr size_t kLdsMax = 160 * 1024;

// Supported (NT, PF) pairs; the double buffer is used for PF <= 8 (it fits the
// LDS for every d_beta <= 8 * NT with L <= 31), a single buffer for PF = 17.
template <typename CT, int NT, int PF, int PV, int NS>
static void launch_ab2_t(const SigmaArgs& a, uint64_t grid, const CT* C, double* sig, hipStream_t s) {
    constexpr bool DB = PF <= 8;
    const size_t lds = sigma_ab2_lds_bytes(a, sizeof(CT) == 4, DB);
    if (lds > kLdsMax) throw Error(ERR_UNSUPPORTED, "sigma: beta space too large for the LDS buffers");
    k_sigma_ab2<CT, PF, PV, NT, DB, NS><<<(unsigned)grid, NT, lds, s>>>(a, C, sig);
}

template <typename CT, int NT, int PF>
static bool launch_ab2_pv(const SigmaArgs& a, uint64_t grid, const CT* C, double* sig, hipStream_t s) {
    const int L2 = a.L * a.L;
    const int ns = a.Sb;
#define CF_AB2_NS(PV)                                                         \
    do {                                                                      \
        if (ns <= 17) launch_ab2_t<CT, NT, PF, PV, 17>(a, grid, C, sig, s);   \
        else if (ns <= 33) launch_ab2_t<CT, NT, PF, PV, 33>(a, grid, C, sig, s); \
        else launch_ab2_t<CT, NT, PF, PV, 65>(a, grid, C, sig, s);            \
    } while (0)
    if (L2 <= NT) CF_AB2_NS(1);
    else if (L2 <= 2 * NT) CF_AB2_NS(2);
    else return false;
#undef CF_AB2_NS
    return true;
}

template <typename CT>
void launch_sigma_ab2(const SigmaArgs& a0, SigmaShape sh, const CT* C, double* sig, hipStream_t s) {
    bool ok = false;
    const bool tables = sizeof(CT) == 4 ? a0.vpair32 != nullptr : a0.vpair64 != nullptr;
    // FP64 rows of more than 8 elements per thread spill the prefetch registers
    // (L = 16: 920 ms against 771 ms for kernel 1, measured on MI355X)
    const bool fits = sizeof(CT) == 4 || (a0.db + sh.threads - 1) / sh.threads <= 8;
    if (sh.kernel == 2 && tables && fits && a0.Sb <= kMaxSingles + 1) {
        SigmaArgs a = a0;
        const int NTr = sh.threads;
        a.nchunk = (uint32_t)((a.db + NTr - 1) / NTr);
        const uint64_t nb = a.da * (uint64_t)a.nchunk;
        if (nb > 0xFFFFFFF0ull) throw Error(ERR_UNSUPPORTED, "too many sigma workgroups");
        a.nblocks = (uint32_t)nb;
        const uint64_t grid = (nb + 7) / 8 * 8;
        const uint64_t pt = (a.db + NTr - 1) / NTr;  // row elements per thread
        switch (NTr) {
            case 128:
                if (pt <= 1) ok = launch_ab2_pv<CT, 128, 1>(a, grid, C, sig, s);
                else if (pt <= 2) ok = launch_ab2_pv<CT, 128, 2>(a, grid, C, sig, s);
                else if (pt <= 4) ok = launch_ab2_pv<CT, 128, 4>(a, grid, C, sig, s);
                break;
            case 256:
                if (pt <= 1) ok = launch_ab2_pv<CT, 256, 1>(a, grid, C, sig, s);
                else if (pt <= 2) ok = launch_ab2_pv<CT, 256, 2>(a, grid, C, sig, s);
                else if (pt <= 4) ok = launch_ab2_pv<CT, 256, 4>(a, grid, C, sig, s);
                break;
            case 512:
                if (pt <= 1) ok = launch_ab2_pv<CT, 512, 1>(a, grid, C, sig, s);
                else if (pt <= 2) ok = launch_ab2_pv<CT, 512, 2>(a, grid, C, sig, s);
                else if (pt <= 8) ok = launch_ab2_pv<CT, 512, 8>(a, grid, C, sig, s);
                break;
            case 768:
                if (pt <= 8) ok = launch_ab2_pv<CT, 768, 8>(a, grid, C, sig, s);
                else if (pt <= 17) ok = launch_ab2_pv<CT, 768, 17>(a, grid, C, sig, s);
                break;
            case 1024:
                if (pt <= 2) ok = launch_ab2_pv<CT, 1024, 2>(a, grid, C, sig, s);
                else if (pt <= 8) ok = launch_ab2_pv<CT, 1024, 8>(a, grid, C, sig, s);
                else if (pt <= 17) ok = launch_ab2_pv<CT, 1024, 17>(a, grid, C, sig, s);
                break;
            default: throw Error(ERR_ARG, "sigma workgroup size must be 128, 256, 512, 768 or 1024");
        }
        if (ok) CF_HIP(hipGetLastError());
    }
    if (!ok) launch_sigma_ab<CT>(a0, sh.threads, C, sig, s);
}
template void launch_sigma_ab2<double>(const SigmaArgs&, SigmaShape, const double*, double*, hipStream_t);
template void launch_sigma_ab2<float>(const SigmaArgs&, SigmaShape, const float*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// Kernel 3.  The same alpha-action loop as kernel 2 with the per-term work cut
// to two LDS reads and three VALU operations.  Measured on kernel 2 (FCI
// L=16, rocprofv3 counters): 12 VALU and 3 SALU wave-instructions per term
// against 2.4 LDS instructions, with the LDS pipe 16% busy -- the kernel was
// bound by instruction issue, not by LDS bandwidth or bank conflicts.  Here:
//  * each beta action of a thread is one register holding two LDS byte
//    offsets, (offset of C(J_b) in the staged row) << 13 | (offset of the V
//    slice entry); the beta sign picks a negated copy of the slice staged
//    next to it, so no sign select is left;
//  * the beta number operators (q in Ib, J = Ib) are ordinary entries of the
//    list, whose exact length NS is a template parameter (no list-length
//    tests, fully unrolled);
//  * FP32 operands: one FP32 product, converted and added in FP64; FP64
//    operands: one FP64 FMA.
// Staged buffer (bytes, element size s): [0, L^2 s) V slice, [L^2 s, 2 L^2 s)
// its negation, [2 L^2 s, 2 L^2 s + d_b s) the C row.  BUFB > 0: two buffers
// BUFB bytes apart (BUFB a compile-time constant, so buffer 1's reads carry
// it as the instruction's immediate offset); BUFB = 0: one buffer and two
// barriers per action.  PIN: the codes are re-declared inside the term loop
// (an empty asm that "modifies" each one), so the compiler unpacks them per
// use (two more VALU per term) instead of hoisting 2 NS unpacked offsets
// into registers -- at NS = 72 that hoisting took 214 registers and left 2
// waves per SIMD.
// ---------------------------------------------------------------------------
template <typename CT, int NT, int PF, int PV, int NS, int BUFB, int LDSB, bool BATCH, bool PIN>
__global__ void __launch_bounds__(NT) k_sigma_ab3(SigmaArgs a0, const CT* __restrict__ C0, double* __restrict__ sig0,
                                                  SigmaBatch sb) {
    // static LDS (the kernel's only LDS): its base is the constant 0, so the
    // packed offsets address it directly
    __shared__ __align__(16) unsigned char smem3[LDSB];
    constexpr uint32_t S = sizeof(CT);
    SigmaArgs a = a0;
    const CT* __restrict__ C = C0;
    double* __restrict__ sig = sig0;
    if (BATCH) {  // lane blockIdx.y of a batched Davidson product
        const int lane = blockIdx.y;
        const BdStep& st = sb.step[lane];
        if (st.mode == BD_IDLE || st.f32 != (S == 4 ? 1 : 0)) return;
        a = sb.args[lane];
        const uint64_t off = ((uint64_t)lane * sb.M + st.m) * sb.n;
        C = S == 4 ? reinterpret_cast<const CT*>(sb.x32 + (uint64_t)lane * sb.n) : reinterpret_cast<const CT*>(sb.V + off);
        sig = sb.S + off;
    }
    const uint64_t db = a.db;
    const uint32_t nchunk = a.nchunk;
    const uint32_t per = gridDim.x >> 3;
    const uint32_t lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    if (lb >= a.nblocks) return;
    const uint64_t Ia = lb / nchunk;
    const uint64_t Ib = (uint64_t)(lb - Ia * nchunk) * NT + threadIdx.x;
    const bool act = Ib < db;
    const int L = a.L, L2 = L * L;
    const uint32_t vneg = (uint32_t)L2 * S, crow = 2u * (uint32_t)L2 * S;
    // this thread's beta actions: the singles of its self + singles list
    // (sorted by target, the self entry J = Ib at position `self`), then its
    // number operators; every register index is a compile-time constant
    uint32_t code[NS];
    {
        const int ns1 = a.Sb - 1;
        const Exc* lb_ = a.subb + (act ? Ib : 0) * (uint64_t)a.Sb;
        int lo = 0, hi = a.Sb - 1;  // position of the self entry (J == Ib)
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (lb_[mid].J < (uint32_t)Ib) lo = mid + 1;
            else hi = mid;
        }
        const int self = lo;
        str_t u = act ? a.sb[Ib] : 0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (j < ns1) {
                const Exc x = lb_[j + (j >= self ? 1 : 0)];
                code[j] = ((crow + x.J * S) << 13) | ((uint32_t)(x.f * L + x.t) * S + (x.sign < 0 ? vneg : 0u));
            } else {  // number operator q in Ib: J = Ib
                const int q = u ? __builtin_ctzll(u) : 0;
                u &= u - 1;
                code[j] = ((crow + (uint32_t)(act ? Ib : 0) * S) << 13) | ((uint32_t)(q * L + q) * S);
            }
        }
    }
    const Exc* la = a.suba + Ia * (uint64_t)a.Sa;
    const CT* vpair = vpair_tab<CT>(a);
    const CT* vself = vself_tab<CT>(a) + Ia * (uint64_t)L2;
    CT pv[PV], pf[PF];
    bool neg_next = false;
    auto fetch = [&](int e) {
        const Exc x = la[e];
        neg_next = x.type == 1 && x.sign < 0;
        const CT* vsrc = x.type == 0 ? vself : vpair + (uint64_t)(x.f * L + x.t) * L2;
#pragma unroll
        for (int k = 0; k < PV; ++k) {
            const int i = k * NT + threadIdx.x;
            pv[k] = i < 2 * L2 ? vsrc[i < L2 ? i : i - L2] : CT(0);
        }
        const CT* src = C + (uint64_t)x.J * db;
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint64_t i = (uint64_t)k * NT + threadIdx.x;
            pf[k] = i < db ? src[i] : CT(0);
        }
    };
    auto store = [&](unsigned char* buf) {
        CT* V = (CT*)buf;
        CT* row = (CT*)(buf + crow);
#pragma unroll
        for (int k = 0; k < PV; ++k) {
            const int i = k * NT + threadIdx.x;
            if (i < 2 * L2) V[i] = i < L2 ? pv[k] : -pv[k];
        }
#pragma unroll
        for (int k = 0; k < PF; ++k) {
            const uint64_t i = (uint64_t)k * NT + threadIdx.x;
            if (i < db) row[i] = pf[k];
        }
    };
    auto terms = [&](const unsigned char* buf) -> double {
        double t = 0.0;
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            // PIN: the code is redefined here, so its two unpacked offsets are
            // formed per use instead of being hoisted out of the action loop
            // into 2 NS more registers
            if (PIN) asm volatile("" : "+v"(code[j]));
            const uint32_t c = code[j];
            const CT v = *(const CT*)(buf + (c & 0x1FFFu));
            const CT x = *(const CT*)(buf + (c >> 13));
            if (sizeof(CT) == 4) t += (double)(v * x);
            else t = fma((double)v, (double)x, t);
        }
        return t;
    };
    fetch(0);
    store(smem3);
    bool neg = neg_next;
    __syncthreads();
    double acc = 0.0;
    const int Sa = a.Sa;
    if (BUFB > 0) {
        for (int e = 0; e < Sa; e += 2) {
            // action e from buffer 0, action e + 1 from buffer 1
            if (e + 1 < Sa) fetch(e + 1);
            if (act) {
                const double t = terms(smem3);
                acc += neg ? -t : t;
            }
            if (e + 1 >= Sa) break;
            store(smem3 + BUFB);
            neg = neg_next;
            __syncthreads();
            if (e + 2 < Sa) fetch(e + 2);
            if (act) {
                const double t = terms(smem3 + BUFB);
                acc += neg ? -t : t;
            }
            if (e + 2 >= Sa) break;
            store(smem3);
            neg = neg_next;
            __syncthreads();
        }
    } else {
        for (int e = 0; e < Sa; ++e) {
            const bool more = e + 1 < Sa;
            if (more) fetch(e + 1);
            if (act) {
                const double t = terms(smem3);
                acc += neg ? -t : t;
            }
            if (!more) break;
            __syncthreads();  // every thread has read the buffer
            store(smem3);
            neg = neg_next;
            __syncthreads();
        }
    }
    if (act) sig[Ia * db + Ib] += acc;
}

// Shapes of kernel 3: NS = beta singles + beta electrons of the space;
// workgroup size, row elements per thread and buffer spacing per NS.
// LDSB: static LDS bytes (two buffers of BUFB, or one buffer of LDSB bytes).
template <typename CT, int NT, int PF, int NS, int BUFB, int LDSB>
static bool ab3_prepare(const SigmaArgs& a0, SigmaArgs& a, uint64_t& grid) {
    static_assert(LDSB <= (int)kLdsMax && (BUFB == 0 || LDSB == 2 * BUFB), "kernel 3 LDS layout");
    constexpr uint32_t S = sizeof(CT);
    const uint64_t L2 = (uint64_t)a0.L * a0.L;
    const uint64_t buf = 2 * L2 * S + a0.db * S;
    if (2 * L2 * S > 0x1FFFu || buf >= (1u << 19)) return false;   // the two packed offsets
    if ((a0.db + NT - 1) / NT > (uint64_t)PF || 2 * L2 > (uint64_t)NT * 2) return false;
    if (buf > (uint64_t)(BUFB > 0 ? BUFB : LDSB)) return false;
    a = a0;
    a.nchunk = (uint32_t)((a.db + NT - 1) / NT);
    const uint64_t nb = a.da * (uint64_t)a.nchunk;
    if (nb > 0xFFFFFFF0ull) return false;
    a.nblocks = (uint32_t)nb;
    grid = (nb + 7) / 8 * 8;
    return true;
}

// sb: a batched product over `lanes` lanes (grid rows), each with its own args.
template <typename CT, int NT, int PF, int NS, int BUFB, int LDSB, bool PIN = false>
static bool launch_ab3_t(const SigmaArgs& a0, const CT* C, double* sig, hipStream_t s, SigmaArgs* prepared = nullptr,
                         const SigmaBatch* sb = nullptr, int lanes = 1) {
    SigmaArgs a;
    uint64_t grid = 0;
    if (!ab3_prepare<CT, NT, PF, NS, BUFB, LDSB>(a0, a, grid)) return false;
    if (prepared) {  // shape query only
        *prepared = a;
        return true;
    }
    const uint64_t L2 = (uint64_t)a0.L * a0.L;
    const bool pv1 = 2 * L2 <= (uint64_t)NT;
    if (sb) {
        const dim3 g((unsigned)grid, (unsigned)lanes);
        if (pv1) k_sigma_ab3<CT, NT, PF, 1, NS, BUFB, LDSB, true, PIN><<<g, NT, 0, s>>>(a, nullptr, nullptr, *sb);
        else k_sigma_ab3<CT, NT, PF, 2, NS, BUFB, LDSB, true, PIN><<<g, NT, 0, s>>>(a, nullptr, nullptr, *sb);
    } else {
        const SigmaBatch none{};
        if (pv1) k_sigma_ab3<CT, NT, PF, 1, NS, BUFB, LDSB, false, PIN><<<(unsigned)grid, NT, 0, s>>>(a, C, sig, none);
        else k_sigma_ab3<CT, NT, PF, 2, NS, BUFB, LDSB, false, PIN><<<(unsigned)grid, NT, 0, s>>>(a, C, sig, none);
    }
    CF_HIP(hipGetLastError());
    return true;
}

template <typename CT>
static bool ab3_dispatch(const SigmaArgs& a, int nb_electrons, const CT* C, double* sig, hipStream_t s,
                         SigmaArgs* prepared, const SigmaBatch* sb, int lanes) {
    const int ns = (a.Sb - 1) + nb_electrons;
    constexpr int S = (int)sizeof(CT);
#define CF_AB3(NT, PF, NSV, BUFB, LDSB) \
    launch_ab3_t<CT, NT, PF, NSV, BUFB, LDSB>(a, C, sig, s, prepared, sb, lanes)
#define CF_AB3P(NT, PF, NSV, BUFB, LDSB) \
    launch_ab3_t<CT, NT, PF, NSV, BUFB, LDSB, true>(a, C, sig, s, prepared, sb, lanes)
    switch (ns) {
        case 12: return CF_AB3(128, 1, 12, 4096, 8192);       // L=6 3b, L=7 2b
        case 20: return CF_AB3(128, 1, 20, 4096, 8192);       // L=8 4b
        case 24: return CF_AB3(128, 1, 24, 4096, 8192);       // L=9 4b
        case 30: return CF_AB3(256, 1, 30, 4096, 8192);       // L=10 5b
        case 35: return CF_AB3(256, 2, 35, 8192, 16384);      // L=11 5b
        // L=12 6b, codes pinned: 0.417 / 0.453 ms against 0.471 / 0.464 ms
        // (FP32 / FP64 sigma, MI355X); at L <= 11 pinning was neutral or slower
        case 42: return CF_AB3P(512, 2, 42, 16384, 32768);
        // L=16 8b: d_b = 12,870.  1024 threads with the codes pinned (at most
        // 128 registers, 4 waves per SIMD); FP32: two buffers of 53,528 bytes,
        // FP64: one buffer of 107,056 bytes.  Measured on MI355X: 251 / 302 ms
        // against 302 / 435 ms for 512 threads with one buffer and hoisted
        // offsets (214 / 248 registers, 2 waves per SIMD).
        case 72:
            if constexpr (S == 4) return CF_AB3P(1024, 13, 72, 53528, 107056);
            else return CF_AB3P(1024, 13, 72, 0, 107056);
        default: return false;
    }
#undef CF_AB3
#undef CF_AB3P
}

template <typename CT>
bool launch_sigma_ab3(const SigmaArgs& a, int nb_electrons, const CT* C, double* sig, hipStream_t s) {
    const bool tables = sizeof(CT) == 4 ? a.vpair32 != nullptr : a.vpair64 != nullptr;
    if (!tables) return false;
    return ab3_dispatch<CT>(a, nb_electrons, C, sig, s, nullptr, nullptr, 1);
}
template bool launch_sigma_ab3<double>(const SigmaArgs&, int, const double*, double*, hipStream_t);
template bool launch_sigma_ab3<float>(const SigmaArgs&, int, const float*, double*, hipStream_t);

bool sigma_ab3_batch_args(SigmaArgs& a, int nb_electrons, bool f32) {
    SigmaArgs out;
    const bool ok = f32 ? ab3_dispatch<float>(a, nb_electrons, nullptr, nullptr, nullptr, &out, nullptr, 1)
                        : ab3_dispatch<double>(a, nb_electrons, nullptr, nullptr, nullptr, &out, nullptr, 1);
    if (ok) a = out;
    return ok;
}

template <typename CT>
void launch_sigma_ab3_batch(const SigmaArgs& shape, int nb_electrons, const SigmaBatch& sb, int lanes, hipStream_t s) {
    if (lanes < 1 || lanes > kBdMaxLanes) throw Error(ERR_ARG, "batched sigma: 1..32 lanes");
    if (!ab3_dispatch<CT>(shape, nb_electrons, nullptr, nullptr, s, nullptr, &sb, lanes))
        throw Error(ERR_UNSUPPORTED, "batched sigma: no kernel 3 instance for this shape");
}
template void launch_sigma_ab3_batch<double>(const SigmaArgs&, int, const SigmaBatch&, int, hipStream_t);
template void launch_sigma_ab3_batch<float>(const SigmaArgs&, int, const SigmaBatch&, int, hipStream_t);

__global__ void k_to_f32(uint64_t n, const double* __restrict__ x, float* __restrict__ y) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = (float)x[i];
}
__global__ void k_to_f64(uint64_t n, const float* __restrict__ x, double* __restrict__ y) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) y[i] = (double)x[i];
}
void launch_to_f32(uint64_t n, const double* x, float* y, hipStream_t s) {
    k_to_f32<<<8192, 256, 0, s>>>(n, x, y);
    CF_HIP(hipGetLastError());
}
void launch_to_f64(uint64_t n, const float* x, double* y, hipStream_t s) {
    k_to_f64<<<8192, 256, 0, s>>>(n, x, y);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// 1-RDM without a stored matrix (large FCI spaces): for every alpha string K
// and single K -> J (p in K, r in J, phase s): gamma^a_pr += s * C(K,:).C(J,:);
// the diagonal gamma^a_pp = sum_K n_p(K) |C(K,:)|^2.  Beta from the transposed
// vector.  One workgroup per K, fixed-order reductions (deterministic).
// out[K][e] = row dot products for the S entries of K's self+singles list.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_row_dots(uint64_t da, uint64_t db, int S, const Exc* __restrict__ sub,
                                                  const double* __restrict__ C, double* __restrict__ out) {
    __shared__ double red[4];
    const uint64_t K = blockIdx.x;
    const double* ck = C + K * db;
    const Exc* lst = sub + K * (uint64_t)S;
    for (int e = 0; e < S; ++e) {
        const double* cj = C + (uint64_t)lst[e].J * db;
        double acc = 0.0;
        for (uint64_t i = threadIdx.x; i < db; i += 256) acc += ck[i] * cj[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) out[K * (uint64_t)S + e] = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
    }
}

void launch_row_dots(uint64_t da, uint64_t db, int S, const Exc* sub, const double* C, double* out, hipStream_t s) {
    if (!da) return;
    k_row_dots<<<(unsigned)da, 256, 0, s>>>(da, db, S, sub, C, out);
    CF_HIP(hipGetLastError());
}

// out = in^T for a rows x cols row-major matrix (tiled through LDS)
__global__ void k_transpose(uint64_t rows, uint64_t cols, const double* __restrict__ in, double* __restrict__ out) {
    __shared__ double tile[32][33];
    const uint64_t c0 = (uint64_t)blockIdx.x * 32, r0 = (uint64_t)blockIdx.y * 32;
    for (int i = threadIdx.y; i < 32; i += 8) {
        const uint64_t r = r0 + i, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[i][threadIdx.x] = in[r * cols + c];
    }
    __syncthreads();
    for (int i = threadIdx.y; i < 32; i += 8) {
        const uint64_t c = c0 + i, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[c * rows + r] = tile[threadIdx.x][i];
    }
}

void launch_transpose(uint64_t rows, uint64_t cols, const double* in, double* out, hipStream_t s) {
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    if (grid.y > 65535) throw Error(ERR_UNSUPPORTED, "transpose too tall");
    k_transpose<<<grid, dim3(32, 8), 0, s>>>(rows, cols, in, out);
    CF_HIP(hipGetLastError());
}

}  // namespace cf
