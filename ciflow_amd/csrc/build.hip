// Slater-Condon CRS build on the GPU: count pass, exclusive scan, fill pass.
//
// Replaces FCI::construct_CI_matrix / DOCI::construct_CI_matrix
// (CIMethod.cpp:611-660, 749-888) driven by CIMethod::build_parallel
// (:489-561) and SparseMatrix_CRS::PushToRow/add_from_list
// (SparseMatrix_CRS.cpp:240-322, 663-696).
//
// The reference stores the upper triangle and its SpMV scatters into y[col];
// the GPU keeps FULL rows so the SpMV is a race-free gather.  Each row is
// produced by one wavefront directly in ascending column order (no insertion
// sort), entries with |v| <= zero are dropped exactly as PushToRow drops them,
// and the surviving entries are compacted with a ballot + mbcnt prefix.
//
// Element values follow the reference's formulas AND summation order, so the
// upper part of every row matches the reference bit for bit:
//   diagonal      (D1a + D1b) + sum_{k in a} sum_{l in b} V(k,l,k,l)   (:398-407)
//   single        [h_pq + sum_z (V_pzqz - V_pzzq)] + sum_{d in other} V_pdqd, then sign (:409-427)
//   ab double     V(a_up1, b_down1, a_up2, b_down2) with signs of both strings (:444-457)
//   same-spin dbl precomputed per string pair (:429-442)
// Lower-triangle entries are evaluated in the orientation the reference uses
// for the mirrored upper entry, so the stored matrix is exactly symmetric.
//
// FCI row layout: for row (ia, ib) the columns are ja*db + jb with ja running
// over the sorted alpha list of ia; an alpha "self" entry expands to the full
// sorted beta list of ib, an alpha single to the beta self+singles list, an
// alpha double to the single column ja*db + ib.  segoff[ia] holds the prefix
// offsets of these segments; a lane finds its segment by binary search.
#include "kernels.h"

#include <hipcub/hipcub.hpp>
#include "string_ops.h"

namespace cf {

__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__global__ void k_fci_segoff(uint64_t da, int Ea, const Exc* __restrict__ full_a, int Eb, int Sb,
                             uint32_t* __restrict__ segoff) {
    uint64_t ia = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ia >= da) return;
    const Exc* lst = full_a + ia * (uint64_t)Ea;
    uint32_t* so = segoff + ia * (uint64_t)(Ea + 1);
    uint32_t acc = 0;
    for (int e = 0; e < Ea; ++e) {
        so[e] = acc;
        int t = lst[e].type;
        acc += t == 0 ? (uint32_t)Eb : (t == 1 ? (uint32_t)Sb : 1u);
    }
    so[Ea] = acc;
}

void launch_fci_segoff(uint64_t da, int Ea, const Exc* full_a, int Eb, int Sb, uint32_t* segoff,
                       hipStream_t stream) {
    if (!da) return;
    k_fci_segoff<<<(unsigned)((da + 255) / 256), 256, 0, stream>>>(da, Ea, full_a, Eb, Sb, segoff);
    CF_HIP(hipGetLastError());
}

// jmix of CIMethod::diagonal (CIMethod.cpp:343-351): sum over alpha k (outer),
// beta l (inner), accumulated from zero.
__device__ __forceinline__ double jmix(str_t a, str_t b, const HamDev& H) {
    double s = 0.0;
    for (str_t t = a; t; t &= t - 1) {
        int k = __builtin_ctzll(t);
        for (str_t u = b; u; u &= u - 1) s += H.Jm[k * H.L + __builtin_ctzll(u)];
    }
    return s;
}

__global__ void k_fci_diag(uint64_t da, uint64_t db, const str_t* __restrict__ sa, const str_t* __restrict__ sb,
                           const double* __restrict__ d1a, const double* __restrict__ d1b, HamDev H,
                           double* __restrict__ diag) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= da * db) return;
    uint64_t ia = r / db, ib = r - ia * db;
    double s = d1a[ia];
    s += d1b[ib];
    s += jmix(sa[ia], sb[ib], H);
    diag[r] = s;
}

void launch_fci_diag(uint64_t da, uint64_t db, const str_t* sa, const str_t* sb, const double* d1a,
                     const double* d1b, HamDev H, double* diag, hipStream_t stream) {
    uint64_t n = da * db;
    if (!n) return;
    k_fci_diag<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(da, db, sa, sb, d1a, d1b, H, diag);
    CF_HIP(hipGetLastError());
}

__global__ void k_doci_diag(uint64_t d, const str_t* __restrict__ s, const double* __restrict__ d1, HamDev H,
                            double* __restrict__ diag) {
    uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d) return;
    double v = d1[r];
    v += d1[r];
    v += jmix(s[r], s[r], H);
    diag[r] = v;
}

void launch_doci_diag(uint64_t d, const str_t* s, const double* d1, HamDev H, double* diag, hipStream_t stream) {
    if (!d) return;
    k_doci_diag<<<(unsigned)((d + 255) / 256), 256, 0, stream>>>(d, s, d1, H, diag);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// FCI rows: one wavefront per row.
// ---------------------------------------------------------------------------
// Element u of the segment that alpha-list entry ea opens in row (ia, ib):
// column and value with the reference's formulas and summation order.
__device__ __forceinline__ void fci_entry(const FciRowsArgs& a, uint64_t ia, uint64_t ib, str_t Sa, str_t Sb,
                                          const Exc& ea, const Exc* lb, const Exc* sbl, uint32_t u, uint32_t& col,
                                          double& val) {
    const int L = a.H.L;
    if (ea.type == 0) {
        const Exc eb = lb[u];
        col = (uint32_t)(ia * a.db + eb.J);
        if (eb.type == 0) {
            val = a.diag[ia * a.db + ib];
        } else if (eb.type == 1) {
            int p = eb.f < eb.t ? eb.f : eb.t, q = eb.f < eb.t ? eb.t : eb.f;
            const double* w = a.H.W + ((size_t)p * L + q) * L;
            val = eb.x;
            for (str_t s = Sa; s; s &= s - 1) val += w[__builtin_ctzll(s)];
            if (eb.sign < 0) val *= -1;
        } else {
            val = eb.x;
        }
    } else if (ea.type == 1) {
        const Exc eb = sbl[u];
        col = (uint32_t)((uint64_t)ea.J * a.db + eb.J);
        if (eb.type == 0) {
            int p = ea.f < ea.t ? ea.f : ea.t, q = ea.f < ea.t ? ea.t : ea.f;
            const double* w = a.H.W + ((size_t)p * L + q) * L;
            val = ea.x;
            for (str_t s = Sb; s; s &= s - 1) val += w[__builtin_ctzll(s)];
            if (ea.sign < 0) val *= -1;
        } else {
            size_t i0, i1, i2, i3;
            if (ea.J > ia) { i0 = ea.f; i1 = eb.f; i2 = ea.t; i3 = eb.t; }
            else { i0 = ea.t; i1 = eb.t; i2 = ea.f; i3 = eb.f; }
            val = a.H.V[((i0 * L + i1) * L + i2) * L + i3];
            if ((ea.sign < 0) != (eb.sign < 0)) val *= -1;
        }
    } else {
        col = (uint32_t)((uint64_t)ea.J * a.db + ib);
        val = ea.x;
    }
}

template <bool FILL>
__global__ void __launch_bounds__(256) k_fci_rows(FciRowsArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t r = a.row0 + (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= a.row0 + a.nrows) return;
    const uint64_t ia = r / a.db, ib = r - ia * a.db;
    const str_t Sa = a.sa[ia], Sb = a.sb[ib];
    const uint32_t* so = a.segoff + ia * (uint64_t)(a.Ea + 1);
    const uint32_t R = so[a.Ea];
    const Exc* la = a.fa + ia * (uint64_t)a.Ea;
    const Exc* lb = a.fb + ib * (uint64_t)a.Eb;
    const Exc* sbl = a.subb + ib * (uint64_t)a.Sb;
    int64_t base = 0;
    if (FILL) base = a.rowptr[r];
    uint32_t cnt = 0;
    for (uint32_t t0 = 0; t0 < R; t0 += 64) {
        const uint32_t t = t0 + lane;
        bool keep = false;
        uint32_t col = 0;
        double val = 0.0;
        if (t < R) {
            int lo = 0, hi = a.Ea;  // largest e with so[e] <= t
            while (hi - lo > 1) {
                int mid = (lo + hi) >> 1;
                if (so[mid] <= t) lo = mid;
                else hi = mid;
            }
            fci_entry(a, ia, ib, Sa, Sb, la[lo], lb, sbl, t - so[lo], col, val);
            keep = fabs(val) > a.zero;
        }
        const uint64_t mask = __ballot(keep);
        if (FILL && keep) {
            const int64_t pos = base + cnt + lanes_below(mask);
            a.col[pos] = col;
            a.val[pos] = val;
            if (a.val32) a.val32[pos] = (float)val;
        }
        cnt += (uint32_t)__popcll(mask);
    }
    if (!FILL && lane == 0) a.rowcnt[r - a.row0] = cnt;
}

// Position-major copy of the complete FCI rows of one alpha string:
// ell[(ia*R + t)*db + ib] = element t of row (ia, ib), 0 where the element is
// not stored (|v| <= zero).  One thread per beta string, so the writes of a
// wavefront are contiguous.
__global__ void __launch_bounds__(256) k_fci_ell(FciRowsArgs a, uint32_t chunks, double* __restrict__ ell,
                                                 float* __restrict__ ell32) {
    const uint64_t ia = blockIdx.x / chunks;
    const uint64_t ib = (uint64_t)(blockIdx.x - ia * chunks) * 256 + threadIdx.x;
    if (ia >= a.da || ib >= a.db) return;
    const str_t Sa = a.sa[ia], Sb = a.sb[ib];
    const uint32_t* so = a.segoff + ia * (uint64_t)(a.Ea + 1);
    const uint64_t R = so[a.Ea];
    const Exc* la = a.fa + ia * (uint64_t)a.Ea;
    const Exc* lb = a.fb + ib * (uint64_t)a.Eb;
    const Exc* sbl = a.subb + ib * (uint64_t)a.Sb;
    for (int k = 0; k < a.Ea; ++k) {
        const Exc ea = la[k];
        const uint32_t len = so[k + 1] - so[k];
        for (uint32_t u = 0; u < len; ++u) {
            uint32_t col;
            double val;
            fci_entry(a, ia, ib, Sa, Sb, ea, lb, sbl, u, col, val);
            if (!(fabs(val) > a.zero)) val = 0.0;
            const uint64_t pos = (ia * R + so[k] + u) * a.db + ib;
            if (ell) ell[pos] = val;
            if (ell32) ell32[pos] = (float)val;
        }
    }
}

void launch_fci_ell(const FciRowsArgs& a, double* ell, float* ell32, hipStream_t stream) {
    const uint64_t chunks = (a.db + 255) / 256;
    const uint64_t grid = a.da * chunks;
    if (!grid) return;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "ELL build grid too large");
    k_fci_ell<<<(unsigned)grid, 256, 0, stream>>>(a, (uint32_t)chunks, ell, ell32);
    CF_HIP(hipGetLastError());
}

void launch_fci_rows(const FciRowsArgs& a, bool fill, hipStream_t stream) {
    if (!a.nrows) return;
    unsigned grid = (unsigned)((a.nrows + 3) / 4);
    if (fill) k_fci_rows<true><<<grid, 256, 0, stream>>>(a);
    else k_fci_rows<false><<<grid, 256, 0, stream>>>(a);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// DOCI rows (CIMethod.cpp:611-660): pair hops j -> k of one string.  Columns in
// ascending order without sorting: hops that lower the string come first
// (occupied j descending, then virtual k < j ascending), then the diagonal,
// then hops that raise it (virtual k ascending, then occupied j < k
// descending).  Element: V(a,a,b,b) with a the orbital occupied in the
// smaller of the two strings, i.e. two_diff_orbitals(up1,up2,up1,up2) of the
// reference's upper triangle; the phase is always +1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int select_bit(str_t v, int m) {  // position of the m-th (0-based) set bit
    for (int i = 0; i < m; ++i) v &= v - 1;
    return __builtin_ctzll(v);
}

template <bool FILL>
__global__ void __launch_bounds__(256) k_doci_rows(DociRowsArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t r = a.row0 + (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= a.row0 + a.nrows) return;
    const int L = a.L;
    const str_t full = (L == 64) ? ~0ull : (((str_t)1 << L) - 1);
    const str_t s = a.s[r];
    const str_t vir = ~s & full;
    uint32_t ndown = 0;
    for (str_t t = s; t; t &= t - 1) {
        int j = __builtin_ctzll(t);
        ndown += (uint32_t)__popcll(vir & (((str_t)1 << j) - 1));
    }
    const uint32_t R = 1u + (uint32_t)a.np * (uint32_t)(L - a.np);
    int64_t base = 0;
    if (FILL) base = a.rowptr[r];
    uint32_t cnt = 0;
    for (uint32_t t0 = 0; t0 < R; t0 += 64) {
        const uint32_t t = t0 + lane;
        bool keep = false;
        uint32_t col = 0;
        double val = 0.0;
        if (t < R) {
            if (t == ndown) {
                col = (uint32_t)r;
                val = a.diag[r];
            } else {
                int j = 0, k = 0;
                if (t < ndown) {
                    uint32_t rem = t;
                    for (int jj = L - 1; jj >= 0; --jj) {
                        if (!((s >> jj) & 1)) continue;
                        str_t below = vir & (((str_t)1 << jj) - 1);
                        uint32_t c = (uint32_t)__popcll(below);
                        if (rem < c) { j = jj; k = select_bit(below, (int)rem); break; }
                        rem -= c;
                    }
                } else {
                    uint32_t rem = t - ndown - 1;
                    for (int kk = 0; kk < L; ++kk) {
                        if ((s >> kk) & 1) continue;
                        str_t below = s & (((str_t)1 << kk) - 1);
                        uint32_t c = (uint32_t)__popcll(below);
                        if (rem < c) { k = kk; j = select_bit(below, (int)(c - 1 - rem)); break; }
                        rem -= c;
                    }
                }
                const str_t tgt = s ^ ((str_t)1 << j) ^ ((str_t)1 << k);
                col = rank_bits(tgt, a.binom);
                const size_t lo = j < k ? j : k, hi = j < k ? k : j;
                val = a.H.V[((lo * L + lo) * L + hi) * L + hi];
            }
            keep = fabs(val) > a.zero;
        }
        const uint64_t mask = __ballot(keep);
        if (FILL && keep) {
            const int64_t pos = base + cnt + lanes_below(mask);
            a.col[pos] = col;
            a.val[pos] = val;
            if (a.val32) a.val32[pos] = (float)val;
        }
        cnt += (uint32_t)__popcll(mask);
    }
    if (!FILL && lane == 0) a.rowcnt[r - a.row0] = cnt;
}

void launch_doci_rows(const DociRowsArgs& a, bool fill, hipStream_t stream) {
    if (!a.nrows) return;
    unsigned grid = (unsigned)((a.nrows + 3) / 4);
    if (fill) k_doci_rows<true><<<grid, 256, 0, stream>>>(a);
    else k_doci_rows<false><<<grid, 256, 0, stream>>>(a);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Determinant-list rows (selected CI).  get_ham_element (CIMethod.cpp:459-487)
// between two arbitrary determinants, restated on bit strings with the
// reference's formulas and summation order; one wavefront per row, the
// lanes test 64 columns at a time (popcount of the differences first), and
// the ballot of |v| > zero gives each survivor its slot.  Lower-triangle
// entries are evaluated as get_ham_element(det_j, det_i), the orientation of
// the reference's upper entry, so the stored matrix is exactly symmetric.
// O(n^2) pair tests per pass, as the reference's construct_CI_matrix.
// ---------------------------------------------------------------------------
// one_diff_orbital(s1, s2, other) (CIMethod.cpp:409-427)
__device__ __forceinline__ double det_one_diff(str_t s1, str_t s2, str_t other, const HamDev& H) {
    const int L = H.L;
    const str_t x = s1 ^ s2, com = s1 & s2;
    const int o0 = __builtin_ctzll(x), o1 = 63 - __builtin_clzll(x);
    double e = H.h[o0 * L + o1];
    for (str_t t = com; t; t &= t - 1) {
        const int z = __builtin_ctzll(t);
        e += H.V[(((size_t)o0 * L + z) * L + o1) * L + z] - H.V[(((size_t)o0 * L + z) * L + z) * L + o1];
    }
    const int diff = stacked(com, o0, o1);
    const double* w = H.W + ((size_t)o0 * L + o1) * L;  // V(o0, d, o1, d)
    for (str_t t = other; t; t &= t - 1) e += w[__builtin_ctzll(t)];
    return (diff & 1) ? -e : e;
}

// two_diff_orbitals(s1, s2), same spin (CIMethod.cpp:429-442): the orbitals of
// s1 first, then those of s2, each ascending.
__device__ __forceinline__ double det_two_same(str_t s1, str_t s2, const HamDev& H) {
    const int L = H.L;
    const str_t com = s1 & s2, a = s1 & ~s2, b = s2 & ~s1;
    const int a0 = __builtin_ctzll(a), a1 = 63 - __builtin_clzll(a);
    const int b0 = __builtin_ctzll(b), b1 = 63 - __builtin_clzll(b);
    const int diff = stacked(com, a0, a1) + stacked(com, b0, b1);
    const double e = H.V[(((size_t)a0 * L + a1) * L + b0) * L + b1] - H.V[(((size_t)a0 * L + a1) * L + b1) * L + b0];
    return (diff & 1) ? -e : e;
}

// two_diff_orbitals(up1, up2, down1, down2) (CIMethod.cpp:444-457)
__device__ __forceinline__ double det_two_mixed(str_t u1, str_t u2, str_t d1, str_t d2, const HamDev& H) {
    const int L = H.L;
    const int o0 = __builtin_ctzll(u1 & ~u2), o1 = __builtin_ctzll(u2 & ~u1);
    const int p0 = __builtin_ctzll(d1 & ~d2), p1 = __builtin_ctzll(d2 & ~d1);
    const int diff = stacked(u1 & u2, o0, o1) + stacked(d1 & d2, p0, p1);
    const double e = H.V[(((size_t)o0 * L + p0) * L + o1) * L + p1];
    return (diff & 1) ? -e : e;
}

__device__ __forceinline__ double det_element(str_t u1, str_t d1, str_t u2, str_t d2, const HamDev& H) {
    const int pu = __popcll(u1 ^ u2), pd = __popcll(d1 ^ d2);
    if (pd > 4 || pu > 4) return 0.0;
    if (pu == 4 && pd == 0) return det_two_same(u1, u2, H);
    if (pu == 2) {
        if (pd == 0) return det_one_diff(u1, u2, d1, H);
        if (pd == 2) return det_two_mixed(u1, u2, d1, d2, H);
        return 0.0;
    }
    if (pd == 4 && pu == 0) return det_two_same(d1, d2, H);
    if (pd == 2 && pu == 0) return det_one_diff(d1, d2, u1, H);
    if (pd == 0 && pu == 0) {  // diagonal (CIMethod.cpp:398-407)
        double s = diag_one_string(u1, H);
        s += diag_one_string(d1, H);
        s += jmix(u1, d1, H);
        return s;
    }
    return 0.0;
}

template <bool FILL>
__global__ void __launch_bounds__(256) k_det_rows(DetRowsArgs a) {
    const int lane = threadIdx.x & 63;
    const uint64_t i = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= a.n) return;
    const str_t ui = a.up[i], di = a.down[i];
    int64_t base = 0;
    if (FILL) base = a.rowptr[i];
    uint32_t cnt = 0;
    for (uint64_t j0 = 0; j0 < a.n; j0 += 64) {
        const uint64_t j = j0 + lane;
        bool keep = false;
        double val = 0.0;
        if (j < a.n) {
            const str_t uj = a.up[j], dj = a.down[j];
            val = j >= i ? det_element(ui, di, uj, dj, a.H) : det_element(uj, dj, ui, di, a.H);
            keep = fabs(val) > a.zero;
        }
        const uint64_t mask = __ballot(keep);
        if (FILL && keep) {
            const int64_t pos = base + cnt + lanes_below(mask);
            a.col[pos] = (uint32_t)j;
            a.val[pos] = val;
            if (a.val32) a.val32[pos] = (float)val;
        }
        cnt += (uint32_t)__popcll(mask);
    }
    if (!FILL && lane == 0) a.rowcnt[i] = cnt;
}

// ---------------------------------------------------------------------------
// Determinant-list rows by excitation lookup (L <= 32).  The list's keys
// (up << 32 | down) are sorted once; the columns of row i are the list
// determinants that at most two excitations of det_i reach: the entries with
// up_i (one binary-searched range, filtered by |down_i ^ down_j| <= 4), those
// whose up string is a single excitation of up_i (their range, filtered by
// |down_i ^ down_j| <= 2, or for long ranges looked up by down_i and its
// single excitations), and the keys (double excitation of up_i, down_i)
// looked up exactly.  That is O(n (E_a log n + group sizes)) work instead of
// n^2 pair tests.  Each candidate's element is get_ham_element in the
// orientation k_det_rows uses; survivors take slots from a per-wavefront LDS
// counter (unordered) and each row is sorted by column afterwards, so the
// CRS equals k_det_rows' bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t lower_bound_key(const uint64_t* __restrict__ a, uint64_t lo, uint64_t hi,
                                                    uint64_t key) {
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ int kth_set_bit(str_t v, int k) {
    for (int i = 0; i < k; ++i) v &= v - 1;
    return __builtin_ctzll(v);
}

// (a, b), a < b, of pair index p among k items in the order (0,1), (0,2), ...
__device__ __forceinline__ void pair_of(int p, int k, int& a, int& b) {
    a = 0;
    while (p >= k - 1 - a) {
        p -= k - 1 - a;
        ++a;
    }
    b = a + 1 + p;
}

// emit(j, u_j, d_j) for every list entry j that at most two excitations of
// det_i reach (row i's candidates), each lane in a fixed order.
template <class Emit>
__device__ __forceinline__ void det_row_candidates(const DetLookupArgs& A, uint64_t i, int lane, Emit&& emit) {
    const DetRowsArgs& a = A.r;
    const int L = A.L;
    const uint64_t n = a.n;
    const str_t ui = a.up[i], di = a.down[i];
    const str_t full = L >= 64 ? ~(str_t)0 : (((str_t)1 << L) - 1);
    const int na = a.nup, va = L - a.nup;
    const int nb = a.ndown, vb = L - a.ndown;
    const uint64_t* __restrict__ key = A.skey;
    // the sorted range of the entries with up string u
    auto group = [&](str_t u, uint64_t& g0, uint64_t& g1) {
        const uint64_t k0 = (uint64_t)u << 32, k1 = k0 + (1ull << 32);
        g0 = lower_bound_key(key, 0, n, k0);
        g1 = k1 == 0 ? n : lower_bound_key(key, g0, n, k1);
    };
    // every list entry with key k in [lo, hi)
    auto take_equal = [&](uint64_t lo, uint64_t hi, uint64_t k, str_t uj, str_t dj) {
        for (uint64_t t = lower_bound_key(key, lo, hi, k); t < hi && key[t] == k; ++t) emit(A.sidx[t], uj, dj);
    };
    // same up string: down within two excitations (diagonal, beta singles and doubles)
    {
        uint64_t g0, g1;
        group(ui, g0, g1);
        for (uint64_t t = g0 + lane; t < g1; t += 64) {
            const str_t d = (str_t)(key[t] & 0xffffffffull);
            if (__popcll(di ^ d) <= 4) emit(A.sidx[t], ui, d);
        }
    }
    // up string one alpha single away: down equal or one beta single away
    for (int sidx = lane; sidx < na * va; sidx += 64) {
        const int o = kth_set_bit(ui, sidx / va), e = kth_set_bit(~ui & full, sidx % va);
        const str_t u2 = ui ^ ((str_t)1 << o) ^ ((str_t)1 << e);
        const uint64_t k0 = (uint64_t)u2 << 32;
        uint64_t g0, g1;
        group(u2, g0, g1);
        if (g1 - g0 <= (uint64_t)(2 * (1 + nb * vb))) {
            for (uint64_t t = g0; t < g1; ++t) {
                const str_t d = (str_t)(key[t] & 0xffffffffull);
                if (__popcll(di ^ d) <= 2) emit(A.sidx[t], u2, d);
            }
        } else {
            take_equal(g0, g1, k0 | di, u2, di);
            for (int q = 0; q < nb * vb; ++q) {
                const int bo = kth_set_bit(di, q / vb), be = kth_set_bit(~di & full, q % vb);
                const str_t d2 = di ^ ((str_t)1 << bo) ^ ((str_t)1 << be);
                take_equal(g0, g1, k0 | d2, u2, d2);
            }
        }
    }
    // up string one alpha double away, down equal
    const int pa = na * (na - 1) / 2, pv = va * (va - 1) / 2;
    for (int didx = lane; didx < pa * pv; didx += 64) {
        int o0, o1, e0, e1;
        pair_of(didx / pv, na, o0, o1);
        pair_of(didx % pv, va, e0, e1);
        const str_t empty = ~ui & full;
        const str_t u2 = ui ^ ((str_t)1 << kth_set_bit(ui, o0)) ^ ((str_t)1 << kth_set_bit(ui, o1)) ^
                         ((str_t)1 << kth_set_bit(empty, e0)) ^ ((str_t)1 << kth_set_bit(empty, e1));
        take_equal(0, n, ((uint64_t)u2 << 32) | di, u2, di);
    }
}

template <bool FILL>
__global__ void __launch_bounds__(256) k_det_rows_lookup(DetLookupArgs A) {
    __shared__ uint32_t wcnt[4];
    const DetRowsArgs& a = A.r;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t i = (uint64_t)blockIdx.x * 4 + w;
    const bool act = i < a.n;
    if (lane == 0) wcnt[w] = 0;
    __syncthreads();
    if (act) {
        const str_t ui = a.up[i], di = a.down[i];
        int64_t base = 0;
        if (FILL) base = a.rowptr[i];
        det_row_candidates(A, i, lane, [&](uint64_t j, str_t uj, str_t dj) {
            const double v = j >= i ? det_element(ui, di, uj, dj, a.H) : det_element(uj, dj, ui, di, a.H);
            if (fabs(v) > a.zero) {
                const uint32_t pos = atomicAdd(&wcnt[w], 1u);
                if (FILL) {
                    a.col[base + pos] = (uint32_t)j;
                    a.val[base + pos] = v;
                }
            }
        });
    }
    __syncthreads();
    if (act && lane == 0 && !FILL) a.rowcnt[i] = wcnt[w];
}

// The product of a determinant list's Hamiltonian computed on the fly
// (CI_Big::mvprod, CIMethod.cpp:943-984): row i, one wavefront, sums
// H_ij x_j over its candidates, every element in the orientation of the
// stored matrix and without the drop rule (the reference's on-the-fly part
// uses every element it evaluates).  Elements with i, j < m belong to the
// stored leading block, whose product the caller has already written to
// y[0, m): those rows add, the others overwrite.  Each lane sums its
// candidates in its fixed traversal order and the wavefront reduces in a
// fixed tree, so repeated products are bitwise equal.
__global__ void __launch_bounds__(256) k_det_sigma_lookup(DetLookupArgs A, const double* __restrict__ x,
                                                          double* __restrict__ y, uint64_t m) {
    const DetRowsArgs& a = A.r;
    const int lane = threadIdx.x & 63;
    const uint64_t i = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= a.n) return;  // a whole wavefront; no barrier follows
    const str_t ui = a.up[i], di = a.down[i];
    double acc = 0.0;
    det_row_candidates(A, i, lane, [&](uint64_t j, str_t uj, str_t dj) {
        if (i < m && j < m) return;
        const double v = j >= i ? det_element(ui, di, uj, dj, a.H) : det_element(uj, dj, ui, di, a.H);
        acc = fma(v, x[j], acc);
    });
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) y[i] = i < m ? y[i] + acc : acc;
}

void launch_det_sigma_lookup(const DetLookupArgs& a, const double* x, double* y, uint64_t m, hipStream_t stream) {
    if (!a.r.n) return;
    if (a.L > 32) throw Error(ERR_UNSUPPORTED, "determinant-list on-the-fly product: L <= 32");
    const uint64_t grid = (a.r.n + 3) / 4;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "determinant list too long");
    k_det_sigma_lookup<<<(unsigned)grid, 256, 0, stream>>>(a, x, y, m);
    CF_HIP(hipGetLastError());
}

void launch_det_rows_lookup(const DetLookupArgs& a, bool fill, hipStream_t stream) {
    if (!a.r.n) return;
    if (a.L > 32) throw Error(ERR_UNSUPPORTED, "determinant-list lookup build: L <= 32");
    const uint64_t grid = (a.r.n + 3) / 4;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "determinant list too long");
    if (fill) k_det_rows_lookup<true><<<(unsigned)grid, 256, 0, stream>>>(a);
    else k_det_rows_lookup<false><<<(unsigned)grid, 256, 0, stream>>>(a);
    CF_HIP(hipGetLastError());
}

__global__ void k_det_keys(uint64_t n, const str_t* __restrict__ up, const str_t* __restrict__ down,
                           uint64_t* __restrict__ key, uint32_t* __restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    key[i] = ((uint64_t)up[i] << 32) | (uint64_t)down[i];
    idx[i] = (uint32_t)i;
}

// Sorted keys (up << 32 | down) of the list with the list index of each.
void sort_det_keys(uint64_t n, const str_t* up, const str_t* down, uint64_t* skey, uint32_t* sidx, int bits,
                   hipStream_t stream) {
    if (!n) return;
    if (n > 0x7fffffffull) throw Error(ERR_UNSUPPORTED, "determinant list too long for the key sort");
    DBuf<uint64_t> key(n);
    DBuf<uint32_t> idx(n);
    k_det_keys<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(n, up, down, key.p, idx.p);
    CF_HIP(hipGetLastError());
    size_t tb = 0;
    CF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, skey, idx.p, sidx, (int)n, 0, bits, stream));
    DBuf<unsigned char> tmp(tb);
    CF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, skey, idx.p, sidx, (int)n, 0, bits, stream));
}

// Each row's (col, val) pairs sorted by column (rows given by rowptr).
void sort_rows_by_column(uint64_t n, uint64_t nnz, const int64_t* rowptr, uint32_t* col, double* val,
                         hipStream_t stream) {
    if (!nnz) return;
    if (nnz > 0x7fffffffull || n > 0x7fffffffull) throw Error(ERR_UNSUPPORTED, "row sort: too many entries");
    DBuf<uint32_t> col2(nnz);
    DBuf<double> val2(nnz);
    size_t tb = 0;
    CF_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, tb, col, col2.p, val, val2.p, (int)nnz, (int)n,
                                                       rowptr, rowptr + 1, 0, 32, stream));
    DBuf<unsigned char> tmp(tb);
    CF_HIP(hipcub::DeviceSegmentedRadixSort::SortPairs(tmp.p, tb, col, col2.p, val, val2.p, (int)nnz, (int)n,
                                                       rowptr, rowptr + 1, 0, 32, stream));
    CF_HIP(hipMemcpyAsync(col, col2.p, nnz * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    CF_HIP(hipMemcpyAsync(val, val2.p, nnz * sizeof(double), hipMemcpyDeviceToDevice, stream));
}

void launch_det_rows(const DetRowsArgs& a, bool fill, hipStream_t stream) {
    if (!a.n) return;
    const uint64_t grid = (a.n + 3) / 4;
    if (grid > 0xffffffffull) throw Error(ERR_UNSUPPORTED, "determinant list too long");
    if (fill) k_det_rows<true><<<(unsigned)grid, 256, 0, stream>>>(a);
    else k_det_rows<false><<<(unsigned)grid, 256, 0, stream>>>(a);
    CF_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Exclusive scan u32 -> i64 (three phases: per-block scan, scan of block
// sums, add).  4096 items per 1024-thread block.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ int64_t wave_incl_scan(int64_t v) {
    const int lane = threadIdx.x & 63;
    for (int off = 1; off < 64; off <<= 1) {
        int64_t o = __shfl_up(v, off, 64);
        if (lane >= off) v += o;
    }
    return v;
}

__device__ int64_t block_excl_scan(int64_t v, int64_t* total) {
    __shared__ int64_t wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int64_t inc = wave_incl_scan(v);
    if (lane == 63) wsum[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        int64_t x = lane < kScanThreads / 64 ? wsum[lane] : 0;
        int64_t xi = wave_incl_scan(x);
        if (lane < kScanThreads / 64) wsum[lane] = xi - x;
        if (lane == kScanThreads / 64 - 1) *total = xi;
    }
    __syncthreads();
    int64_t r = inc - v + wsum[wid];
    __syncthreads();
    return r;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const uint32_t* __restrict__ cnt, int64_t* __restrict__ out,
                                                          uint64_t n, int64_t* __restrict__ bsum) {
    __shared__ int64_t tot;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
    for (int i = 0; i < kScanItems; ++i) {
        v[i] = (base + i < n) ? (int64_t)cnt[base + i] : 0;
        s += v[i];
    }
    int64_t pre = block_excl_scan(s, &tot);
    for (int i = 0; i < kScanItems; ++i) {
        if (base + i < n) out[base + i] = pre;
        pre += v[i];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_bsum(int64_t* __restrict__ bsum, uint64_t nb,
                                                         int64_t* __restrict__ grand) {
    __shared__ int64_t tot;
    int64_t carry = 0;
    for (uint64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
        uint64_t i = b0 + threadIdx.x;
        int64_t v = i < nb ? bsum[i] : 0;
        int64_t pre = block_excl_scan(v, &tot);
        if (i < nb) bsum[i] = pre + carry;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) *grand = carry;
}

__global__ void k_scan_add(int64_t* __restrict__ out, uint64_t n, const int64_t* __restrict__ bsum) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += bsum[i / kScanTile];
}

void exclusive_scan_u32_to_i64(const uint32_t* cnt, int64_t* out, uint64_t n, hipStream_t stream) {
    uint64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb == 0) nb = 1;
    DBuf<int64_t> bsum(nb);
    k_scan_tiles<<<(unsigned)nb, kScanThreads, 0, stream>>>(cnt, out, n, bsum.p);
    CF_HIP(hipGetLastError());
    k_scan_bsum<<<1, kScanThreads, 0, stream>>>(bsum.p, nb, out + n);
    CF_HIP(hipGetLastError());
    if (n) {
        k_scan_add<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(out, n, bsum.p);
        CF_HIP(hipGetLastError());
    }
    CF_HIP(hipStreamSynchronize(stream));  // bsum is freed on return
}

// ---------------------------------------------------------------------------
// Export helpers.
// ---------------------------------------------------------------------------
__global__ void k_count_upper(const int64_t* __restrict__ rowptr, const uint32_t* __restrict__ col, uint64_t r0,
                              uint64_t r1, uint64_t* __restrict__ cnt) {
    uint64_t r = r0 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= r1) return;
    int64_t lo = rowptr[r], hi = rowptr[r + 1];
    const int64_t end = hi;
    while (lo < hi) {  // first col >= r
        int64_t mid = (lo + hi) >> 1;
        if (col[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    cnt[r - r0] = (uint64_t)(end - lo);
}

void launch_count_upper(const int64_t* rowptr, const uint32_t* col, uint64_t r0, uint64_t r1, uint64_t* cnt,
                        hipStream_t stream) {
    if (r1 <= r0) return;
    k_count_upper<<<(unsigned)((r1 - r0 + 255) / 256), 256, 0, stream>>>(rowptr, col, r0, r1, cnt);
    CF_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(256) k_export_rows(const int64_t* __restrict__ rowptr, const uint32_t* __restrict__ col,
                                                     const double* __restrict__ val, uint64_t r0, uint64_t r1,
                                                     bool upper_only, const uint64_t* __restrict__ off,
                                                     uint32_t* __restrict__ ocol, double* __restrict__ oval) {
    const int lane = threadIdx.x & 63;
    const uint64_t r = r0 + (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= r1) return;
    int64_t b = rowptr[r], e = rowptr[r + 1];
    if (upper_only) {
        int64_t lo = b, hi = e;
        while (lo < hi) {
            int64_t mid = (lo + hi) >> 1;
            if (col[mid] < r) lo = mid + 1;
            else hi = mid;
        }
        b = lo;
    }
    const uint64_t o = off[r - r0];
    for (int64_t k = b + lane; k < e; k += 64) {
        ocol[o + (uint64_t)(k - b)] = col[k];
        oval[o + (uint64_t)(k - b)] = val[k];
    }
}

void launch_export_rows(const int64_t* rowptr, const uint32_t* col, const double* val, uint64_t r0, uint64_t r1,
                        bool upper_only, const uint64_t* out_off, uint32_t* ocol, double* oval, hipStream_t stream) {
    if (r1 <= r0) return;
    k_export_rows<<<(unsigned)((r1 - r0 + 3) / 4), 256, 0, stream>>>(rowptr, col, val, r0, r1, upper_only, out_off,
                                                                       ocol, oval);
    CF_HIP(hipGetLastError());
}

}  // namespace cf
