// Device data structures and kernel launchers of the ciflow_amd library.
#pragma once

#include "common.h"

namespace cf {

// One entry of a string's excitation list, sorted by target rank J.
//  type 0: the string itself          x = one-string diagonal part
//  type 1: single, orbital f -> t     x = h_pq + sum_z (V_pzqz - V_pzzq), sign = (-1)^(stacked)
//  type 2: same-spin double           x = signed matrix element (canonical orientation)
struct Exc {
    uint32_t J;
    uint8_t type;
    int8_t sign;
    uint8_t f;
    uint8_t t;
    double x;
};
static_assert(sizeof(Exc) == 16, "Exc must stay 16 bytes");

// Device-side view of a Hamiltonian.
struct HamDev {
    int L;
    const double* h;   // L*L
    const double* V;   // L^4 physics, V[((i*L+j)*L+k)*L+l]
    const double* W;   // L^3, W[(p*L+q)*L+d] = V(p,d,q,d)
    const double* Jm;  // L^2, Jm[k*L+l] = V(k,l,k,l)
};

// launchers (all asynchronous on `stream`)
void launch_unrank(int L, int nel, uint64_t first, uint64_t count, const uint64_t* binom, str_t* out,
                   hipStream_t stream);
void launch_rank(int L, int nel, const str_t* strings, uint64_t count, const uint64_t* binom, uint32_t* out,
                 hipStream_t stream);
void launch_diag1(int nel, const str_t* strs, uint64_t d, HamDev H, double* out, hipStream_t stream);
// Excitation lists of all strings: full (E per string) and self+singles (S per string).
void launch_exc_lists(int L, int nel, const str_t* strs, uint64_t d, const uint64_t* binom, HamDev H,
                      const double* diag1, int E, int S, Exc* full, Exc* sub, hipStream_t stream);
// Row-layout offsets of the FCI rows of each alpha string: segoff[ia*(Ea+1)+e].
void launch_fci_segoff(uint64_t da, int Ea, const Exc* full_a, int Eb, int Sb, uint32_t* segoff,
                       hipStream_t stream);
void launch_fci_diag(uint64_t da, uint64_t db, const str_t* sa, const str_t* sb, const double* d1a,
                     const double* d1b, HamDev H, double* diag, hipStream_t stream);
void launch_doci_diag(uint64_t d, const str_t* s, const double* d1, HamDev H, double* diag, hipStream_t stream);

struct FciRowsArgs {
    uint64_t da, db;
    int Ea, Sa, Eb, Sb;
    const str_t* sa;
    const str_t* sb;
    const Exc* fa;   // full alpha lists
    const Exc* fb;   // full beta lists
    const Exc* subb; // beta self+singles lists
    const uint32_t* segoff;
    const double* diag;
    HamDev H;
    double zero;
    uint64_t row0, nrows;  // row range handled by this launch
    // count pass
    uint32_t* rowcnt;
    // fill pass
    const int64_t* rowptr;
    uint32_t* col;
    double* val;
    float* val32;
};
void launch_fci_rows(const FciRowsArgs& a, bool fill, hipStream_t stream);
// Position-major (per alpha string) copy of the complete rows, zeros where
// |v| <= zero; either output may be null.
void launch_fci_ell(const FciRowsArgs& a, double* ell, float* ell32, hipStream_t stream);

struct DociRowsArgs {
    uint64_t d;
    int L, np;
    const str_t* s;
    const uint64_t* binom;
    const double* diag;
    HamDev H;
    double zero;
    uint64_t row0, nrows;
    uint32_t* rowcnt;
    const int64_t* rowptr;
    uint32_t* col;
    double* val;
    float* val32;
};
void launch_doci_rows(const DociRowsArgs& a, bool fill, hipStream_t stream);

// Selected-CI rows over a determinant list (FCI_File::construct_CI_matrix,
// CIMethod.cpp:890-912): row i in list order, full rows, columns ascending.
struct DetRowsArgs {
    uint64_t n;
    int nup, ndown;
    const str_t* up;
    const str_t* down;
    HamDev H;
    double zero;
    uint32_t* rowcnt;
    const int64_t* rowptr;
    uint32_t* col;
    double* val;
    float* val32;
};
void launch_det_rows(const DetRowsArgs& a, bool fill, hipStream_t stream);
// The same rows by excitation lookup in the sorted keys (L <= 32); the fill
// writes each row's entries unordered (sort_rows_by_column orders them).
struct DetLookupArgs {
    DetRowsArgs r;                  // n, nup, ndown, up, down, H, zero, rowcnt / rowptr, col, val (val32 unused)
    const uint64_t* skey = nullptr;  // sorted keys up << 32 | down
    const uint32_t* sidx = nullptr;  // list index of each sorted key
    int L = 0;
};
void launch_det_rows_lookup(const DetLookupArgs& a, bool fill, hipStream_t stream);
// y = H x computed on the fly from the sorted keys (all entries but those of
// the leading m x m block; rows < m add to y, the others overwrite it).
void launch_det_sigma_lookup(const DetLookupArgs& a, const double* x, double* y, uint64_t m, hipStream_t stream);
// bits: significant key bits (32 + L)
void sort_det_keys(uint64_t n, const str_t* up, const str_t* down, uint64_t* skey, uint32_t* sidx, int bits,
                   hipStream_t stream);
void sort_rows_by_column(uint64_t n, uint64_t nnz, const int64_t* rowptr, uint32_t* col, double* val,
                         hipStream_t stream);

// Exclusive scan of 32-bit counts into 64-bit offsets (n+1 entries).
void exclusive_scan_u32_to_i64(const uint32_t* cnt, int64_t* out, uint64_t n, hipStream_t stream);

// Upper-triangle / full export of a row range.
void launch_count_upper(const int64_t* rowptr, const uint32_t* col, uint64_t r0, uint64_t r1, uint64_t* cnt,
                        hipStream_t stream);
void launch_export_rows(const int64_t* rowptr, const uint32_t* col, const double* val, uint64_t r0, uint64_t r1,
                        bool upper_only, const uint64_t* out_off, uint32_t* ocol, double* oval, hipStream_t stream);

// SpMV y = A x on full rows.
void launch_spmv_f64(uint64_t n, const int64_t* rowptr, const uint32_t* col, const double* val, const double* x,
                     double* y, int lanes_per_row, hipStream_t stream);
void launch_spmv_f32(uint64_t n, const int64_t* rowptr, const uint32_t* col, const float* val, const double* x,
                     double* y, int lanes_per_row, hipStream_t stream);
// Vectorised-stream row gather: lanes_per_row 4, 8, 16 or 32, unroll 1 or 2.
void launch_spmv_v4_f64(uint64_t n, const int64_t* rowptr, const uint32_t* col, const double* val, const double* x,
                        double* y, int lanes_per_row, int unroll, hipStream_t stream);
void launch_spmv_v4_f32(uint64_t n, const int64_t* rowptr, const uint32_t* col, const float* val, const double* x,
                        double* y, int lanes_per_row, int unroll, hipStream_t stream);
// Sliced row gather over the chunked copy (64 rows per chunk, quad-interleaved
// entries; qoff = nchunk + 1 exclusive quad offsets): waves per chunk 1, 2, 4,
// 8 or 16, unroll 1 or 2 (4 with 4, 8 or 16 waves).  A launch with few
// chunks splits each into up to kSellMaxSplit quad ranges whose sums go
// through part (part_cap doubles; splits needing more than part_cap are not
// used) and a second pass.
constexpr uint32_t kSellMaxSplit = 4;
void launch_spmv_sell_f64(uint64_t n, const int64_t* qoff, const uint32_t* col, const double* val, const double* x,
                          double* y, double* part, uint64_t part_cap, int waves, int unroll, hipStream_t stream);
void launch_spmv_sell_f32(uint64_t n, const int64_t* qoff, const uint32_t* col, const float* val, const double* x,
                          double* y, double* part, uint64_t part_cap, int waves, int unroll, hipStream_t stream);
// Chunked copy of CRS rows: quads of each chunk's longest row, then the fill
// (val32 / sval32 may be NULL).
void launch_sell_quads(uint64_t n, const int64_t* rowptr, uint32_t* quads, hipStream_t stream);
void launch_sell_fill(uint64_t n, const int64_t* rowptr, const uint32_t* col, const double* val, const float* val32,
                      const int64_t* qoff, uint32_t* scol, double* sval, float* sval32, hipStream_t stream);

// Krylov vector kernels (vectors are columns of a column-major n x k block).
void launch_multi_dot(uint64_t n, int k, const double* Vb, const double* w, double* partial, int nblocks,
                      hipStream_t stream);
void launch_reduce_partials(int k, int nblocks, const double* partial, double* out, hipStream_t stream);
int multi_dot_blocks(uint64_t n);
// Davidson basis limit and the Ritz coefficients, passed to the kernels by
// value (host y, k <= kDavMaxBasis).
constexpr int kDavMaxBasis = 24;
struct DavCoeffs {
    double c[kDavMaxBasis];
};
// dst = src / sqrt(*norm2), the squared norm in device memory (zeros when it
// is <= 1e-28)
void launch_scale_rnorm(uint64_t n, const double* src, const double* norm2, double* dst, hipStream_t stream);
// Fused Davidson step for n <= kFusedMaxN: x, hx, t and per-block |r|^2 parts
// (fused_axpy_blocks(n) of them) in one pass; then, after two multi_axpy_r
// passes over t, dst = t / |t| with out[0] = |r|^2 and out[1] = |t|^2.
void launch_combine_residual(uint64_t n, int k, const double* V, const double* S, const double* y, double theta,
                             double shift_min, const double* diag, double* x, double* hx, double* t, double* rparts,
                             double* tparts, hipStream_t stream);
// (with tparts: |t|^2 parts before the first pass; pnorm2 the parts after a
// selective second pass, used unless the first pass sufficed)
void launch_scale_rnorm_parts(uint64_t n, const double* src, const double* pnorm, const double* rparts, double* dst,
                              double* out, hipStream_t stream, const double* tparts = nullptr,
                              const double* pnorm2 = nullptr, double eta = 0.0);
// Large-n Davidson step in two basis passes (nblocks = multi_dot_blocks(n),
// partial of 64 rows x nblocks): x = V y, r = S y - theta x and the correction
// t = r / (theta - diag) (|theta - diag| clamped to >= shift_min; t = r for
// shift_min 0), with parts of V^T t in rows 0..k-1, |t|^2 in row_t and
// |r|^2 in row_r; then t -= V c with parts of V^T t (rows 0..k-1) and |t|^2
// (row_t) of the updated t.
void launch_combine_res_dot(uint64_t n, int k, const double* V, const double* S, const double* y, double theta,
                            double shift_min, const double* diag, double* x, double* t, double* partial, int nblocks,
                            int row_t, int row_r, hipStream_t stream);
void launch_axpy_dot(uint64_t n, int k, const double* V, const double* c, double* t, double* partial, int nblocks,
                     int row_t, hipStream_t stream);

// w -= Vb[:, 0:k] c; with gate_before/gate_after (|w|^2 before and after a
// first Gram-Schmidt pass, device memory) skipped when after > eta^2 before.
void launch_multi_axpy(uint64_t n, int k, const double* Vb, const double* c, double* w, hipStream_t stream,
                       const double* gate_before = nullptr, const double* gate_after = nullptr, double eta = 0.0);
void launch_local3(uint64_t n, const double* a, const double* v1, const double* b, const double* v0, double* w,
                   hipStream_t stream);
void launch_finish_step(uint64_t n, int j, const double* c1, const double* c2, const double* wnorm2,
                        const double* w, double* vnext, double* alpha, double* beta, hipStream_t stream);
void launch_scale_copy(uint64_t n, const double* src, double s, double* dst, hipStream_t stream);

// Step kernels that reduce the partials of the preceding product themselves
// (every workgroup sums them in the same fixed order), so a Lanczos step needs
// no separate reduction launches.  For n <= kFusedMaxN: the product runs with
// fused_dot_blocks(n) workgroups of 1024 threads (a power of two <= 64).
constexpr uint64_t kFusedMaxN = 262144;
int fused_dot_blocks(uint64_t n);
int fused_axpy_blocks(uint64_t n);
// With skip_before/skip_after (per-block |.|^2 parts, fused_axpy_blocks(n) of
// each): a selective second Gram-Schmidt pass, skipped inside the kernel when
// sum(after) > eta^2 sum(before).
void launch_multi_dot_wide(uint64_t n, int k, const double* Vb, const double* w, double* partial, int nblocks,
                           hipStream_t stream, const double* skip_before = nullptr,
                           const double* skip_after = nullptr, double eta = 0.0);
// w -= a v1 + b v0, a = sum of the nblocks partials (written to a_out), b = *b_in
void launch_local3_r(uint64_t n, int nblocks, const double* partial, const double* v1, const double* b_in,
                     const double* v0, double* w, double* a_out, hipStream_t stream);
// w -= Vb[:, 0:k] c, c_i = sum of partial[i*nblocks + 0:nblocks] (written to c_out);
// pnorm[block] = that workgroup's part of |w|^2 after the update (fused_axpy_blocks(n) workgroups)
void launch_multi_axpy_r(uint64_t n, int k, int nblocks, const double* partial, const double* Vb, double* w,
                         double* c_out, double* pnorm, hipStream_t stream, const double* skip_before = nullptr,
                         const double* skip_after = nullptr, double eta = 0.0);
// finish_step with |w|^2 = the sum of the fused_axpy_blocks(n) entries of pnorm
void launch_finish_step_r(uint64_t n, int j, const double* c1, const double* c2, const double* pnorm,
                          const double* w, double* vnext, double* alpha, double* beta, hipStream_t stream);
void launch_rotate_basis(uint64_t n, int m, int k, double* Vb, const double* Y, hipStream_t stream);
void launch_fill_start(uint64_t n, uint64_t seed, double* v, hipStream_t stream);
void launch_copy(uint64_t n, const double* src, double* dst, hipStream_t stream);

// ---------------------------------------------------------------------------
// Batched Davidson (davbatch.hip): `lanes` requests of one shape (same n)
// advanced in lock step, one launch per step for all of them.
// ---------------------------------------------------------------------------
constexpr int kBdMaxLanes = 32;
constexpr int kBdPartial = 64 * 64;  // per lane: dot partials, rows x fused_dot_blocks(n)
constexpr int kBdParts = 5 * 256;    // per lane: pnorm, pnorm2, rparts, tparts, r1parts
enum BdMode : int { BD_IDLE = 0, BD_RITZ = 1, BD_START_X = 2, BD_START_SEED = 3 };
// One lane's part of an iteration, written by the host before each one.
struct BdStep {
    double ys[kDavMaxBasis];  // BD_RITZ: coefficients of the lowest Ritz vector
    double y1[kDavMaxBasis];  // second Ritz vector (second-pair residual)
    double theta, theta1, shift;
    uint64_t seed;            // BD_START_SEED
    int mode;
    int m;        // basis vectors in use; the step forms and multiplies vector m
    int f32;      // the product takes FP32 operand values
    int restart;  // first rotate V and S to their `keep` Ritz vectors (Y: mr x mr)
    int mr, keep;
    int r2;       // second-pair residual requested
    int pad;
};
// Per-lane operator data, fixed for a batch.
struct BdOpLane {
    const double* diag = nullptr;  // preconditioner diagonal
    const double* Ha = nullptr;    // FCI: dense alpha string Hamiltonian (da x da)
    const double* Hb = nullptr;    // dense beta string Hamiltonian (db x db)
    const double* vk64 = nullptr;  // DOCI: pair table
    const float* vk32 = nullptr;
};
// The batch's buffers: lane l's vectors at a fixed stride from the bases.
struct BdView {
    uint64_t n = 0;
    int M = 0;           // basis vectors per lane
    int Ga = 0, Gd = 0;  // fused_axpy_blocks(n), fused_dot_blocks(n)
    double* V = nullptr;  // [lanes][M][n]
    double* S = nullptr;  // [lanes][M][n] images H V
    double* x = nullptr;  // [lanes][n] Ritz vector
    double* t = nullptr;  // correction / new vector before normalisation
    float* x32 = nullptr;  // [lanes][n] FP32 copy of the new vector (or null)
    double* partial = nullptr;  // [lanes][kBdPartial]
    double* parts = nullptr;    // [lanes][kBdParts]
    double* dg = nullptr;       // [lanes][64]: rows 0..m dots V^T S_m, 62 |r|^2, 63 |t|^2
    double* Y = nullptr;        // [lanes][kDavMaxBasis^2] restart rotations
    const BdStep* step = nullptr;  // [lanes]
    const BdOpLane* op = nullptr;  // [lanes]
};
// Restart rotations, then the new vector of every lane: x = V y, r = S y -
// theta x, t = r / (theta - diag), two selective Gram-Schmidt passes, v_m =
// t / |t| (and its FP32 copy when x32 is set and the lane's product is FP32).
// Every lane's sums run in a fixed order that depends on n alone.
void bd_form_vectors(const BdView& b, int lanes, bool any_restart, hipStream_t s);
// dg rows 0..m: V^T S_m after the product.
void bd_new_column(const BdView& b, int lanes, hipStream_t s);
// r1parts of |S y1 - theta1 V y1|^2 for lanes with r2 set.
void bd_resid2(const BdView& b, int lanes, hipStream_t s);
// FCI same-spin part S_m = Ha C + C Hb (C = V_m as a row-major da x db matrix),
// FP64 MFMA; overwrites S_m (the alpha-beta kernel adds to it).
void bd_same_spin(const BdView& b, int lanes, uint64_t da, uint64_t db, hipStream_t s);

}  // namespace cf
