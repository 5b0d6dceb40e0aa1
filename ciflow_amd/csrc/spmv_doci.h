// DOCI SpMV on position-major pair-hop columns (see spmv_doci.hip).
#pragma once

#include "kernels.h"

namespace cf {

// Every DOCI row has the same np * nv pair hops (np pairs, nv = L - np empty
// orbitals), and the element of a hop j -> k depends on (j, k) only.  Slot
// t = a * nv + b of a row is the hop of its a-th occupied orbital to its b-th
// empty orbital, both counted from the top; col[t * n + i] is the rank of the
// target string of row i.  Values come from an L x L pair table in LDS.
struct DociEllArgs {
    uint64_t n = 0;
    int L = 0, np = 0, nv = 0;
    const str_t* s = nullptr;        // row strings (rank order)
    const uint32_t* col = nullptr;   // [np * nv][n]
    const double* diag = nullptr;    // diagonal elements
    double zero = 0.0;               // elements with |value| <= zero count as absent
};

// Workgroup size (64..1024, multiple of 64) and lanes per row (1, 2, 4, 8).
struct DociShape {
    int threads = 256;
    int sub = 1;
};
// Lanes per row that fill the GPU for n rows (256-thread workgroups).
DociShape doci_default_shape(uint64_t n);

// Block-staged DOCI product (no stored columns).  Orbitals split into w low
// and L - w high ones; the rows whose strings share the high part `hi` (and so
// hold c = np - popcount(hi) low pairs) are one contiguous rank range,
// base(hi, c) + rank_lo(lo), a "block".  A workgroup owns a block and stages
// in LDS its own x block and, one at a time, the x blocks that the hops
// crossing the split reach; hops inside the high part read the neighbour
// block with coalesced loads.
struct DociBlock {
    uint32_t hi;    // high part (bit q = orbital w + q)
    uint32_t base;  // rank of its first row
    uint32_t c;     // low pairs
    uint32_t R;     // rows, C(w, c)
};
struct DociBlockArgs {
    uint64_t n = 0;
    int L = 0, np = 0, w = 0;
    const DociBlock* blocks = nullptr;
    uint32_t nblocks = 0;
    const uint16_t* rank_lo = nullptr;  // [2^w]: rank of a low pattern among those of its popcount
    const uint16_t* lo_of = nullptr;    // [2^w]: low patterns by popcount, rank order
    const uint32_t* lo_off = nullptr;   // [w + 2]: start of each popcount in lo_of
    const uint64_t* binom = nullptr;
    const double* diag = nullptr;
    double zero = 0.0;
};
// The lanes of a batched Davidson product (block-staged kernel): lane l
// (grid row l) takes its pair table and diagonal from op[l] and multiplies
// the new vector of its step, V_m, into S_m; lanes whose step is idle or of
// the other value precision do nothing.
struct DociBatch {
    const BdStep* step = nullptr;
    const BdOpLane* op = nullptr;
    const double* V = nullptr;
    double* S = nullptr;
    uint64_t n = 0;
    int M = 0;
};
// Largest low part: C(12, 6) = 924 rows per block, 4 per thread of 256.
constexpr int kDociMaxW = 12;
// Low orbitals for L orbitals (the block shape the kernel is run with by default).
int doci_block_w(int L);
void launch_spmv_doci_blk_f64(const DociBlockArgs& a, const double* vk, const double* x, double* y, hipStream_t s);
void launch_spmv_doci_blk_f32(const DociBlockArgs& a, const float* vk, const double* x, double* y, hipStream_t s);
// a: the shared block tables of the shape (its diag is not used).
void launch_spmv_doci_blk_batch(const DociBlockArgs& a, bool f32, const DociBatch& bt, int lanes, hipStream_t s);

// col of DociEllArgs from the row strings (binom: the rank table).
void launch_doci_ell_fill(const DociEllArgs& a, const uint64_t* binom, uint32_t* col, hipStream_t s);
// Pair table: vk[j * L + k] = V(m, m, M, M) with m = min(j, k), M = max(j, k),
// zeroed where |value| <= zero, as the CRS build drops such elements.
void launch_doci_pair_table(const DociEllArgs& a, HamDev H, double* vk64, float* vk32, hipStream_t s);
void launch_spmv_doci_f64(const DociEllArgs& a, DociShape sh, const double* vk, const double* x, double* y,
                          hipStream_t s);
void launch_spmv_doci_f32(const DociEllArgs& a, DociShape sh, const float* vk, const double* x, double* y,
                          hipStream_t s);

}  // namespace cf
