// On-the-fly FCI sigma (see sigma.hip).
#pragma once

#include "kernels.h"

namespace cf {

struct SigmaArgs {
    uint64_t da = 0, db = 0;
    int L = 0, Sa = 0, Sb = 0;
    uint32_t nchunk = 0, nblocks = 0;
    const str_t* sa = nullptr;
    const str_t* sb = nullptr;
    const Exc* suba = nullptr;  // [da][Sa] self + singles
    const Exc* subb = nullptr;  // [db][Sb]
    const double* V = nullptr;  // L^4
    // slice tables of the prefetching kernel, per staged precision:
    //   vpair[(p*L + r)*L^2 + q*L + s] = V[p][q][r][s]    (alpha single p -> r)
    //   vself[Ia*L^2 + q*L + s] = sum_{p in Ia} V[p][q][p][s]  (alpha number operators)
    const double* vpair64 = nullptr;
    const double* vself64 = nullptr;
    const float* vpair32 = nullptr;
    const float* vself32 = nullptr;
};

// Alpha-beta kernel choice: 1 = stage the V slice and C row per alpha action
// (gathers V from L^4), 2 = slice tables with the next action's V slice and C
// row prefetched into registers (double-buffered LDS up to 8 row elements per
// thread), 3 = kernel 2's scheme with packed LDS offsets and exact list
// lengths (its own workgroup sizes; shapes without an instance run kernel 2).
struct SigmaShape {
    int threads = 512;
    int kernel = 3;
};

// The lanes of a batched Davidson product (kernel 3): lane l (grid row l)
// runs on args[l] with the new vector of its step as operand (V_m, or its
// FP32 copy in x32) and adds into S_m; lanes whose step is idle or of the
// other operand precision do nothing.
struct SigmaBatch {
    const SigmaArgs* args = nullptr;  // [lanes], nchunk / nblocks from sigma_ab3_batch_args
    const BdStep* step = nullptr;
    const double* V = nullptr;
    double* S = nullptr;
    const float* x32 = nullptr;
    uint64_t n = 0;
    int M = 0;
};

size_t sigma_ab_lds_bytes(const SigmaArgs& a, bool f32);
// threads: workgroup size of the alpha-beta kernel (128, 256, 512, 768 or 1024);
// a workgroup owns one alpha string and `threads` beta strings.
template <typename CT>
void launch_sigma_ab(const SigmaArgs& a, int threads, const CT* C, double* sig, hipStream_t s);
// Builds vpair / vself (CT precision) from V and the alpha strings.
template <typename CT>
void launch_sigma_tables(const SigmaArgs& a, CT* vpair, CT* vself, hipStream_t s);
// Kernel 2 (falls back to kernel 1 where the prefetch registers cannot hold a row).
template <typename CT>
void launch_sigma_ab2(const SigmaArgs& a, SigmaShape sh, const CT* C, double* sig, hipStream_t s);
size_t sigma_ab2_lds_bytes(const SigmaArgs& a, bool f32, bool dbuf);
// Kernel 3 (packed LDS offsets, negated slice copy, exact list lengths); false
// when the space's shape has no instance (the caller runs kernel 2).
// nb_electrons: beta electrons (the number operators of every beta string).
template <typename CT>
bool launch_sigma_ab3(const SigmaArgs& a, int nb_electrons, const CT* C, double* sig, hipStream_t s);
// Sets a.nchunk / a.nblocks of kernel 3's instance for this shape and operand
// precision; false when there is none.
bool sigma_ab3_batch_args(SigmaArgs& a, int nb_electrons, bool f32);
// shape: any lane's args (the lanes share da, db, L and the list lengths).
template <typename CT>
void launch_sigma_ab3_batch(const SigmaArgs& shape, int nb_electrons, const SigmaBatch& sb, int lanes, hipStream_t s);
template <typename T>
void launch_dense_string_ham(uint64_t d, int E, const Exc* full, T* H, hipStream_t s);
void launch_to_f32(uint64_t n, const double* x, float* y, hipStream_t s);
void launch_to_f64(uint64_t n, const float* x, double* y, hipStream_t s);
void launch_row_dots(uint64_t da, uint64_t db, int S, const Exc* sub, const double* C, double* out, hipStream_t s);
void launch_transpose(uint64_t rows, uint64_t cols, const double* in, double* out, hipStream_t s);

}  // namespace cf
