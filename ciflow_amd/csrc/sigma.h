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
};

template <typename T>
void launch_dense_string_ham(uint64_t d, int E, const Exc* full, T* H, hipStream_t s);
size_t sigma_ab_lds_bytes(const SigmaArgs& a, bool f32);
// threads: workgroup size of the alpha-beta kernel (128, 256, 512, 768 or 1024);
// a workgroup owns one alpha string and `threads` beta strings.
template <typename CT>
void launch_sigma_ab(const SigmaArgs& a, int threads, const CT* C, double* sig, hipStream_t s);
void launch_to_f32(uint64_t n, const double* x, float* y, hipStream_t s);
void launch_to_f64(uint64_t n, const float* x, double* y, hipStream_t s);
void launch_row_dots(uint64_t da, uint64_t db, int S, const Exc* sub, const double* C, double* out, hipStream_t s);
void launch_transpose(uint64_t rows, uint64_t cols, const double* in, double* out, hipStream_t s);

}  // namespace cf
