// FCI-structured SpMV with LDS-staged x segments (see spmv_fci.hip).
#pragma once

#include "kernels.h"

namespace cf {

struct FciSpmvArgs {
    uint64_t db = 0;
    int R = 0, W = 16, NS = 4, nd = 0, nd_max = 1, nsd_max = 1, nchunk = 0, Ea = 0, nsd = 0;
    uint32_t tiles_per_alpha = 0, ntiles = 0;
    const Exc* alist = nullptr;            // [da][Ea] alpha excitation lists
    const uint32_t* segoff = nullptr;      // [da][Ea+1] row segment offsets
    const int64_t* rowptr = nullptr;
    const uint32_t* col = nullptr;
    const int32_t* desc = nullptr;         // [da][R]   LDS offset descriptor per row position
    const uint32_t* sd_J = nullptr;        // [da][nsd] alpha target of each self/single segment
    const uint16_t* chunk_first = nullptr; // [da][nchunk+1] first self/single ordinal starting in chunk c
    const uint32_t* dbl_J = nullptr;       // [da][nd]  alpha target of each double
};

struct FciSpmvPlan {
    FciSpmvArgs a;
    DBuf<int32_t> desc;
    DBuf<uint32_t> sdJ, dJ;
    DBuf<uint16_t> cfirst;
    // R: full row length, nd: doubles per alpha string, Sb: beta self+singles list length
    void init(uint64_t da, uint64_t db, int Ea, const Exc* full_a, const uint32_t* segoff, int R, int nd, int Sb,
              int W, hipStream_t stream);
};

size_t fci_spmv_lds_bytes(const FciSpmvArgs& a, int W);
void launch_spmv_fci_f64(const FciSpmvArgs& a, const double* val, const double* x, double* y, hipStream_t stream);
void launch_spmv_fci_f32(const FciSpmvArgs& a, const float* val, const double* x, double* y, hipStream_t stream);

}  // namespace cf
