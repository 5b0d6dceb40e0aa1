// FCI SpMV on position-major complete rows with LDS-staged x blocks (see spmv_ell.hip).
#pragma once

#include "kernels.h"

namespace cf {

struct EllArgs {
    uint64_t da = 0, db = 0;
    int Ea = 0, Sb = 0, Eb = 0;
    const Exc* fa = nullptr;          // full alpha lists [da][Ea]
    const uint32_t* segoff = nullptr;  // [da][Ea+1]
    const uint16_t* jsub = nullptr;   // [Sb][db]: beta self+singles targets, target-sorted
    const uint16_t* jfull = nullptr;  // [Eb][db]: full beta list targets
};

// jsub[u*db + ib] = sub(ib)[u].J, jfull[u*db + ib] = full(ib)[u].J (db <= 65535).
void launch_ell_index(uint64_t db, int Sb, int Eb, const Exc* subb, const Exc* fullb, uint16_t* jsub,
                      uint16_t* jfull, hipStream_t s);
// Largest beta dimension the LDS-staged kernel supports.
constexpr uint64_t kEllMaxDb = 8192;
void launch_spmv_ell_f64(const EllArgs& a, int threads, const double* ell, const double* x, double* y,
                         hipStream_t s);
void launch_spmv_ell_f32(const EllArgs& a, int threads, const float* ell, const double* x, double* y,
                         hipStream_t s);

}  // namespace cf
