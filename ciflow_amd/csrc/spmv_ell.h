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
    // alpha entries regrouped per alpha string (same count for every string):
    // ent[ia*Ea + q] = {segment start, Ja}, doubles q < nda, then the nsa
    // singles, then the self entry at q = nda + nsa
    const uint2* ent = nullptr;
    int nda = 0, nsa = 0;
    uint64_t R = 0;  // positions per alpha string
};

// Workgroup shape of the ELL kernel: threads (multiple of 64, <= 1024) and
// beta strings per thread (1, or 2 with paired 8/16-byte loads; 2 needs an
// even d_beta).  vec 0 selects the single-buffered kernel.
struct EllShape {
    int threads = 0;
    int vec = 1;
};

// Regroups the alpha lists into ent (see EllArgs).
void launch_ell_entries(uint64_t da, int Ea, int nda, int nsa, const Exc* fa, const uint32_t* segoff, uint2* ent,
                        hipStream_t s);
// Default shape for a beta dimension and value size (4 or 8 bytes).
EllShape ell_default_shape(uint64_t db, int value_bytes);

// jsub[u*db + ib] = sub(ib)[u].J, jfull[u*db + ib] = full(ib)[u].J (db <= 65535).
void launch_ell_index(uint64_t db, int Sb, int Eb, const Exc* subb, const Exc* fullb, uint16_t* jsub,
                      uint16_t* jfull, hipStream_t s);
// Largest beta dimension the LDS-staged kernel supports.
constexpr uint64_t kEllMaxDb = 8192;
void launch_spmv_ell_f64(const EllArgs& a, EllShape sh, const double* ell, const double* x, double* y,
                         hipStream_t s);
void launch_spmv_ell_f32(const EllArgs& a, EllShape sh, const float* ell, const double* x, double* y,
                         hipStream_t s);

}  // namespace cf
