// Batched small DOCI solves, one workgroup per request (see batch.hip).
#pragma once

#include "common.h"

namespace cf {

constexpr int kBatchMaxL = 32;        // orbitals
constexpr int kBatchMaxSteps = 400;   // Lanczos steps (products) per request

// One request of a batch (device memory).
struct DociBatchJob {
    const double* h = nullptr;     // L x L (rotated) one-electron integrals
    const double* V = nullptr;     // L^4 physics two-electron integrals
    const str_t* s = nullptr;      // the n pair strings of the (L, np) space, rank order
    const uint32_t* col = nullptr; // hop columns [np * nv][n] of that space
    double tol = 1e-12;            // ||r|| <= tol |theta|
    double tol_energy = 0.0;       // Kato-Temple estimate stop (0: off)
    uint64_t seed = 0;             // start vector
    double* out = nullptr;         // eigval, residual, products, status (0 converged, 1 step limit)
    double* scratch = nullptr;     // 3 n doubles: diagonal, previous Krylov vector, w
    int L = 0, np = 0, nv = 0, n = 0;
    int check = 2;                 // convergence test every `check` steps (>= 1)
    int max_steps = kBatchMaxSteps;
};

// Largest DOCI space one workgroup solves.
int doci_batch_capacity();
// max_n: largest n in the batch (picks the workgroup shape).
void launch_doci_batch_lanczos(const DociBatchJob* d_jobs, int count, int max_n, hipStream_t s);

}  // namespace cf
