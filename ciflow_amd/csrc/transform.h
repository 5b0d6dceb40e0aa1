// Integral rotation on FP64 MFMA (see transform.hip).
#pragma once

#include "common.h"

namespace cf {

// out[b][m][a] = sum_i in[b][i][m] U[b][a][i]   (i, a < L; m < M)
void launch_quarter(int L, uint64_t M, int batch, const double* in, const double* U, double* out, hipStream_t s);
// Batched h' = U h U^T, V'_abcd = sum U_ai U_bj U_ck U_dl V_ijkl; tmp holds batch*L^4 doubles.
void rotate_integrals_dev(int L, int batch, const double* h, const double* V, const double* U, double* h_out,
                          double* V_out, double* tmp, hipStream_t s);
// W[p][q][d] = V(p,d,q,d), Jm[k][l] = V(k,l,k,l)
void launch_derived(int L, const double* V, double* W, double* Jm, hipStream_t s);

// host helpers
void gen_rotation(int L, uint64_t seed, double sigma, double* K);  // antisymmetric, K_ij ~ N(0, sigma^2)
void expm_skew(int L, const double* K, double* U);                 // U = exp(K)

}  // namespace cf
