// 1- and 2-RDMs on the GPU (replaces CIDens::construct_density and the
// DensFCI / DensDOCI builders, CIDens.cpp:323-334, 533-673, 1073-1224,
// 1530-1631).
#pragma once

#include "kernels.h"

namespace cf {

// Dense RDMs of the CI vector c (host, n entries):
//   gamma[s][p][q]        = <a+_ps a_qs>                     (host, 2*L*L)
//   Gamma[b][i][j][k][l]  = <a+_is a+_jt a_lt a_ks>, b: 0 aa, 1 ab, 2 bb  (host, 3*L^4, or null)
void rdm_dense(hipStream_t stream, int kind, int L, int nup, int ndown, uint64_t da, uint64_t db, const str_t* sa,
               const str_t* sb, const Exc* suba, const Exc* subb, int Sa, int Sb, const uint64_t* binom,
               const double* c_host, uint64_t n, double* gamma, double* Gamma);

// Dense RDMs of a determinant-list space (DensFILE, CIDens.cpp:1468-1526):
// up/down are the device strings of the n list determinants (list order),
// c_host the CI vector.  L <= 32; the determinants must be distinct.
void rdm_dets(hipStream_t stream, int L, uint64_t n, const str_t* up, const str_t* down, const double* c_host,
              double* gamma, double* Gamma);

// Pack dense RDMs into the reference storage (CIDens.cpp:166-288,
// matrix.cpp:674-689): one[2][L(L+1)/2], two[3][L^2(L^2+1)/2] (two may be null).
void pack_rdm(int L, const double* gamma, const double* Gamma, double* one, double* two);

}  // namespace cf
