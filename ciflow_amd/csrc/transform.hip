// Orbital rotation of the integrals on FP64 MFMA (replaces
// OrbitalTransform::fillHamCI / buildOneBodyMatrixElements,
// OrbitalTransform.cpp:64-238, and the exp(X) of UnitaryMatrix::updateUnitary,
// UnitaryMatrix.cpp:229-340).
//
//   h'_ab   = sum_ij U_ai U_bj h_ij
//   V'_abcd = sum_ijkl U_ai U_bj U_ck U_dl V_ijkl      (physics V; U_ai = new orbital a in old i)
//
// The reference does four quarter-transform DGEMMs per irrep block.  Here each
// quarter transform contracts the LEADING index and appends the new index at
// the end, out[(jkl), a] = sum_i in[i, (jkl)] U[a, i], so four identical
// passes return V' in its natural (a,b,c,d) order and every pass is one
// GEMM (L^3 x L) = (L x L^3)^T (L x L)^T, batched over independent
// Hamiltonians.  8 L^5 flop per Hamiltonian (C1 symmetry), on
// v_mfma_f64_16x16x4f64 with K = L in steps of 4.
#include "transform.h"

namespace cf {

typedef double v4d __attribute__((ext_vector_type(4)));

// in:  batch x [K][M]   (leading index contracted), K = L
// U:   batch x [L][L]   U[a*L + i]
// out: batch x [M][L]
// One wavefront per MT x 16 columns m computes the 16 x 16 output tiles of all
// NA = ceil(L / 16) blocks of the new index: its L x (16 MT) input block is
// read once (every k-step's loads issued before the MFMAs) and serves every
// tile.  Each tile accumulates its k-steps in order 0, 4, 8, ... (KS k-steps
// of registers, KS >= ceil(L / 4)).
template <int NA, int KS, int MT>
__global__ void __launch_bounds__(64) k_quarter_mfma(int L, uint64_t M, const double* __restrict__ in,
                                                    const double* __restrict__ U, double* __restrict__ out) {
    const uint64_t b = blockIdx.z;
    const uint64_t m0 = (uint64_t)blockIdx.x * 16 * MT;
    const int lane = threadIdx.x;
    const int r = lane & 15, kk = lane >> 4;
    const double* inb = in + b * (uint64_t)L * M;
    const double* Ub = U + b * (uint64_t)L * L;
    double* ob = out + b * M * (uint64_t)L;
    const int ks = (L + 3) / 4;
    double av[MT][KS];  // A[m][k] for m = m0 + 16 u + r, k = 4 s + kk
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        const int k = 4 * s + kk;
#pragma unroll
        for (int u = 0; u < MT; ++u) {
            const uint64_t m = m0 + 16 * u + r;
            av[u][s] = (s < ks && k < L && m < M) ? inb[(uint64_t)k * M + m] : 0.0;
        }
    }
    v4d acc[MT][NA];
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int t = 0; t < NA; ++t) acc[u][t] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < KS; ++s) {
        if (s < ks) {
            const int k = 4 * s + kk;
#pragma unroll
            for (int t = 0; t < NA; ++t) {
                const int a = 16 * t + r;
                const double bv = (k < L && a < L) ? Ub[(uint64_t)a * L + k] : 0.0;  // B[k][a]
#pragma unroll
                for (int u = 0; u < MT; ++u)
                    acc[u][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u][s], bv, acc[u][t], 0, 0, 0);
            }
        }
    }
    // f64 16x16x4 accumulator layout: D[i][j] sits in lane j + 16*(i%4),
    // element i/4 (row = 4*element + lane/16, column = lane%16)
#pragma unroll
    for (int u = 0; u < MT; ++u)
#pragma unroll
        for (int t = 0; t < NA; ++t)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint64_t m = m0 + 16 * u + 4 * v + kk;
                const int a = 16 * t + r;
                if (m < M && a < L) ob[m * L + a] = acc[u][t][v];
            }
}

void launch_quarter(int L, uint64_t M, int batch, const double* in, const double* U, double* out, hipStream_t s) {
    if (batch <= 0) return;
    if (L < 1 || L > kMaxL) throw Error(ERR_ARG, "rotation: L out of range");
    constexpr int MT = 4;
    const dim3 grid((unsigned)((M + 16 * MT - 1) / (16 * MT)), 1, (unsigned)batch);
    if (L <= 16) k_quarter_mfma<1, 4, MT><<<grid, 64, 0, s>>>(L, M, in, U, out);
    else if (L <= 32) k_quarter_mfma<2, 8, MT><<<grid, 64, 0, s>>>(L, M, in, U, out);
    else if (L <= 48) k_quarter_mfma<3, 12, MT><<<grid, 64, 0, s>>>(L, M, in, U, out);
    else k_quarter_mfma<4, 16, MT><<<grid, 64, 0, s>>>(L, M, in, U, out);
    CF_HIP(hipGetLastError());
}

void rotate_integrals_dev(int L, int batch, const double* h, const double* V, const double* U, double* h_out,
                          double* V_out, double* tmp, hipStream_t s) {
    const uint64_t L2 = (uint64_t)L * L, L3 = L2 * L;
    // one-body: (i, j) -> (j, a) -> (a, b)
    launch_quarter(L, L, batch, h, U, tmp, s);
    launch_quarter(L, L, batch, tmp, U, h_out, s);
    // two-body: four cyclic passes, V -> tmp -> V_out -> tmp -> V_out
    launch_quarter(L, L3, batch, V, U, tmp, s);
    launch_quarter(L, L3, batch, tmp, U, V_out, s);
    launch_quarter(L, L3, batch, V_out, U, tmp, s);
    launch_quarter(L, L3, batch, tmp, U, V_out, s);
}

// Derived tables of a Hamiltonian already on the device.
__global__ void k_derived(int L, const double* __restrict__ V, double* __restrict__ W, double* __restrict__ Jm) {
    const uint64_t L2 = (uint64_t)L * L, L3 = L2 * L;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < L3) {
        const uint64_t p = t / L2, q = (t / L) % L, d = t % L;
        W[t] = V[((p * L + d) * L + q) * L + d];
    }
    if (t < L2) {
        const uint64_t k = t / L, l = t % L;
        Jm[t] = V[((k * L + l) * L + k) * L + l];
    }
}

void launch_derived(int L, const double* V, double* W, double* Jm, hipStream_t s) {
    const uint64_t L3 = (uint64_t)L * L * L;
    k_derived<<<(unsigned)((L3 + 255) / 256), 256, 0, s>>>(L, V, W, Jm);
    CF_HIP(hipGetLastError());
}

}  // namespace cf
