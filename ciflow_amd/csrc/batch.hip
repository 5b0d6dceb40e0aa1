// Many small DOCI ground-state solves in one launch: the served stream's
// small requests (DOCI spaces of up to 13,312 determinants, L <= 16 at half
// filling), which the reference solves one after another
// (OrbitalOptimization::update_cim -> set_ham -> build_ci -> solve,
// SimulatedAnnealing.cpp:83-87, CIMethod.cpp:117-126,241-268).  Solved one
// request at a time they are bound by their chain of about 600 short kernel
// launches, not by arithmetic; here one workgroup runs a whole solve on one
// CU and a launch carries every request of a batch, of any mix of sizes.
//
// Per workgroup: the DOCI matrix of DOCI::construct_CI_matrix
// (CIMethod.cpp:611-660) applied without storing it -- the diagonal
// 2 sum h_ii + sum_i (ii|ii) + sum_{k != l} (2 (kk|ll) - (kl|lk)) per row once,
// and the pair hops j -> k with element V(j,j,k,k) from an L x L table in LDS
// and target ranks from the position-major hop columns shared by every
// request of the same (L, pairs) shape -- and plain FP64 Lanczos on it: the
// Krylov vector v_j in LDS (the gathers of the product read it there), w,
// v_{j-1} and the diagonal in the request's global scratch (rows tid + k NT,
// each touched by one thread only), alpha_j and beta_j in LDS.  Every
// `check` steps wave 0 finds the two lowest eigenvalues of the tridiagonal
// T_m by Sturm-count multisection (64 lanes, one candidate each) and the
// last components of their eigenvectors by inverse iteration, and applies
// the solver's stopping tests: ||r_0|| = |beta_m z_m| <= tol |theta_0|, or the
// Kato-Temple estimate ||r_0||^2 / (theta_1 - ||r_1|| - theta_0) <= tol_energy.
// Without reorthogonalisation converged Ritz values reappear as copies; a
// second value within 1e-10 |theta_0| of the first is such a copy and the
// gap is taken to the next distinct one.  All reductions have a fixed order,
// so a request's result does not depend on the rest of its batch.
#include "batch.h"

namespace cf {

namespace {

constexpr int kMaxSteps = kBatchMaxSteps;

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sturm count: eigenvalues of T_m below x (LDL^T pivots of T - x I).
__device__ int sturm_count(const double* al, const double* be, int m, double x) {
    int c = 0;
    double q = al[0] - x;
    if (q < 0.0) ++c;
    for (int i = 1; i < m; ++i) {
        if (q == 0.0) q = 1e-300;
        q = (al[i] - x) - be[i - 1] * be[i - 1] / q;
        if (q < 0.0) ++c;
    }
    return c;
}

// k-th smallest eigenvalue (k = 0, 1, ...) of T_m by multisection over one
// wave: every lane tests one point, the bracket shrinks 65-fold per round.
__device__ double kth_eigenvalue(const double* al, const double* be, int m, int k, double lo, double hi) {
    const int lane = threadIdx.x & 63;
    for (int round = 0; round < 12; ++round) {
        const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
        const int c = sturm_count(al, be, m, x);
        // the first lane whose point has more than k eigenvalues below it
        const unsigned long long ball = __ballot(c > k);
        const int first = ball ? __builtin_ctzll(ball) : 64;
        const double nlo = lo + (hi - lo) * (double)first / 65.0;
        const double nhi = first < 64 ? lo + (hi - lo) * (double)(first + 1) / 65.0 : hi;
        if (nhi - nlo <= 1e-16 * fmax(1.0, fabs(nlo))) {
            lo = nlo;
            hi = nhi;
            break;
        }
        lo = nlo;
        hi = nhi;
    }
    return 0.5 * (lo + hi);
}

// Last component of the normalised eigenvector of T_m for eigenvalue theta:
// two steps of inverse iteration (tridiagonal elimination), one thread.
__device__ double last_component(const double* al, const double* be, int m, double theta, double* wk) {
    double* d = wk;        // pivots
    double* z = wk + kMaxSteps;
    for (int i = 0; i < m; ++i) z[i] = 1.0;
    for (int it = 0; it < 2; ++it) {
        // forward elimination of (T - theta I) z_new = z
        d[0] = al[0] - theta;
        if (fabs(d[0]) < 1e-150) d[0] = 1e-150;
        for (int i = 1; i < m; ++i) {
            const double l = be[i - 1] / d[i - 1];
            d[i] = (al[i] - theta) - l * be[i - 1];
            if (fabs(d[i]) < 1e-150) d[i] = 1e-150;
            z[i] -= l * z[i - 1];
        }
        z[m - 1] /= d[m - 1];
        for (int i = m - 2; i >= 0; --i) z[i] = (z[i] - be[i] * z[i + 1]) / d[i];
        double zmax = 0.0;  // scale first: the solve is near-singular by design
        for (int i = 0; i < m; ++i) zmax = fmax(zmax, fabs(z[i]));
        if (!(zmax > 0.0) || !(zmax < 1e300)) return 0.0;
        double nrm = 0.0;
        for (int i = 0; i < m; ++i) {
            z[i] /= zmax;
            nrm += z[i] * z[i];
        }
        nrm = 1.0 / sqrt(nrm);
        for (int i = 0; i < m; ++i) z[i] *= nrm;
    }
    return z[m - 1];
}

__device__ __forceinline__ double start_value(uint64_t seed, uint64_t q) {
    uint64_t z = seed + 0x9E3779B97F4A7C15ULL * (q + 1);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    return 1.0 + 0.5 * ((double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
}

}  // namespace

// NT threads, rows tid + k * NT for k < R (n <= NT * R = MAXN).
template <int NT, int R>
__global__ void __launch_bounds__(NT) k_doci_batch_lanczos(const DociBatchJob* __restrict__ jobs) {
    constexpr int MAXN = NT * R;
    constexpr int NW = NT / 64;
    __shared__ double x[MAXN];                       // v_j
    __shared__ double hop[kBatchMaxL * kBatchMaxL];  // V(j,j,k,k)
    __shared__ double pd[kBatchMaxL * kBatchMaxL];   // 2 (kk|ll) - (kl|lk), k != l
    __shared__ double d1[kBatchMaxL];                // 2 h_ii + (ii|ii)
    __shared__ double al[kMaxSteps], be[kMaxSteps];
    __shared__ double wk[2 * kMaxSteps];
    __shared__ double red[2][NW];
    __shared__ double ctl[4];                        // beta, stop flag, theta0, res0
    const DociBatchJob J = jobs[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int L = J.L, np = J.np, nv = J.nv;
    const int n = J.n;
    for (int t = tid; t < L * L; t += NT) {
        const int j = t / L, k = t - j * L;
        const int lo = j < k ? j : k, hi = j < k ? k : j;
        hop[t] = J.V[(((size_t)lo * L + lo) * L + hi) * L + hi];
        pd[t] = j == k ? 0.0
                       : 2.0 * J.V[(((size_t)j * L + k) * L + j) * L + k] - J.V[(((size_t)j * L + k) * L + k) * L + j];
    }
    for (int t = tid; t < L; t += NT) d1[t] = 2.0 * J.h[t * L + t] + J.V[(((size_t)t * L + t) * L + t) * L + t];
    __syncthreads();
    // this thread's rows r = tid, tid + NT, ...: diagonal, previous Krylov
    // vector and w in the request's scratch (each row touched by one thread only)
    double* dg = J.scratch;
    double* vp = J.scratch + n;
    double* w = J.scratch + 2 * (size_t)n;
    double part = 0.0;
    for (int r = tid; r < n; r += NT) {
        const str_t s = J.s[r];
        double e = 0.0;
        for (str_t u = s; u; u &= u - 1) {
            const int i = __builtin_ctzll(u);
            e += d1[i];
            for (str_t v = s; v; v &= v - 1) e += pd[i * L + __builtin_ctzll(v)];
        }
        dg[r] = e;
        vp[r] = 0.0;
        const double v0 = start_value(J.seed, (uint64_t)r);
        w[r] = v0;
        part += v0 * v0;
    }
    auto block_sum2 = [&](double a, double b, double* outa, double* outb) {
        a = wave_sum64(a);
        b = wave_sum64(b);
        if (lane == 0) {
            red[0][wv] = a;
            red[1][wv] = b;
        }
        __syncthreads();
        double sa = 0.0, sb = 0.0;
        for (int i = 0; i < NW; ++i) {
            sa += red[0][i];
            sb += red[1][i];
        }
        *outa = sa;
        *outb = sb;
        __syncthreads();
    };
    double nrm2, dummy;
    block_sum2(part, 0.0, &nrm2, &dummy);
    {
        const double sc = 1.0 / sqrt(nrm2);
        for (int r = tid; r < n; r += NT) x[r] = w[r] * sc;
    }
    __syncthreads();
    const str_t full = ((str_t)1 << L) - 1;
    double beta_prev = 0.0, theta0 = 0.0, res0 = 0.0;
    int steps = 0, status = 1;
    const int maxs = J.max_steps < kMaxSteps ? J.max_steps : kMaxSteps;
    for (int j = 0; j < maxs; ++j) {
        // w = H v_j
        double pa = 0.0;
        for (int r = tid; r < n; r += NT) {
            const str_t s = J.s[r], vir = ~s & full;
            const double xr = x[r];
            double acc = dg[r] * xr;
            const uint32_t* cp = J.col + r;
            uint32_t t = 0;
            str_t o = s;
            for (int a = 0; a < np; ++a) {
                const int jj = 63 - __builtin_clzll(o);
                o ^= (str_t)1 << jj;
                str_t v = vir;
                for (int b = 0; b < nv; ++b) {
                    const int kk = 63 - __builtin_clzll(v);
                    v ^= (str_t)1 << kk;
                    acc += hop[jj * L + kk] * x[cp[(size_t)t * n]];
                    ++t;
                }
            }
            w[r] = acc;
            pa += xr * acc;
        }
        double alpha;
        block_sum2(pa, 0.0, &alpha, &dummy);
        double pb = 0.0;
        for (int r = tid; r < n; r += NT) {
            const double wr = w[r] - alpha * x[r] - beta_prev * vp[r];
            w[r] = wr;
            pb += wr * wr;
        }
        double b2;
        block_sum2(pb, 0.0, &b2, &dummy);
        const double beta = sqrt(b2);
        if (tid == 0) {
            al[j] = alpha;
            be[j] = beta;
        }
        steps = j + 1;
        // convergence test on T_{j+1} (wave 0)
        const bool test = (j + 1 >= 4 && (j + 1) % J.check == 0) || j + 1 == maxs || !(beta > 1e-14 * fabs(alpha));
        if (test) {
            __syncthreads();
            if (wv == 0) {
                const int m = j + 1;
                // Gershgorin bracket
                double glo = 1e300, ghi = -1e300;
                for (int i = lane; i < m; i += 64) {
                    const double rad = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i < m - 1 ? fabs(be[i]) : 0.0);
                    glo = fmin(glo, al[i] - rad);
                    ghi = fmax(ghi, al[i] + rad);
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    glo = fmin(glo, __shfl_xor(glo, off, 64));
                    ghi = fmax(ghi, __shfl_xor(ghi, off, 64));
                }
                const double pad = 1e-12 * fmax(1.0, fmax(fabs(glo), fabs(ghi)));
                glo -= pad;
                ghi += pad;
                const double th0 = kth_eigenvalue(al, be, m, 0, glo, ghi);
                double th1 = 0.0;
                int k1 = 1;
                bool have1 = false;
                for (; k1 < m && k1 < 6; ++k1) {  // next distinct value (copies of theta0 skipped)
                    th1 = kth_eigenvalue(al, be, m, k1, glo, ghi);
                    if (th1 - th0 > 1e-10 * fmax(1.0, fabs(th0))) {
                        have1 = true;
                        break;
                    }
                }
                if (lane == 0) {
                    const double z0 = last_component(al, be, m, th0, wk);
                    const double r0 = fabs(beta * z0);
                    double r1 = 0.0;
                    if (have1) r1 = fabs(beta * last_component(al, be, m, th1, wk));
                    bool stop = r0 <= J.tol * fabs(th0) || !(beta > 1e-14 * fabs(alpha));
                    if (!stop && J.tol_energy > 0.0 && have1) {
                        const double mu = th1 - r1;
                        stop = mu > th0 && r0 * r0 <= J.tol_energy * (mu - th0);
                    }
                    ctl[1] = stop ? 1.0 : 0.0;
                    ctl[2] = th0;
                    ctl[3] = r0;
                }
            }
            __syncthreads();
            theta0 = ctl[2];
            res0 = ctl[3];
            if (ctl[1] != 0.0) {
                status = 0;
                break;
            }
        }
        if (j + 1 == maxs) break;
        // v_{j+1} = w / beta; v_j kept in registers
        __syncthreads();  // every gather of v_j is done
        const double inv = 1.0 / beta;
        for (int r = tid; r < n; r += NT) {
            vp[r] = x[r];
            x[r] = w[r] * inv;
        }
        beta_prev = beta;
        __syncthreads();
    }
    if (tid == 0) {
        J.out[0] = theta0;
        J.out[1] = res0;
        J.out[2] = (double)steps;
        J.out[3] = (double)status;
    }
}

int doci_batch_capacity() { return 1024 * 13; }

void launch_doci_batch_lanczos(const DociBatchJob* d_jobs, int count, int max_n, hipStream_t s) {
    if (count <= 0) return;
    if (max_n <= 1024) k_doci_batch_lanczos<256, 4><<<(unsigned)count, 256, 0, s>>>(d_jobs);
    else if (max_n <= 4096) k_doci_batch_lanczos<512, 8><<<(unsigned)count, 512, 0, s>>>(d_jobs);
    else if (max_n <= 1024 * 13) k_doci_batch_lanczos<1024, 13><<<(unsigned)count, 1024, 0, s>>>(d_jobs);
    else throw Error(ERR_UNSUPPORTED, "batched DOCI solve: more than 13,312 determinants");
    CF_HIP(hipGetLastError());
}

}  // namespace cf
