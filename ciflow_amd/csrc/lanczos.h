// Thick-restart Lanczos for the lowest eigenpair, written as a
// reverse-communication state machine so the same control logic drives
//   * the GPU-resident solver (Krylov vectors in HBM, operator = GPU SpMV),
//   * the dsaupd_-shaped host RCI entry point (cf_rci_*), which the
//     reference's arpackDiagonalize loop shape expects
//     (SparseMatrix_CRS.cpp:441-449: call, get ido == 1, apply mvprod, call).
//
// ARPACK (arpack-ng 3.1.5-3.3.0, an external Fortran dependency of the
// reference) is not available; this is the published thick-restart Lanczos
// of Wu & Simon (SIAM J. Matrix Anal. Appl. 22 (2000) 602) with full
// classical Gram-Schmidt re-orthogonalisation done twice, keeping ARPACK's
// parameters (which="SA", nev=1, ncv=min(42,n), convergence
// ||r|| <= tol*|theta|).  The returned eigenvalue is the Rayleigh quotient of
// the normalised Ritz vector, evaluated with one extra operator application.
#pragma once

#include <cstdint>
#include <functional>
#include <vector>

namespace cf {

// Vector storage + operations of one Lanczos run.  Indices refer to basis
// columns 0..ncv (ncv+1 columns); "w" is the work vector.
struct KrylovBackend {
    virtual ~KrylovBackend() {}
    virtual void start_random(uint64_t seed) = 0;  // v0 = normalised deterministic random vector
    virtual void start_keep() = 0;                 // v0 already holds the start vector: normalise it
    virtual void apply(int j) = 0;                 // w = A v_j (not used by the RCI backend)
    // Two passes of classical Gram-Schmidt of w against v_0..v_j; alpha = total
    // coefficient on v_j, beta = ||w|| afterwards.
    virtual void orth(int j, double* alpha, double* beta) = 0;
    virtual void set_next(int j1, double beta) = 0;  // v_j1 = w / beta
    virtual void rotate(int m, int k, const double* Y) = 0;  // v_0..v_k-1 = V_m Y[:, 0:k] (Y m x m row-major)
    virtual void move(int from, int to) = 0;
    virtual void normalize0() = 0;
    // with w = A v_0: theta = v_0.w, resid = ||w - theta v_0||
    virtual void rayleigh(double* theta, double* resid) = 0;
    // Deferred coefficients (a backend whose vectors live on a device): orth
    // plus v_{j+1} = w / beta_j entirely on the device, alpha_j and beta_j kept
    // there; fetch_coeffs returns those of steps [j0, j1) with one wait.  The
    // arithmetic is that of orth + set_next, so results do not change.
    virtual bool deferred() const { return false; }
    virtual void orth_deferred(int /*j*/) {}
    virtual void fetch_coeffs(int /*j0*/, int /*j1*/, double* /*alpha*/, double* /*beta*/) {}
};

// Symmetric eigen-decomposition (Householder + implicit QL), ascending;
// Z columns = vectors.
void sym_eig(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& Z);
// The same eigenvalues with only the last row of Z (zlast[c] = Z[m-1][c]),
// O(m^2) when A is tridiagonal apart from a leading arrow block.
void sym_eig_last_row(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& zlast);

// Kato-Temple stop for the lowest Ritz pair (theta0, residual res) of a
// Rayleigh-Ritz step: lambda_0 >= theta0 - res^2 / (mu - theta0) for any mu
// with theta0 < mu <= lambda_1.  The second Ritz value theta1 lies at or above
// lambda_1 (interlacing), so it is no valid mu by itself; mu = theta1 - res1,
// with res1 the residual norm of the second Ritz pair, moves it down by the
// distance within which that pair's eigenvalue is known.  Accepts when
// mu > theta0 and res^2 <= tol * (mu - theta0).
inline bool temple_accept(double res, double theta0, double theta1, double res1, double tol) {
    const double mu = theta1 - res1;
    return mu > theta0 && res * res <= tol * (mu - theta0);
}

class ThickRestartLanczos {
public:
    enum Event { NEED_APPLY = 1, DONE = 99 };
    // check_every > 0 (deferred backends): also test convergence every
    // check_every steps inside a cycle, not only when the basis is full.
    // tol_energy > 0: also accept the Ritz pair once the Kato-Temple bound on
    // its eigenvalue error, ||r||^2 / (mu - theta_0), is <= tol_energy, with
    // mu = theta_1 - ||r_1|| (temple_accept).
    ThickRestartLanczos(KrylovBackend* b, uint64_t n, int ncv, double tol, int maxrestart, uint64_t seed,
                        bool start_given, int check_every = 0, double tol_energy = 0.0);
    // Advance until the operator must be applied to v_{pending()} (NEED_APPLY,
    // the result is expected in the backend's w) or the run ended (DONE).
    Event step();
    int pending() const { return pending_; }
    double eigval() const { return theta_; }
    double resid() const { return resid_; }
    int n_apply() const { return napply_; }
    int n_restart() const { return nrestart_; }
    int status() const { return status_; }  // 0 converged, 1 restart limit
    // Called at every convergence test: (products so far, lowest Ritz value, residual estimate).
    void set_progress(std::function<void(int, double, double)> f) { progress_ = std::move(f); }

private:
    enum Phase { INIT, LOOP_APPLY, AFTER_APPLY, FINAL_APPLY, AFTER_FINAL, FINISHED };
    void eig_and_restart();
    bool absorb(int j1);
    bool converged_early(int m);
    KrylovBackend* b_;
    uint64_t n_;
    int M_, m_ = 0, j_ = 0, k_ = 0;
    double tol_;
    int maxrestart_;
    uint64_t seed_;
    bool start_given_;
    int check_every_ = 0, filled_ = 0;
    double tol_energy_ = 0.0;
    std::function<void(int, double, double)> progress_;
    bool accept(double res, const std::vector<double>& th, double res1) const;
    Phase phase_ = INIT;
    int pending_ = 0;
    std::vector<double> T_;
    double beta_last_ = 0.0, anorm_ = 0.0;
    bool breakdown_ = false;
    double theta_ = 0.0, resid_ = 0.0;
    int napply_ = 0, nrestart_ = 0, status_ = 1;
};

}  // namespace cf
