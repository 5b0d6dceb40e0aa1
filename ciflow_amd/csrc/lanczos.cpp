// Thick-restart Lanczos state machine (see lanczos.h).
#include "lanczos.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>

namespace cf {

// Householder reduction to tridiagonal form followed by the implicit QL
// algorithm with Wilkinson-type shifts (the classical symmetric eigensolver,
// e.g. Golub & Van Loan, Matrix Computations, 4th ed., sections 8.3.1 and
// 8.3.3).  O(m^3) with a small constant; the Krylov matrices here have m <= 64.
namespace {

// A := Q^T A Q tridiagonal, reducing rows m-1 .. 1 in turn, so index m-1 is
// never touched by a reflection (row m-1 of Q is e_{m-1}).  On return d holds
// h of each reflection (0 when the row needed none), e the sub-diagonal
// (e[i] couples i-1 and i) and A the reflectors, ready for accumulation.
void householder(int m, std::vector<double>& A, std::vector<double>& d, std::vector<double>& e) {
    d.assign(m, 0.0);
    e.assign(m, 0.0);
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * m + j]; };
    for (int i = m - 1; i > 0; --i) {
        const int l = i - 1;
        double h = 0.0;
        if (l > 0) {
            // a row that is already tridiagonal needs no reflection: the
            // Krylov matrices are tridiagonal apart from the arrow block a
            // thick restart leaves, so this skips most rows
            double below = 0.0;
            for (int k = 0; k < l; ++k) below += std::fabs(a(i, k));
            const double scale = below + std::fabs(a(i, l));
            if (below == 0.0) {
                e[i] = a(i, l);
            } else {
                for (int k = 0; k <= l; ++k) {
                    a(i, k) /= scale;
                    h += a(i, k) * a(i, k);
                }
                const double f = a(i, l);
                const double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
                e[i] = scale * g;
                h -= f * g;
                a(i, l) = f - g;
                double ff = 0.0;
                for (int j = 0; j <= l; ++j) {
                    a(j, i) = a(i, j) / h;
                    double gg = 0.0;
                    for (int k = 0; k <= j; ++k) gg += a(j, k) * a(i, k);
                    for (int k = j + 1; k <= l; ++k) gg += a(k, j) * a(i, k);
                    e[j] = gg / h;
                    ff += e[j] * a(i, j);
                }
                const double hh = ff / (h + h);
                for (int j = 0; j <= l; ++j) {
                    const double fj = a(i, j);
                    const double gj = e[j] - hh * fj;
                    e[j] = gj;
                    for (int k = 0; k <= j; ++k) a(j, k) -= fj * e[k] + gj * a(i, k);
                }
            }
        } else {
            e[i] = a(i, l);
        }
        d[i] = h;
    }
    e[0] = 0.0;
}

// Implicit QL on the tridiagonal (w diagonal, e sub-diagonal with e[i]
// coupling i-1 and i).  Every plane rotation of columns (i, i+1) is passed to
// rot(i, s, c) for the caller to apply to its vectors.
template <class Rot>
void implicit_ql(int m, std::vector<double>& w, std::vector<double>& e, Rot rot) {
    for (int i = 1; i < m; ++i) e[i - 1] = e[i];
    if (m > 0) e[m - 1] = 0.0;
    for (int l = 0; l < m; ++l) {
        int iter = 0;
        for (;;) {
            int mm = l;
            for (; mm < m - 1; ++mm) {
                const double dd = std::fabs(w[mm]) + std::fabs(w[mm + 1]);
                if (std::fabs(e[mm]) <= std::numeric_limits<double>::epsilon() * dd) break;
            }
            if (mm == l) break;
            if (++iter > 60) break;  // not reached for symmetric input
            double g = (w[l + 1] - w[l]) / (2.0 * e[l]);
            // plain square roots: the entries are bounded by the operator
            // norm, far from overflow, and std::hypot costs several times more
            double r = std::sqrt(g * g + 1.0);
            g = w[mm] - w[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
            double s = 1.0, c = 1.0, p = 0.0;
            int i = mm - 1;
            for (; i >= l; --i) {
                const double f = s * e[i];
                const double b = c * e[i];
                r = std::sqrt(f * f + g * g);
                e[i + 1] = r;
                if (r == 0.0) {
                    w[i + 1] -= p;
                    e[mm] = 0.0;
                    break;
                }
                s = f / r;
                c = g / r;
                g = w[i + 1] - p;
                r = (w[i] - g) * s + 2.0 * c * b;
                p = s * r;
                w[i + 1] = g + p;
                g = c * r - b;
                rot(i, s, c);
            }
            if (r == 0.0 && i >= l) continue;
            w[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
}

std::vector<int> ascending(const std::vector<double>& w) {
    std::vector<int> idx(w.size());
    for (size_t i = 0; i < w.size(); ++i) idx[i] = (int)i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return w[x] < w[y]; });
    return idx;
}

}  // namespace

void sym_eig(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& Z) {
    std::vector<double> e;
    householder(m, A, w, e);
    auto a = [&](int i, int j) -> double& { return A[(size_t)i * m + j]; };
    // accumulate Q into A, leaving the diagonal in w
    for (int i = 0; i < m; ++i) {
        const int l = i - 1;
        if (w[i] != 0.0) {
            for (int j = 0; j <= l; ++j) {
                double g = 0.0;
                for (int k = 0; k <= l; ++k) g += a(i, k) * a(k, j);
                for (int k = 0; k <= l; ++k) a(k, j) -= g * a(k, i);
            }
        }
        w[i] = a(i, i);
        a(i, i) = 1.0;
        for (int j = 0; j <= l; ++j) a(j, i) = a(i, j) = 0.0;
    }
    implicit_ql(m, w, e, [&](int i, double s, double c) {
        for (int k = 0; k < m; ++k) {
            const double f = a(k, i + 1);
            a(k, i + 1) = s * a(k, i) + c * f;
            a(k, i) = c * a(k, i) - s * f;
        }
    });
    // ascending order, vectors as columns of Z
    const std::vector<int> idx = ascending(w);
    std::vector<double> ws(m);
    Z.assign((size_t)m * m, 0.0);
    for (int c = 0; c < m; ++c) {
        ws[c] = w[idx[c]];
        for (int r = 0; r < m; ++r) Z[(size_t)r * m + c] = a(r, idx[c]);
    }
    w.swap(ws);
}

void sym_eig_last_row(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& zlast) {
    std::vector<double> e;
    householder(m, A, w, e);
    for (int i = 0; i < m; ++i) w[i] = A[(size_t)i * m + i];
    // row m-1 of Q is e_{m-1}; only that row of Q R is followed through QL
    std::vector<double> z(m, 0.0);
    if (m > 0) z[m - 1] = 1.0;
    implicit_ql(m, w, e, [&](int i, double s, double c) {
        const double f = z[i + 1];
        z[i + 1] = s * z[i] + c * f;
        z[i] = c * z[i] - s * f;
    });
    const std::vector<int> idx = ascending(w);
    std::vector<double> ws(m);
    zlast.assign(m, 0.0);
    for (int c = 0; c < m; ++c) {
        ws[c] = w[idx[c]];
        zlast[c] = z[idx[c]];
    }
    w.swap(ws);
}

ThickRestartLanczos::ThickRestartLanczos(KrylovBackend* b, uint64_t n, int ncv, double tol, int maxrestart,
                                         uint64_t seed, bool start_given, int check_every, double tol_energy)
    : b_(b), n_(n), tol_(tol), maxrestart_(maxrestart), seed_(seed), start_given_(start_given),
      check_every_(check_every > 0 ? check_every : 0), tol_energy_(tol_energy > 0.0 ? tol_energy : 0.0) {
    M_ = (uint64_t)ncv > n ? (int)n : ncv;
    if (M_ < 1) M_ = 1;
    T_.assign((size_t)M_ * M_, 0.0);
}

ThickRestartLanczos::Event ThickRestartLanczos::step() {
    for (;;) {
        switch (phase_) {
            case INIT:
                if (start_given_) b_->start_keep();
                else b_->start_random(seed_);
                j_ = k_ = filled_ = 0;
                phase_ = LOOP_APPLY;
                break;
            case LOOP_APPLY:
                pending_ = j_;
                phase_ = AFTER_APPLY;
                return NEED_APPLY;
            case AFTER_APPLY: {
                ++napply_;
                if (b_->deferred()) {
                    b_->orth_deferred(j_);
                    ++j_;
                    if (j_ == M_) {
                        if (!absorb(j_)) {
                            m_ = M_;
                            eig_and_restart();
                        }
                        break;
                    }
                    if (check_every_ && (j_ - k_) % check_every_ == 0 && (absorb(j_) || converged_early(j_)))
                        break;
                    phase_ = LOOP_APPLY;
                    break;
                }
                double a = 0.0, beta = 0.0;
                b_->orth(j_, &a, &beta);
                T_[(size_t)j_ * M_ + j_] = a;
                beta_last_ = beta;
                anorm_ = std::max(anorm_, std::fabs(a) + beta);
                const bool breakdown = beta <= 1e-15 * anorm_ || (uint64_t)(j_ + 1) == n_;
                if (breakdown) {
                    m_ = j_ + 1;
                    breakdown_ = true;
                    eig_and_restart();
                    break;
                }
                b_->set_next(j_ + 1, beta);
                if (j_ + 1 < M_) {
                    T_[(size_t)j_ * M_ + j_ + 1] = beta;
                    T_[(size_t)(j_ + 1) * M_ + j_] = beta;
                }
                ++j_;
                if (j_ < M_) {
                    phase_ = LOOP_APPLY;
                } else {
                    m_ = M_;
                    eig_and_restart();
                }
                break;
            }
            case FINAL_APPLY:
                pending_ = 0;
                phase_ = AFTER_FINAL;
                return NEED_APPLY;
            case AFTER_FINAL:
                ++napply_;
                b_->rayleigh(&theta_, &resid_);
                phase_ = FINISHED;
                return DONE;
            case FINISHED:
                return DONE;
        }
    }
}

// Steps filled_..j1-1 ran without host round trips: read their coefficients
// and replay them in order with the breakdown test of the immediate path.
// Returns true when a breakdown ended the cycle (products applied after it
// are discounted: their vectors are not used).
bool ThickRestartLanczos::absorb(int j1) {
    const int j0 = filled_;
    if (j1 <= j0) return false;
    std::vector<double> al(j1 - j0), be(j1 - j0);
    b_->fetch_coeffs(j0, j1, al.data(), be.data());
    for (int jj = j0; jj < j1; ++jj) {
        const double a = al[jj - j0], beta = be[jj - j0];
        T_[(size_t)jj * M_ + jj] = a;
        beta_last_ = beta;
        anorm_ = std::max(anorm_, std::fabs(a) + beta);
        if (beta <= 1e-15 * anorm_ || (uint64_t)(jj + 1) == n_) {
            napply_ -= j_ - 1 - jj;
            m_ = jj + 1;
            breakdown_ = true;
            eig_and_restart();
            return true;
        }
        if (jj + 1 < M_) {
            T_[(size_t)jj * M_ + jj + 1] = beta;
            T_[(size_t)(jj + 1) * M_ + jj] = beta;
        }
    }
    filled_ = j1;
    return false;
}

bool ThickRestartLanczos::accept(double res, const std::vector<double>& th, double res1) const {
    if (progress_) progress_(napply_, th[0], res);
    if (res <= tol_ * std::fabs(th[0])) return true;
    return tol_energy_ > 0.0 && th.size() >= 2 && temple_accept(res, th[0], th[1], res1, tol_energy_);
}

// Convergence test on the leading m x m block in the middle of a cycle.
bool ThickRestartLanczos::converged_early(int m) {
    std::vector<double> A((size_t)m * m);
    for (int a = 0; a < m; ++a)
        for (int c = 0; c < m; ++c) A[(size_t)a * m + c] = T_[(size_t)a * M_ + c];
    // eigenvalues and the last row of the eigenvectors decide; the full
    // decomposition is formed only once the test passes
    std::vector<double> th, zlast;
    sym_eig_last_row(m, A, th, zlast);
    const double res = std::fabs(beta_last_ * zlast[0]);
    if (!accept(res, th, m >= 2 ? std::fabs(beta_last_ * zlast[1]) : 0.0)) return false;
    std::vector<double> Y;
    sym_eig(m, std::move(A), th, Y);
    m_ = m;
    status_ = 0;
    b_->rotate(m, 1, Y.data());
    b_->normalize0();
    phase_ = FINAL_APPLY;
    return true;
}

void ThickRestartLanczos::eig_and_restart() {
    const int m = m_;
    std::vector<double> A((size_t)m * m);
    for (int a = 0; a < m; ++a)
        for (int c = 0; c < m; ++c) A[(size_t)a * m + c] = T_[(size_t)a * M_ + c];
    std::vector<double> th, Y;
    sym_eig(m, A, th, Y);
    const double res = std::fabs(beta_last_ * Y[(size_t)(m - 1) * m + 0]);
    const double res1 = m >= 2 ? std::fabs(beta_last_ * Y[(size_t)(m - 1) * m + 1]) : 0.0;
    const bool converged = accept(res, th, res1) || m < M_ || breakdown_;
    const bool limit = nrestart_ >= maxrestart_;
    if (converged || limit) {
        status_ = converged ? 0 : 1;
        b_->rotate(m, 1, Y.data());
        b_->normalize0();
        phase_ = FINAL_APPLY;
        return;
    }
    // keep the kk lowest Ritz vectors plus the residual direction v_m
    const int kk = std::max(1, std::min(m / 2, m - 2));
    b_->rotate(m, kk, Y.data());
    b_->move(m, kk);
    std::fill(T_.begin(), T_.end(), 0.0);
    for (int c = 0; c < kk; ++c) {
        T_[(size_t)c * M_ + c] = th[c];
        const double s = beta_last_ * Y[(size_t)(m - 1) * m + c];
        T_[(size_t)c * M_ + kk] = s;
        T_[(size_t)kk * M_ + c] = s;
    }
    k_ = kk;
    j_ = kk;
    filled_ = kk;
    ++nrestart_;
    phase_ = LOOP_APPLY;
}

}  // namespace cf
