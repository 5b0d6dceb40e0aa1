// Thick-restart Lanczos state machine (see lanczos.h).
#include "lanczos.h"

#include <algorithm>
#include <cmath>

namespace cf {

void sym_eig(int m, std::vector<double> A, std::vector<double>& w, std::vector<double>& Z) {
    Z.assign((size_t)m * m, 0.0);
    for (int i = 0; i < m; ++i) Z[(size_t)i * m + i] = 1.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                double a = A[(size_t)i * m + j] * A[(size_t)i * m + j];
                tot += a;
                if (i != j) off += a;
            }
        if (off == 0.0 || off <= 1e-34 * tot) break;
        for (int p = 0; p < m - 1; ++p)
            for (int q = p + 1; q < m; ++q) {
                const double apq = A[(size_t)p * m + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * m + p], aqq = A[(size_t)q * m + q];
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < m; ++r) {
                    double arp = A[(size_t)r * m + p], arq = A[(size_t)r * m + q];
                    A[(size_t)r * m + p] = c * arp - s * arq;
                    A[(size_t)r * m + q] = s * arp + c * arq;
                }
                for (int r = 0; r < m; ++r) {
                    double apr = A[(size_t)p * m + r], aqr = A[(size_t)q * m + r];
                    A[(size_t)p * m + r] = c * apr - s * aqr;
                    A[(size_t)q * m + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < m; ++r) {
                    double zrp = Z[(size_t)r * m + p], zrq = Z[(size_t)r * m + q];
                    Z[(size_t)r * m + p] = c * zrp - s * zrq;
                    Z[(size_t)r * m + q] = s * zrp + c * zrq;
                }
            }
    }
    std::vector<int> idx(m);
    for (int i = 0; i < m; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](int a, int b) { return A[(size_t)a * m + a] < A[(size_t)b * m + b]; });
    w.resize(m);
    std::vector<double> Zs((size_t)m * m);
    for (int c = 0; c < m; ++c) {
        w[c] = A[(size_t)idx[c] * m + idx[c]];
        for (int r = 0; r < m; ++r) Zs[(size_t)r * m + c] = Z[(size_t)r * m + idx[c]];
    }
    Z.swap(Zs);
}

ThickRestartLanczos::ThickRestartLanczos(KrylovBackend* b, uint64_t n, int ncv, double tol, int maxrestart,
                                         uint64_t seed, bool start_given)
    : b_(b), n_(n), tol_(tol), maxrestart_(maxrestart), seed_(seed), start_given_(start_given) {
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
                j_ = k_ = 0;
                phase_ = LOOP_APPLY;
                break;
            case LOOP_APPLY:
                pending_ = j_;
                phase_ = AFTER_APPLY;
                return NEED_APPLY;
            case AFTER_APPLY: {
                ++napply_;
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

void ThickRestartLanczos::eig_and_restart() {
    const int m = m_;
    std::vector<double> A((size_t)m * m);
    for (int a = 0; a < m; ++a)
        for (int c = 0; c < m; ++c) A[(size_t)a * m + c] = T_[(size_t)a * M_ + c];
    std::vector<double> th, Y;
    sym_eig(m, A, th, Y);
    const double res = std::fabs(beta_last_ * Y[(size_t)(m - 1) * m + 0]);
    const bool converged = res <= tol_ * std::fabs(th[0]) || m < M_ || breakdown_;
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
    ++nrestart_;
    phase_ = LOOP_APPLY;
}

}  // namespace cf
