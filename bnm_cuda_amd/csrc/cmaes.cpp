// cmaes.cpp -- CMA-ES driver state for the simulation population (host side).
//
// Restates the semantics of cmaes.cpp (CMAES_Points) of the reference: N.
// Hansen's CMA-ES (2006 variant by default, 2016 with negative weights as
// "variant 16"), sample points clipped into the unit box, the IQR-based
// boundary penalty of Hansen et al. 2009 (per-coordinate weight / scale terms added to the cost
// of clipped points), early stop when the best value has not moved by more
// than `early_stop_tol` over `early_stop_gens` generations.
//
// Differences from the reference: the symmetric eigendecomposition is a cyclic
// Jacobi sweep (the reference calls gsl_eigen_symmv) and eigenvectors may
// differ in sign, which changes the sampled points after the first update;
// the state lives in one object instead of globals.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/bnm_hip.h"

namespace {

// C = V diag(w) V^T for symmetric C (cyclic Jacobi), eigenvalues ascending by
// absolute value with V's columns alongside (gsl_eigen_symmv_sort ABS_ASC).
void jacobi_eig(int n, std::vector<double> A, std::vector<double>& w, std::vector<double>& V) {
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) off += A[(size_t)p * n + q] * A[(size_t)p * n + q];
        if (off < 1e-30) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (std::fabs(apq) < 1e-300) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                const double theta = (aqq - app) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    const double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) w[i] = A[(size_t)i * n + i];
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return std::fabs(w[a]) < std::fabs(w[b]); });
    std::vector<double> w2(n), V2((size_t)n * n);
    for (int i = 0; i < n; ++i) {
        w2[i] = w[idx[i]];
        for (int k = 0; k < n; ++k) V2[(size_t)k * n + i] = V[(size_t)k * n + idx[i]];
    }
    w = w2;
    V = V2;
}

double median_of(std::vector<double> v) {
    std::sort(v.begin(), v.end());
    const size_t n = v.size();
    return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// quantile of sorted data with linear interpolation (gsl_stats_quantile_from_sorted_data)
double quantile_sorted(const std::vector<double>& v, double f) {
    const double idx = f * (v.size() - 1);
    const size_t lo = (size_t)std::floor(idx);
    const double d = idx - lo;
    return lo + 1 < v.size() ? (1 - d) * v[lo] + d * v[lo + 1] : v[lo];
}

}  // namespace

struct bnm_cmaes {
    int D, lambda, mue, variant, it = 0, early_gens;
    unsigned seed_mw;
    double sigma, alphacov, gamma_scale, soft_edge, early_tol;
    double cc = 0, ccov = 0, csigma = 0, dsigma = 0, chiN = 0, muecov = 0, mueeff = 0, c1 = 0, cmue = 0;
    std::vector<double> weights, bound_weight, bound_scale, psigma, pc, m, m_prev, m_feasible, iqr_history, C, B, Dd;
    std::vector<double> X, Xfeas, Xpref;   // [D][lambda]
    std::vector<double> pool;              // [n_pool][D], raw (unclipped) points
    int n_pool = 0;
    std::vector<double> best_x, tvec;
    double best_val = -INFINITY, best_raw = 0;
    int stopped = 0;
    int hist_len;
};

extern "C" {

bnm_cmaes* bnm_cmaes_create(int dim, int lambda, unsigned seed_mw, double sigma0, int variant) {
    if (dim < 1 || lambda < 2 || (variant != 6 && variant != 16)) return nullptr;
    auto* e = new bnm_cmaes;
    e->D = dim;
    e->lambda = lambda;
    e->mue = lambda / 2;                 // run_CMAES.cpp:711
    e->variant = variant;
    e->seed_mw = seed_mw;
    e->sigma = sigma0;
    e->alphacov = 2;                     // constants.h defaults
    e->gamma_scale = 2.0;
    e->soft_edge = 0.0;
    e->early_gens = 30;
    e->early_tol = 1e-3;
    e->hist_len = (int)(20 + std::ceil((float)(3 * dim) / (float)lambda));
    const int D = dim, L = lambda, mu = e->mue;
    e->weights.assign(L, 0.0);
    double s1 = 0, s2 = 0;
    for (int i = 0; i < mu; ++i) {
        e->weights[i] = variant == 6 ? std::log(mu + 1.0) - std::log(i + 1.0)
                                     : std::log((L + 1) / 2.) - std::log(i + 1.0);
        s1 += e->weights[i];
        s2 += e->weights[i] * e->weights[i];
    }
    for (int i = 0; i < mu; ++i) e->weights[i] /= s1;
    e->mueeff = s1 * s1 / s2;
    e->muecov = e->mueeff;
    if (variant == 16) {
        double n1 = 0, n2 = 0;
        for (int i = mu; i < L; ++i) {
            e->weights[i] = std::log((L + 1) / 2.) - std::log(i + 1.0);
            n1 += e->weights[i];
            n2 += e->weights[i] * e->weights[i];
        }
        const double mueeffminus = n1 * n1 / n2;
        e->c1 = e->alphacov / ((D + 1.3) * (D + 1.3) + e->mueeff);
        const double t1 = e->alphacov * (e->mueeff - 2 + 1 / e->mueeff) /
                          ((D + 2) * (D + 2) + e->alphacov * e->mueeff / 2);
        e->cmue = std::min(1 - e->c1, t1);
        const double am = 1 + e->c1 / e->cmue, ame = 1 + 2 * mueeffminus / (e->mueeff + 2),
                     apd = (1 - e->c1 - e->cmue) / (D * e->cmue);
        const double t2 = std::min(std::min(am, ame), apd);
        for (int i = mu; i < L; ++i) e->weights[i] = e->weights[i] / (-n1) * t2;
        e->cc = (4 + e->mueeff / D) / (D + 4 + 2 * e->mueeff / D);
        e->csigma = (e->mueeff + 2) / (D + e->mueeff + 5);
    } else {
        e->cc = 4. / (D + 4);
        e->csigma = (e->mueeff + 2) / (D + e->mueeff + 3);
    }
    const double t1 = 2. / ((D + std::sqrt(2.0)) * (D + std::sqrt(2.0)));
    const double t2 = std::min(1.0, (2 * e->mueeff - 1) / ((D + 2) * (D + 2) + e->mueeff));
    e->ccov = (1 / e->muecov) * t1 + (1 - 1 / e->muecov) * t2;
    e->dsigma = 1 + 2 * std::max(0., std::sqrt((e->mueeff - 1) / (D + 1.)) - 1) + e->csigma;
    e->chiN = std::sqrt((double)D) * (1 - 1. / (4 * D) + 1. / (21. * D * D));
    e->bound_weight.assign(D, 0.0);
    e->bound_scale.assign(D, 1.0);
    e->psigma.assign(D, 0.0);
    e->pc.assign(D, 0.0);
    e->m_prev.assign(D, 0.0);
    e->m_feasible.assign(D, 0.0);
    e->iqr_history.assign(e->hist_len, 0.0);
    e->C.assign((size_t)D * D, 0.0);
    e->B.assign((size_t)D * D, 0.0);
    e->Dd.assign(D, 1.0);
    for (int i = 0; i < D; ++i) e->C[(size_t)i * D + i] = e->B[(size_t)i * D + i] = 1.0;
    e->m.assign(D, 0.0);
    std::default_random_engine gmw;
    gmw.seed(seed_mw);
    std::uniform_real_distribution<double> uni(0, 1.0);
    for (int i = 0; i < D; ++i) e->m[i] = uni(gmw);   // feasible first mean
    e->X.assign((size_t)D * L, 0.0);
    e->Xfeas.assign((size_t)D * L, 0.0);
    e->Xpref.assign((size_t)D * L, 0.0);
    e->best_x.assign(D, 0.0);
    e->tvec.assign(e->early_gens, 0.0);
    // first generation: x = m + sigma * z (C = I), generator seeded with 11 + it
    std::default_random_engine gad;
    gad.seed(11 + e->it);
    std::normal_distribution<double> nd(0, 1.0);
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < L; ++j) e->X[(size_t)i * L + j] = e->m[i] + e->sigma * nd(gad);
    auto clip = [&](bnm_cmaes* q) {
        for (int i = 0; i < q->D; ++i) {
            for (int j = 0; j < q->lambda; ++j) {
                const double v = q->X[(size_t)i * q->lambda + j];
                q->Xfeas[(size_t)i * q->lambda + j] = std::min(1.0, std::max(0.0, v));
                q->Xpref[(size_t)i * q->lambda + j] =
                    std::min(1.0 - q->soft_edge, std::max(0.0 + q->soft_edge, v));
            }
            q->m_feasible[i] = std::min(1.0 - q->soft_edge, std::max(0.0 + q->soft_edge, q->m[i]));
        }
    };
    clip(e);
    e->it = 1;
    return e;
}

void bnm_cmaes_destroy(bnm_cmaes* e) { delete e; }

// Current sample points in the unit box (feasible, clipped): x[j*D + i] for
// candidate j, parameter i.
int bnm_cmaes_ask(bnm_cmaes* e, double* x_feas) {
    if (!e || !x_feas) return BNM_E_ARG;
    for (int j = 0; j < e->lambda; ++j)
        for (int i = 0; i < e->D; ++i) x_feas[(size_t)j * e->D + i] = e->Xfeas[(size_t)i * e->lambda + j];
    return BNM_OK;
}

// Reports the costs (lower is better) of the current points: computes the
// boundary penalty from the IQR of the raw costs, adds it, updates the
// distribution and samples the next generation.  Returns 1 when the run has
// stopped (early stop, or C lost positive definiteness), else 0.
int bnm_cmaes_tell(bnm_cmaes* e, const double* cost, double* penalized) {
    if (!e || !cost) return BNM_E_ARG;
    const int D = e->D, L = e->lambda, mu = e->mue;
    std::vector<double> raw(cost, cost + L), corr(cost, cost + L);
    // ---- penalty pass (cmaes.cpp:498-553) ----
    {
        std::vector<double> sorted = raw;
        std::sort(sorted.begin(), sorted.end());
        double iqr_now = quantile_sorted(sorted, 0.75) - quantile_sorted(sorted, 0.25);
        double diag = 0;
        for (int i = 0; i < D; ++i) diag += e->C[(size_t)i * D + i];
        iqr_now /= (e->sigma * e->sigma * diag);
        e->iqr_history[(e->it - 1) % e->hist_len] = iqr_now;
        double logdiag = 0;
        for (int i = 0; i < D; ++i) {
            const double cii = e->C[(size_t)i * D + i];
            if (cii <= 0) { e->stopped = 1; return 1; }
            logdiag += std::log(cii);
            if (e->m[i] < e->soft_edge || e->m[i] > 1 - e->soft_edge || e->it == 1) {
                const int nh = e->it < e->hist_len ? e->it + 1 : e->hist_len;
                std::vector<double> h(e->iqr_history.begin(), e->iqr_history.begin() + std::min(nh, e->hist_len));
                e->bound_weight[i] = 5.0001 * D * median_of(h);
                const double dm = e->m[i] - e->m_feasible[i];
                if (std::fabs(dm) > 3 * e->sigma * std::sqrt(cii) * std::max(1.0, std::sqrt((double)D) / e->mueeff) &&
                    ((dm < 0) == (e->m[i] - e->m_prev[i] < 0)))
                    e->bound_weight[i] *= 1.1;
                e->bound_weight[i] *= e->gamma_scale;
            }
        }
        for (int i = 0; i < D; ++i)
            e->bound_scale[i] = std::exp(0.9 * (std::log(e->C[(size_t)i * D + i]) - logdiag / D));
        for (int j = 0; j < L; ++j)
            for (int i = 0; i < D; ++i) {
                const double d = e->Xpref[(size_t)i * L + j] - e->X[(size_t)i * L + j];
                corr[j] += e->bound_weight[i] / e->bound_scale[i] / D * d * d;
            }
    }
    if (penalized) std::copy(corr.begin(), corr.end(), penalized);
    // ---- update (cmaes.cpp:250-496) ----
    std::vector<int> idx(L);
    for (int j = 0; j < L; ++j) idx[j] = j;
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return corr[a] < corr[b]; });
    if (-corr[idx[0]] > e->best_val) {
        e->best_val = -corr[idx[0]];
        e->best_raw = raw[idx[0]];
        for (int i = 0; i < D; ++i) e->best_x[i] = e->Xfeas[(size_t)i * L + idx[0]];
    }
    std::vector<double> mean_step(D), tmp(D);
    for (int i = 0; i < D; ++i) {
        e->m_prev[i] = e->m[i];
        double s = 0;
        for (int k = 0; k < mu; ++k) s += e->weights[k] * e->X[(size_t)i * L + idx[k]];
        e->m[i] = s;
        mean_step[i] = std::sqrt(e->mueeff) * (e->m[i] - e->m_prev[i]) / e->sigma;
    }
    for (int i = 0; i < D; ++i) {
        double s = 0;
        for (int j = 0; j < D; ++j) s += e->B[(size_t)j * D + i] * mean_step[j];
        tmp[i] = s / e->Dd[i];
    }
    double np2 = 0;
    for (int i = 0; i < D; ++i) {
        double s = 0;
        for (int j = 0; j < D; ++j) s += e->B[(size_t)i * D + j] * tmp[j];
        e->psigma[i] = (1 - e->csigma) * e->psigma[i] + std::sqrt(e->csigma * (2 - e->csigma)) * s;
        np2 += e->psigma[i] * e->psigma[i];
    }
    const double lhs = std::sqrt(np2) / std::sqrt(1 - std::pow(1 - e->csigma, 2 * e->it)) / e->chiN;
    const int hsig = e->variant == 16 ? (lhs < 1.4 + 2. / (D + 1)) : (lhs < 1.5 + 1. / (D - 0.5));
    for (int i = 0; i < D; ++i)
        e->pc[i] = (1 - e->cc) * e->pc[i] + hsig * std::sqrt(e->cc * (2 - e->cc)) * mean_step[i];
    e->sigma *= std::exp((std::sqrt(np2) / e->chiN - 1) * e->csigma / e->dsigma);
    std::vector<double>& C = e->C;
    const double s2 = e->sigma * e->sigma;
    if (e->variant == 6) {
        for (int i = 0; i < D; ++i)
            for (int j = 0; j <= i; ++j) {
                double v = (1 - e->ccov) * C[(size_t)i * D + j] +
                           (e->ccov / e->muecov) * (e->pc[i] * e->pc[j] + (1 - hsig) * e->cc * (2 - e->cc) * C[(size_t)i * D + j]);
                for (int k = 0; k < mu; ++k)
                    v += e->ccov * (1 - 1 / e->muecov) * e->weights[k] *
                         (e->X[(size_t)i * L + idx[k]] - e->m_prev[i]) * (e->X[(size_t)j * L + idx[k]] - e->m_prev[j]) / s2;
                C[(size_t)i * D + j] = C[(size_t)j * D + i] = v;
            }
    } else {
        double wsum = 0;
        for (int k = 0; k < L; ++k) wsum += e->weights[k];
        std::vector<double> Cold = C;
        for (int i = 0; i < D; ++i)
            for (int j = 0; j <= i; ++j) {
                double v = (1 - e->c1 - e->cmue * wsum) * Cold[(size_t)i * D + j] +
                           e->c1 * (e->pc[i] * e->pc[j] + (1 - hsig) * e->cc * (2 - e->cc) * Cold[(size_t)i * D + j]);
                for (int k = 0; k < mu; ++k)
                    v += e->cmue * e->weights[k] * (e->X[(size_t)i * L + idx[k]] - e->m_prev[i]) *
                         (e->X[(size_t)j * L + idx[k]] - e->m_prev[j]) / s2;
                for (int k = mu; k < L; ++k) {
                    std::vector<double> y(D), t(D), h(D);
                    for (int z = 0; z < D; ++z) y[z] = (e->X[(size_t)z * L + idx[k]] - e->m[z]) / e->sigma;
                    for (int z = 0; z < D; ++z) {
                        double s = 0;
                        for (int zz = 0; zz < D; ++zz) s += e->B[(size_t)zz * D + z] * y[zz];
                        t[z] = s / e->Dd[z];
                    }
                    double nh = 0;
                    for (int z = 0; z < D; ++z) {
                        double s = 0;
                        for (int zz = 0; zz < D; ++zz) s += e->B[(size_t)z * D + zz] * t[zz];
                        nh += s * s;
                    }
                    v += e->cmue * e->weights[k] * (D / nh) * (e->X[(size_t)i * L + idx[k]] - e->m_prev[i]) *
                         (e->X[(size_t)j * L + idx[k]] - e->m_prev[j]) / s2;
                }
                C[(size_t)i * D + j] = C[(size_t)j * D + i] = v;
            }
    }
    std::vector<double> w, V;
    jacobi_eig(D, C, w, V);
    for (int i = 0; i < D; ++i) {
        if (w[i] <= 0) { e->stopped = 1; return 1; }
        e->Dd[i] = std::sqrt(w[i]);
    }
    e->B = V;
    std::default_random_engine gad;
    gad.seed(11 + e->it);
    std::normal_distribution<double> nd(0, 1.0);
    for (int k = 0; k < L; ++k) {
        for (int i = 0; i < D; ++i) tmp[i] = e->Dd[i] * nd(gad);
        for (int i = 0; i < D; ++i) {
            double s = 0;
            for (int j = 0; j < D; ++j) s += e->B[(size_t)i * D + j] * tmp[j];
            e->X[(size_t)i * L + k] = e->m[i] + e->sigma * s;
        }
    }
    for (int i = 0; i < D; ++i) {
        for (int j = 0; j < L; ++j) {
            const double v = e->X[(size_t)i * L + j];
            e->Xfeas[(size_t)i * L + j] = std::min(1.0, std::max(0.0, v));
            e->Xpref[(size_t)i * L + j] = std::min(1.0 - e->soft_edge, std::max(e->soft_edge, v));
        }
        e->m_feasible[i] = std::min(1.0 - e->soft_edge, std::max(e->soft_edge, e->m[i]));
    }
    e->tvec[e->it % e->early_gens] = e->best_val;
    if (e->it > e->early_gens) {
        const auto mm = std::minmax_element(e->tvec.begin(), e->tvec.end());
        if (*mm.second - *mm.first < e->early_tol) { e->stopped = 1; ++e->it; return 1; }
    }
    ++e->it;
    return 0;
}

// Best feasible point so far (unit box) and its raw cost.
int bnm_cmaes_best(bnm_cmaes* e, double* x, double* raw_cost) {
    if (!e) return BNM_E_ARG;
    if (x) std::copy(e->best_x.begin(), e->best_x.end(), x);
    if (raw_cost) *raw_cost = e->best_raw;
    return BNM_OK;
}

// Strategy parameters: [mueeff, cc, csigma, ccov, dsigma, chiN, c1, cmue,
// weights[0..lambda-1]]; returns the number of values written.
int bnm_cmaes_strategy(bnm_cmaes* e, double* out, int n) {
    if (!e || !out) return BNM_E_ARG;
    std::vector<double> v = {e->mueeff, e->cc, e->csigma, e->ccov, e->dsigma, e->chiN, e->c1, e->cmue};
    v.insert(v.end(), e->weights.begin(), e->weights.end());
    const int k = std::min(n, (int)v.size());
    std::copy(v.begin(), v.begin() + k, out);
    return k;
}

// Pre-selection support.  Draws a pool of n >= lambda candidates from the
// current distribution N(m, sigma^2 C): the first lambda are the current
// sample (what bnm_cmaes_ask returns), the other n - lambda come from a
// generator seeded with `seed`.  x_feas[j*D + i] receives the clipped points.
int bnm_cmaes_ask_pool(bnm_cmaes* e, int n, unsigned seed, double* x_feas) {
    if (!e || !x_feas || n < e->lambda) return BNM_E_ARG;
    const int D = e->D, L = e->lambda;
    e->pool.assign((size_t)n * D, 0.0);
    for (int j = 0; j < L; ++j)
        for (int i = 0; i < D; ++i) e->pool[(size_t)j * D + i] = e->X[(size_t)i * L + j];
    std::default_random_engine g;
    g.seed(seed);
    std::normal_distribution<double> nd(0, 1.0);
    std::vector<double> t(D);
    for (int j = L; j < n; ++j) {
        for (int i = 0; i < D; ++i) t[i] = e->Dd[i] * nd(g);
        for (int i = 0; i < D; ++i) {
            double s = 0;
            for (int k = 0; k < D; ++k) s += e->B[(size_t)i * D + k] * t[k];
            e->pool[(size_t)j * D + i] = e->m[i] + e->sigma * s;
        }
    }
    for (size_t q = 0; q < e->pool.size(); ++q) x_feas[q] = std::min(1.0, std::max(0.0, e->pool[q]));
    e->n_pool = n;
    return BNM_OK;
}

// Makes pool members idx[0..lambda) the current sample (the points the next
// bnm_cmaes_tell scores), e.g. the lambda best by a surrogate.
int bnm_cmaes_select(bnm_cmaes* e, const int* idx) {
    if (!e || !idx || e->n_pool < e->lambda) return BNM_E_ARG;
    const int D = e->D, L = e->lambda;
    for (int j = 0; j < L; ++j)
        if (idx[j] < 0 || idx[j] >= e->n_pool) return BNM_E_ARG;
    for (int j = 0; j < L; ++j)
        for (int i = 0; i < D; ++i) {
            const double v = e->pool[(size_t)idx[j] * D + i];
            e->X[(size_t)i * L + j] = v;
            e->Xfeas[(size_t)i * L + j] = std::min(1.0, std::max(0.0, v));
            e->Xpref[(size_t)i * L + j] = std::min(1.0 - e->soft_edge, std::max(e->soft_edge, v));
        }
    e->n_pool = 0;
    return BNM_OK;
}

double bnm_cmaes_sigma(bnm_cmaes* e) { return e ? e->sigma : -1; }
int bnm_cmaes_iteration(bnm_cmaes* e) { return e ? e->it : -1; }

}  // extern "C"
