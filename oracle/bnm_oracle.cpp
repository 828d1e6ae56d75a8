// bnm_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU, FP64 restatement of the reference's batched simulation path, used as
// the correctness checker for the HIP library.  Only tests/, the smoke() entry
// of __graft_entry__.py and bench.py's cpu_baseline leg may load it.  The
// product (bnm_cuda_amd/) never links, loads or calls anything in oracle/.
//
// What it restates (reference files cited per function below):
//   constants.h:24-116       model / FIC / Balloon-Windkessel constants
//   bnm.cpp:26-429           rWW Euler-Maruyama loop, BW, BOLD sampling, FIC
//   check_fit.cpp:19-217     FC / FCD lower triangles, goodness of fit
//   fic.cpp:71-207           analytical FIC (own Brent root finder)
//   run_CMAES.cpp:476-594    output_ts, pair table, window table
//   run_CMAES.cpp:640-657    noise array layout [ms][substep][node][E,I]
//
// Third-party pieces the reference calls but which are absent here:
//   GSL 2.7 gsl_stats_correlation  -> textbook one-pass co-moment update with
//                                     long double accumulators (pinned against
//                                     numpy.corrcoef in tests)
//   GSL 2.7 Brent fsolver           -> Brent (1973) zero algorithm, same
//                                     interval test (epsabs 0, epsrel 1e-3)
//   libks ks_stat (unpinned HEAD)   -> two-sample KS D = sup|F1-F2|, evaluated
//                                     at every data point (pinned against
//                                     scipy.stats.ks_2samp in tests)
//
// Build: see oracle/Makefile (g++ -O3 -fopenmp, no -ffast-math).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include <omp.h>

namespace {

// ---- constants.h:27-90 (FP64 build: u_real == double) ---------------------
const double kDt = 0.1;            // constants.h:27
const double kModelDt = 0.001;     // constants.h:28
const double kSqrtDt = 0.316227;   // constants.h:29 (truncated on purpose)
const double kJ_NMDA = 0.15;       // constants.h:33
const double kaE = 310, kbE = 125, kdE = 0.16;   // constants.h:34-37
const double kaI = 615, kbI = 177, kdI = 0.087;  // constants.h:38-40
const double kbaRatioE = kbE / kaE;              // constants.h:36
const double kGammaE = 0.641 / 1000.0;           // constants.h:41
const double kGammaI = 1.0 / 1000.0;             // constants.h:42
const double kTauE = 100, kTauI = 10;            // constants.h:43-44
const double kSigma = 0.01;                      // constants.h:47
const double kI0 = 0.382, kwE = 1.0, kwI = 0.7;  // constants.h:48-50
const double kwE_I0 = kwE * kI0;                 // constants.h:53
const double kwI_I0 = kwI * kI0;                 // constants.h:54
// FIC steady state (constants.h:57-68)
const double kTauI_s = 0.01, kGammaI_s = 1.0;
const double kI_E_ss = 0.3773805650, kS_E_ss = 0.1647572075;
// Balloon-Windkessel (constants.h:73-90, Demirtas 2019 3T k1..k3)
const double kRho = 0.34, kAlpha = 0.32, kTau = 0.98;
const double kY = 1.0 / 0.41, kKappa = 1.0 / 0.65, kV0 = 0.02 * 100;
const double kIAlpha = 1.0 / kAlpha, kITau = 1.0 / kTau, kOneMinRho = 1.0 - kRho;
// k1..k3: 3T values by default (constants.h:86-89); orc_set_bw_friston2003(1)
// switches to the BW_FRISTON2003 build's 7 rho, 2, 2 rho - 0.2 (:82-85)
double kK1 = 3.72, kK2 = 0.527, kK3 = 0.53;

inline double clip01(double v) { return std::max(0.0, std::min(v, 1.0)); }

// phi_I of fic.cpp:103-105
double phi_I(double II) {
    return ((kaI * II) - kbI) / (1 - std::exp(-1 * kdI * ((kaI * II) - kbI)));
}

// Brent's zero finder (R. P. Brent, "Algorithms for Minimization without
// Derivatives", 1973, ch. 4), driven like gsl_fsolve in fic.cpp:71-100:
// bracket [x_lo, x_hi], at most 100 iterations, stop when the bracket passes
// the interval test |hi-lo| < epsabs + epsrel*min(|lo|,|hi|) (epsabs 0,
// epsrel 1e-3; bracket straddling 0 uses tolerance epsabs).  Returns -1 on
// failure, as gsl_fsolve does.
template <class F>
double brent_root(F f, double x_lo, double x_hi, double epsrel, int max_iter) {
    double a = x_lo, b = x_hi;
    double fa = f(a), fb = f(b);
    if ((fa < 0.0 && fb < 0.0) || (fa > 0.0 && fb > 0.0)) return -1;
    double c = b, fc = fb, d = b - a, e = b - a;
    for (int iter = 1; iter <= max_iter; ++iter) {
        // keep the root bracketed by [b, c] with |f(b)| <= |f(c)|
        if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
            c = a; fc = fa; d = b - a; e = b - a;
        }
        if (std::fabs(fc) < std::fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tol = 0.5 * 2.220446049250313e-16 * std::fabs(b);
        const double m = 0.5 * (c - b);
        if (fb == 0) return b;
        if (std::fabs(m) <= tol) return b;
        if (std::fabs(e) < tol || std::fabs(fa) <= std::fabs(fb)) {
            d = m; e = m;  // bisection
        } else {
            double p, q, r, s = fb / fa;
            if (a == c) {  // secant
                p = 2 * m * s;
                q = 1 - s;
            } else {       // inverse quadratic interpolation
                q = fa / fc;
                r = fb / fc;
                p = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
                q = (q - 1) * (r - 1) * (s - 1);
            }
            if (p > 0) q = -q; else p = -p;
            if (2 * p < std::min(3 * m * q - std::fabs(tol * q), std::fabs(e * q))) {
                e = d; d = p / q;
            } else {
                d = m; e = m;
            }
        }
        a = b; fa = fb;
        if (std::fabs(d) > tol) b += d; else b += (m > 0 ? tol : -tol);
        fb = f(b);
        // keep [b, c] bracketing the root before the interval test
        if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
            c = a; fc = fa; d = b - a; e = b - a;
        }
        // interval test on the current bracket (lower, upper)
        const double lo = std::min(b, c), hi = std::max(b, c);
        const double abs_lo = std::fabs(lo), abs_hi = std::fabs(hi);
        const double min_abs = ((lo > 0 && hi > 0) || (lo < 0 && hi < 0)) ? std::min(abs_lo, abs_hi) : 0;
        const double tolerance = epsrel * min_abs;
        if (std::fabs(hi - lo) < tolerance) return b;
    }
    return -1;
}

// gsl_stats_correlation as called in check_fit.cpp:48,122,172: one-pass
// co-moment update (West 1979 / Welford) with long double accumulators.
double pearson(const double* x, size_t sx, const double* y, size_t sy, size_t n) {
    long double sum_xsq = 0, sum_ysq = 0, sum_cross = 0;
    long double mean_x = x[0], mean_y = y[0];
    for (size_t i = 1; i < n; ++i) {
        const long double ratio = i / (i + 1.0L);
        const long double dx = x[i * sx] - mean_x;
        const long double dy = y[i * sy] - mean_y;
        sum_xsq += dx * dx * ratio;
        sum_ysq += dy * dy * ratio;
        sum_cross += dx * dy * ratio;
        mean_x += dx / (i + 1.0L);
        mean_y += dy / (i + 1.0L);
    }
    return (double)(sum_cross / (std::sqrt(sum_xsq) * std::sqrt(sum_ysq)));
}

double mean_of(const double* x, size_t n) {
    // gsl_stats_mean: running mean update
    long double m = 0;
    for (size_t i = 0; i < n; ++i) m += (x[i] - m) / (i + 1);
    return (double)m;
}

int n_pairs_of(int nodes, int exc) {
    int n = nodes * (nodes - 1) / 2;
    if (exc) n -= (nodes / 2) * (nodes / 2);
    return n;
}

}  // namespace

extern "C" {

// Selects the BOLD weights for every later call (not thread-safe: set it
// before running simulations).
void orc_set_bw_friston2003(int on) {
    if (on) {
        kK1 = 7 * kRho;
        kK2 = 2.0;
        kK3 = 2 * kRho - 0.2;
    } else {
        kK1 = 3.72;
        kK2 = 0.527;
        kK3 = 0.53;
    }
}

// ---- index tables ---------------------------------------------------------

// run_CMAES.cpp:476-477 (u_real == double): int(T / (TR_s / model_dt)) + 1
int orc_output_ts(int time_steps, int bold_tr) {
    double TR = (double)bold_tr / 1000;
    return (int)(time_steps / (TR / kModelDt)) + 1;
}

// run_CMAES.cpp:488-524 / check_fit.cpp:24-32: i>j row-major lower triangle,
// skipping left-right pairs when exc_interhemispheric (assumes even N, L then R)
int orc_pairs(int nodes, int exc, int* pairs_i, int* pairs_j) {
    if (exc && (nodes % 2)) return -1;
    const int rh = nodes / 2;
    int idx = 0;
    for (int i = 0; i < nodes; ++i)
        for (int j = 0; j < i; ++j) {
            if (exc && ((i < rh) ^ (j < rh))) continue;
            if (pairs_i) { pairs_i[idx] = i; pairs_j[idx] = j; }
            ++idx;
        }
    return idx;
}

// run_CMAES.cpp:529-594 (drop_edges = true): window centres from
// n_vols_remove + ws/2 to n_vols_remove + corr_len - 1 - ws/2 in steps of
// `step`; window = [c - ws/2, c + ws/2] clipped to [0, output_ts - 1].
// Indices are absolute volume indices.  starts/ends may be NULL (count only).
int orc_windows(int output_ts, int n_vols_remove, int window_size, int window_step,
                int* starts, int* ends) {
    const int corr_len = output_ts - n_vols_remove;
    int first = window_size / 2 + n_vols_remove;
    int last = corr_len - 1 - window_size / 2 + n_vols_remove;
    int n = 0;
    for (int c = first; c <= last; c += window_step) {
        int s = c - window_size / 2, e = c + window_size / 2;
        if (s < 0) s = 0;
        if (e >= output_ts) e = output_ts - 1;
        if (starts) { starts[n] = s; ends[n] = e; }
        ++n;
    }
    return n;
}

// ---- noise (run_CMAES.cpp:640-657; bnm.cpp:39-43) ------------------------

// `count` N(0,1) floats from std::mt19937(seed) + std::normal_distribution<float>,
// in draw order.  Index ((ms*10 + substep)*N + node)*2 + {0:E, 1:I}.
void orc_noise(unsigned seed, long long count, float* out) {
    std::mt19937 gen(seed);
    std::normal_distribution<float> nd(0, 1);
    for (long long i = 0; i < count; ++i) out[i] = nd(gen);
}

// raw mt19937 output, for the C++-standard known answer (10000th value of the
// default seed 5489 is 4123659995)
unsigned orc_mt19937_nth(unsigned seed, long long n) {
    std::mt19937 gen(seed);
    unsigned v = 0;
    for (long long i = 0; i < n; ++i) v = gen();
    return v;
}

// ---- regional parameters (run_CMAES.cpp:61-143) --------------------------

// p_j = base * (1 + sum_m scale_m * map_m[j]); shift all so that min >= 0.001
void orc_regional_params(int N, int n_maps, const double* maps, double base, const double* scales,
                         double* out) {
    double min_p = 1000, max_p = -1000;
    for (int j = 0; j < N; ++j) {
        double scaler = 1;
        for (int m = 0; m < n_maps; ++m) scaler += scales[m] * maps[(size_t)m * N + j];
        out[j] = base * scaler;
        min_p = std::min(min_p, out[j]);
        max_p = std::max(max_p, out[j]);
    }
    if (min_p < 0.001)
        for (int j = 0; j < N; ++j) out[j] -= min_p - 0.001;
}

// ---- analytical FIC (fic.cpp:130-207) -----------------------------------

// Root of the inhibitory-current fixed point (fic.cpp:115-119) for one node.
double orc_fic_inh_root(double w_EI) {
    auto f = [&](double x) {
        return kwI_I0 + w_EI * kS_E_ss - 1.0 * kGammaI_s * kTauI_s * phi_I(x) - x;
    };
    return brent_root(f, 0.0, 2.0, 0.001, 100);
}

// Per-node wIE that keeps r_E ~ 3 Hz.  SC row-major [N*N] (row = target node),
// G, wEE/wEI per node.  Writes wIE[N]; returns 1 (and writes nothing) when the
// root finder fails or any J < 0 ("unstable" in fic.cpp:183-186, 199-202).
int orc_analytical_fic(int N, const double* SC, double G, const double* wEE,
                       const double* wEI, double* wIE) {
    std::vector<double> out(N);
    for (int j = 0; j < N; ++j) {
        const double II = orc_fic_inh_root(wEI[j]);
        if (II == -1) return 1;
        const double rI = phi_I(II);
        const double SI = rI * kTauI_s * kGammaI_s;
        // K_EE = G*J_NMDA*SC + diag(wEE) (fic.cpp:148-152); row j . S_E_ss
        double dot = 0;
        for (int k = 0; k < N; ++k) {
            double kee = SC[(size_t)j * N + k] * (G * kJ_NMDA);
            if (k == j) kee += wEE[j];
            dot += kee * kS_E_ss;
        }
        const double J = (-1 / SI) * (kI_E_ss - 0.0 - kwE_I0 - dot);
        if (J < 0) return 1;
        out[j] = J;
    }
    std::copy(out.begin(), out.end(), wIE);
    return 0;
}

// ---- one simulation (bnm.cpp:26-290) -------------------------------------

// Integrates one simulation with the noise taken from `noise`, which holds
// `noise_period_ms` ms of draws in [ms][10][N][2] order and is indexed with
// (ts mod period); period >= time_steps+1 reproduces bnm.cpp's stream.  If
// `noise` is NULL, draws are made on the fly from mt19937(seed) exactly as in
// bnm.cpp (re-seeded on each FIC restart).
//
// wIE [N] is the initial wIE and receives the adjusted wIE.  bold
// [output_ts*N] (time-major) may be NULL.  ext_means [6*N] receives the means
// of S_E, I_E, r_E, S_I, I_I, r_I over TR samples with vol >= n_vols_remove.
// Returns the number of BOLD samples written (== output_ts).
int orc_simulate(int N, const double* SC, double G, const double* wEE, const double* wEI,
                 double* wIE, const float* noise, long long noise_period_ms, unsigned seed,
                 int time_steps, int bold_tr, int n_vols_remove,
                 int adjust_fic, int max_fic_trials, int i_start, int i_end, double init_delta,
                 double* bold, double* ext_means, int* fic_failed, int* fic_trials) {
    std::vector<double> S_E(N, 0.001), S_I(N, 0.001), S_prev(N), w_IE(wIE, wIE + N);
    std::vector<double> bx(N, 0.0), bf(N, 1.0), bnu(N, 1.0), bq(N, 1.0);
    std::vector<double> rE(N, 0), rI(N, 0), IE(N, 0), II(N, 0);
    std::vector<double> meanIE(N, 0), delta(N, init_delta);
    std::vector<double> ext(6 * (size_t)N, 0.0);
    const unsigned i_dur = (unsigned)(i_end - i_start + 1);
    std::mt19937 gen(seed);
    std::normal_distribution<float> nd(0, 1);
    int trial = 0, vol = 0, ts = 0;
    bool failed = false, adjusting = adjust_fic != 0;
    while (ts <= time_steps) {
        const long long ms = noise ? (ts % noise_period_ms) : 0;
        for (int s = 0; s < 10; ++s) {
            for (int j = 0; j < N; ++j) S_prev[j] = S_E[j];            // bnm.cpp:171-173
            for (int j = 0; j < N; ++j) {
                double cin = 0;                                       // bnm.cpp:176-179
                const double* row = SC + (size_t)j * N;
                for (int k = 0; k < N; ++k) cin += row[k] * S_prev[k];
                IE[j] = kwE_I0 + wEE[j] * S_E[j] + cin * G * kJ_NMDA - w_IE[j] * S_I[j];
                II[j] = kwI_I0 + wEI[j] * S_E[j] - S_I[j];
                const double eE = std::exp(-kdE * (kaE * IE[j] - kbE));
                const double eI = std::exp(-kdI * (kaI * II[j] - kbI));
                rE[j] = (kaE * IE[j] - kbE) / (1 - eE);
                rI[j] = (kaI * II[j] - kbI) / (1 - eI);
                double xE, xI;
                if (noise) {
                    const size_t base = (((size_t)ms * 10 + s) * N + j) * 2;
                    xE = noise[base];
                    xI = noise[base + 1];
                } else {
                    xE = nd(gen);
                    xI = nd(gen);
                }
                const double dSE = xE * kSigma * kSqrtDt + kDt * ((1 - S_E[j]) * kGammaE * rE[j] - (S_E[j] / kTauE));
                const double dSI = xI * kSigma * kSqrtDt + kDt * (kGammaI * rI[j] - (S_I[j] / kTauI));
                S_E[j] = clip01(S_E[j] + dSE);
                S_I[j] = clip01(S_I[j] + dSI);
            }
        }
        for (int j = 0; j < N; ++j) {                                 // bnm.cpp:205-211
            bx[j] = bx[j] + kModelDt * (S_E[j] - kKappa * bx[j] - kY * (bf[j] - 1.0));
            const double f_new = bf[j] + kModelDt * bx[j];
            bnu[j] = bnu[j] + kModelDt * kITau * (bf[j] - std::pow(bnu[j], kIAlpha));
            bq[j] = bq[j] + kModelDt * kITau *
                    (bf[j] * (1.0 - std::pow(kOneMinRho, (1.0 / bf[j]))) / kRho -
                     std::pow(bnu[j], kIAlpha) * bq[j] / bnu[j]);
            bf[j] = f_new;
        }
        if (ts % bold_tr == 0) {                                      // bnm.cpp:212-226
            for (int j = 0; j < N; ++j) {
                if (bold)
                    bold[(size_t)vol * N + j] = kV0 * (kK1 * (1 - bq[j]) + kK2 * (1 - bq[j] / bnu[j]) + kK3 * (1 - bnu[j]));
                if (vol >= n_vols_remove) {
                    ext[0 * (size_t)N + j] += S_E[j];
                    ext[1 * (size_t)N + j] += IE[j];
                    ext[2 * (size_t)N + j] += rE[j];
                    ext[3 * (size_t)N + j] += S_I[j];
                    ext[4 * (size_t)N + j] += II[j];
                    ext[5 * (size_t)N + j] += rI[j];
                }
            }
            ++vol;
        }
        ++ts;
        if (adjusting) {                                              // bnm.cpp:229-289
            if ((unsigned)ts >= (unsigned)i_start && (unsigned)ts <= (unsigned)i_end)
                for (int j = 0; j < N; ++j) meanIE[j] += IE[j];
            if (ts == i_end) {
                bool needs = false;
                for (int j = 0; j < N; ++j) {
                    meanIE[j] /= i_dur;
                    const double diff = meanIE[j] - kbaRatioE;
                    if (std::fabs(diff + 0.026) > 0.005) {
                        needs = true;
                        if (trial < max_fic_trials) {
                            if (diff < -0.026) {
                                w_IE[j] -= delta[j];
                                delta[j] -= 0.001;
                                delta[j] = fmaxf((float)delta[j], 0.001f);  // bnm.cpp:250
                            } else {
                                w_IE[j] += delta[j];
                            }
                        }
                    }
                }
                if (needs) {
                    if (trial < max_fic_trials) {
                        for (int j = 0; j < N; ++j) {
                            S_E[j] = 0.001; S_I[j] = 0.001;
                            bx[j] = 0.0; bf[j] = 1.0; bnu[j] = 1.0; bq[j] = 1.0;
                            meanIE[j] = 0;
                        }
                        std::fill(ext.begin(), ext.end(), 0.0);
                        ts = 0;
                        vol = 0;
                        gen = std::mt19937(seed);
                        nd = std::normal_distribution<float>(0, 1);
                        ++trial;
                    } else {
                        failed = true;
                        adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
    }
    const int n_ext = vol - n_vols_remove;
    if (ext_means)
        for (size_t i = 0; i < ext.size(); ++i) ext_means[i] = n_ext > 0 ? ext[i] / n_ext : 0.0;
    std::copy(w_IE.begin(), w_IE.end(), wIE);
    if (fic_failed) *fic_failed = failed ? 1 : 0;
    if (fic_trials) *fic_trials = trial;
    return vol;
}

// ---- FC / FCD / KS / fit (check_fit.cpp) -------------------------------

double orc_pearson(const double* x, const double* y, long long n) {
    return pearson(x, 1, y, 1, (size_t)n);
}

// FC lower triangle of bold [n_vols x N] (check_fit.cpp:19-62).  Returns 0,
// or 1 if any correlation is NaN.
int orc_fc_tril(const double* bold, int n_vols, int N, int exc, double* out) {
    const int rh = N / 2;
    int idx = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < i; ++j) {
            if (exc && ((i < rh) ^ (j < rh))) continue;
            const double r = pearson(bold + i, N, bold + j, N, n_vols);
            if (std::isnan(r)) return 1;
            out[idx++] = r;
        }
    return 0;
}

// FCD lower triangle (check_fit.cpp:64-136, drop_edges = true) of bold
// [n_vols x N] (volumes after n_vols_remove).  window_fc (optional) receives
// [n_windows x n_pairs].  Returns 0, or 1 if any correlation is NaN.
int orc_fcd_tril(const double* bold, int n_vols, int N, int step, int window_size, int exc,
                 double* fcd_out, double* window_fc) {
    const int np = n_pairs_of(N, exc);
    std::vector<int> st(n_vols + 1), en(n_vols + 1);
    const int nw = orc_windows(n_vols, 0, window_size, step, st.data(), en.data());
    std::vector<double> wfc((size_t)nw * np);
    for (int w = 0; w < nw; ++w) {
        const int len = en[w] - st[w] + 1;
        if (orc_fc_tril(bold + (size_t)st[w] * N, len, N, exc, wfc.data() + (size_t)w * np)) return 1;
    }
    if (window_fc) std::copy(wfc.begin(), wfc.end(), window_fc);
    int idx = 0;
    for (int wi = 0; wi < nw; ++wi)
        for (int wj = 0; wj < wi; ++wj) {
            const double r = pearson(wfc.data() + (size_t)wi * np, 1, wfc.data() + (size_t)wj * np, 1, np);
            if (std::isnan(r)) return 1;
            fcd_out[idx++] = r;
        }
    return 0;
}

// The window FCs of orc_fcd_tril alone ([n_windows x n_pairs]; the windows
// are independent, so they are spread over OpenMP threads).  Returns the
// window count, or -1 if any correlation is NaN.
int orc_window_fc(const double* bold, int n_vols, int N, int step, int window_size, int exc,
                  double* window_fc) {
    const int np = n_pairs_of(N, exc);
    std::vector<int> st(n_vols + 1), en(n_vols + 1);
    const int nw = orc_windows(n_vols, 0, window_size, step, st.data(), en.data());
    int bad = 0;
#pragma omp parallel for schedule(dynamic) reduction(|| : bad)
    for (int w = 0; w < nw; ++w) {
        const int len = en[w] - st[w] + 1;
        if (orc_fc_tril(bold + (size_t)st[w] * N, len, N, exc, window_fc + (size_t)w * np)) bad = 1;
    }
    return bad ? -1 : nw;
}

// Two-sample Kolmogorov-Smirnov statistic D = sup_x |F_x(x) - F_y(x)|,
// evaluated at every data point of both samples (libks ks_stat call site:
// check_fit.cpp:204-209, bnm.cu:736-741).
double orc_ks(const double* x, long long nx, const double* y, long long ny) {
    std::vector<double> a(x, x + nx), b(y, y + ny);
    std::sort(a.begin(), a.end());
    std::sort(b.begin(), b.end());
    double d = 0;
    auto eval = [&](double v) {
        const double fa = (double)(std::upper_bound(a.begin(), a.end(), v) - a.begin()) / nx;
        const double fb = (double)(std::upper_bound(b.begin(), b.end(), v) - b.begin()) / ny;
        d = std::max(d, std::fabs(fa - fb));
    };
    for (double v : a) eval(v);
    for (double v : b) eval(v);
    return d;
}

// check_fit_bold (check_fit.cpp:138-217): cost = 1 - r(emp_FC, sim_FC)
// [+ |mean(emp_FC) - mean(sim_FC)|] [+ KS(emp_FCD, sim_FCD)], or max_cost if
// a NaN appears.  bold is [n_vols x N] after n_vols_remove.  fc_out / fcd_out
// (optional) receive the simulated triangles.  Returns -1 on a size mismatch.
double orc_check_fit(const double* bold, int n_vols, int N, int step, int window_size,
                     const double* emp_fc, int n_emp_fc, const double* emp_fcd, int n_emp_fcd,
                     int no_fcd, int exc, int use_fc_diff, double max_cost,
                     double* fc_out, double* fcd_out) {
    const int np = n_pairs_of(N, exc);
    if (np != n_emp_fc) return -1;
    std::vector<double> fc(np);
    if (orc_fc_tril(bold, n_vols, N, exc, fc.data())) return max_cost;
    if (fc_out) std::copy(fc.begin(), fc.end(), fc_out);
    double cost = 1 - pearson(emp_fc, 1, fc.data(), 1, np);
    if (use_fc_diff) cost += std::fabs(mean_of(emp_fc, np) - mean_of(fc.data(), np));
    if (!no_fcd) {
        const int nw = orc_windows(n_vols, 0, window_size, step, nullptr, nullptr);
        std::vector<double> fcd((size_t)nw * (nw - 1) / 2 + 1);
        if (orc_fcd_tril(bold, n_vols, N, step, window_size, exc, fcd.data(), nullptr)) return max_cost;
        const long long nfcd = (long long)nw * (nw - 1) / 2;
        if (fcd_out) std::copy(fcd.begin(), fcd.begin() + nfcd, fcd_out);
        cost += orc_ks(emp_fcd, n_emp_fcd, fcd.data(), nfcd);
    }
    return cost;
}

// FIC penalty (bnm.cpp:390-405; bnm.cu:752-767)
double orc_fic_penalty(const double* mean_rE, int N, double scale, int reject_failed, int failed) {
    double p = 0;
    for (int j = 0; j < N; ++j) {
        const double d = std::fabs(mean_rE[j] - 3);
        if (d > 1) p += 1 - std::exp(-0.05 * (d - 1));
    }
    p *= scale;
    p /= N;
    if (reject_failed && failed) p += 1;
    return p;
}

// ---- batch driver (bnm.cpp:431-466: one simulation per OpenMP thread) -----

// Runs B simulations sharing SC and noise.  Parameters per sim: G[B],
// wEE/wEI/wIE[B*N].  When do_fic, wIE_fic[B*N] holds the initial wIE (the
// analytical FIC is computed here when analytical_init != 0) and receives the
// adjusted wIE; otherwise wIE is used.  Outputs: corr[B] (goodness-of-fit
// cost without penalty), fic_pen[B], and optionally bold[B*output_ts*N],
// fc[B*n_pairs], fcd[B*n_window_pairs], ext[B*6*N], failed[B], trials[B].
// n_threads <= 0 uses OpenMP's default.
int orc_run_simulations(int B, int N, const double* SC, const double* G, const double* wEE,
                        const double* wEI, const double* wIE, double* wIE_fic, int do_fic,
                        int analytical_init, const float* noise, long long noise_period_ms,
                        int time_steps, int bold_tr, int n_vols_remove, int adjust_fic,
                        int max_fic_trials, int i_start, int i_end, double init_delta,
                        int window_size, int window_step, int exc, int no_fcd, int use_fc_diff,
                        double max_cost, double fic_scale, int fic_reject, int sim_only,
                        const double* emp_fc, int n_emp_fc, const double* emp_fcd, int n_emp_fcd,
                        double* corr, double* fic_pen, double* bold, double* fc, double* fcd,
                        double* ext, int* failed, int* trials, int n_threads) {
    const int ots = orc_output_ts(time_steps, bold_tr);
    const int np = n_pairs_of(N, exc);
    const int corr_len = ots - n_vols_remove;
    const int nw = orc_windows(corr_len, 0, window_size, window_step, nullptr, nullptr);
    const long long nwp = (long long)nw * (nw - 1) / 2;
    if (n_threads > 0) omp_set_num_threads(n_threads);
    int status = 0;
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < B; ++b) {
        std::vector<double> w(N), bl((size_t)ots * N), e(6 * (size_t)N);
        const double* wee = wEE + (size_t)b * N;
        const double* wei = wEI + (size_t)b * N;
        if (do_fic) {
            if (analytical_init) {
                if (orc_analytical_fic(N, SC, G[b], wee, wei, w.data()))
                    std::fill(w.begin(), w.end(), 1.0);
            } else {
                std::copy(wIE_fic + (size_t)b * N, wIE_fic + (size_t)(b + 1) * N, w.begin());
            }
        } else {
            std::copy(wIE + (size_t)b * N, wIE + (size_t)(b + 1) * N, w.begin());
        }
        int f = 0, t = 0;
        orc_simulate(N, SC, G[b], wee, wei, w.data(), noise, noise_period_ms, 0, time_steps, bold_tr,
                     n_vols_remove, do_fic && adjust_fic, max_fic_trials, i_start, i_end, init_delta,
                     bl.data(), e.data(), &f, &t);
        if (do_fic) std::copy(w.begin(), w.end(), wIE_fic + (size_t)b * N);
        if (bold) std::copy(bl.begin(), bl.end(), bold + (size_t)b * ots * N);
        if (ext) std::copy(e.begin(), e.end(), ext + (size_t)b * 6 * N);
        if (failed) failed[b] = f;
        if (trials) trials[b] = t;
        if (fic_pen) fic_pen[b] = orc_fic_penalty(e.data() + 2 * (size_t)N, N, fic_scale, fic_reject, f);
        if (sim_only) {
            corr[b] = max_cost;
        } else {
            const double c = orc_check_fit(bl.data() + (size_t)n_vols_remove * N, corr_len, N, window_step,
                                           window_size, emp_fc, n_emp_fc, emp_fcd, n_emp_fcd, no_fcd, exc,
                                           use_fc_diff, max_cost, fc ? fc + (size_t)b * np : nullptr,
                                           (fcd && !no_fcd) ? fcd + (size_t)b * nwp : nullptr);
            if (c == -1) status = -1;
            corr[b] = c;
        }
    }
    return status;
}

}  // extern "C"
