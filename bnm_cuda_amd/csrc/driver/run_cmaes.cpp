// run_cmaes.cpp -- CMA-ES fitting of the reduced Wong-Wang network to an
// empirical FC / FCD on the GPU: the command-line program of the reference's
// run_CMAES.cpp (GPU build, main at run_CMAES.cpp:176-906) over libbnm_hip.
//
// Same 19 positional arguments, input files, BNM_* environment, output
// folders, log lines and best-simulation text files as the reference:
//   run_cmaes <sc_path> <out_prefix|'same'> <fc_tril_path> <fcd_tril_path|no_fcd>
//             <maps_path|no_maps> <n_regions> <G_min[-G_max]> <wee_min[-wee_max]>
//             <wei_min[-wei_max]> <wie_min[-wie_max]> <het_params> <time_steps>
//             <BOLD_TR> <fcd_step> <fcd_window> <rand_seed> <lambda> <itMax> <SeedMW>
//
// Additions (environment):
//   BNM_PRECISION      f64 (default, as the reference's double build) | f32 | bf16x3
//   BNM_NOISE_CHUNKED  1: parallel chunked noise generator (a different stream)
//   BNM_SURROGATE      1: surrogate-guided search (below)
//   BNM_PROGRESS_CSV   1 (default): per-generation progress file next to the log
//   BNM_SURR_FIC       0: surrogate search without the learned FIC start
//
// Surrogate-guided search (BNM_SURROGATE=1):
//  * FIC initialiser: an MLP (inputs G, wEE_j, wEI_j, SC strength_j, z_j =
//    the first map's value at node j, or 0 without maps; 4 x 256 GELU; BF16
//    MFMA) predicts the converged wIE_j minus the analytical one.  With
//    BNM_SURR_FIC_MODEL=<file> it is the head trained offline on the
//    200k-simulation training set (bnm_cuda_amd/trainset.py); otherwise,
//    before the search, BNM_SURR_FIC_SIMS parameter sets drawn uniformly from
//    the search box run the FIC phase only (BNM_SURR_FIC_TRIALS restarts
//    allowed) and train it.  Every simulation of the search then starts FIC
//    from analytical + predicted wIE.
//  * online adaptation (BNM_SURR_FIC_ADAPT=1, default): the head is trained
//    further on the converged wIE of the search's own simulations (the last
//    BNM_SURR_FIC_WINDOW generations, BNM_SURR_FIC_STEPS Adam steps per
//    generation); the first generation starts from the analytical wIE, and
//    a generation uses the learned start only when the head's mean distance
//    to the converged wIE of the previous generation was below
//    BNM_SURR_FIC_GATE (0.9) times the analytical start's.
//  * plain-consistency guard (with the learned start): the jump from the
//    analytical to the learned start stands for k plain restarts (+init_delta
//    per restart upwards; init_delta, then 0.001 less per restart, downwards);
//    a candidate whose learned-start run failed, or whose k plus the restarts
//    it still took reach the plain budget, might have failed FIC with the
//    plain method, so it is simulated again with the plain method and that
//    result is what the search sees (the re-runs are counted as full
//    simulations).  BNM_SURR_GUARD=0 turns the guard off.
//  * fit surrogate: an MLP on the unit-box coordinates predicts
//    (FC r, FCD KS, |mean FC difference|, FIC penalty, cost) of a candidate.
//    It is trained on the simulations of the last BNM_SURR_WINDOW
//    generations, inputs and outputs standardised over them.  From generation
//    BNM_SURR_WARMUP + 1 on, each generation draws BNM_SURR_POOL x lambda
//    candidates, ranks them by predicted cost and simulates the lambda best in
//    full; when the rank correlation of predicted and simulated cost of the
//    previous generation falls below BNM_SURR_MIN_RHO, the next generation is
//    the plain CMA-ES sample (fallback).
//  * BNM_SURR_MODEL=1 ranks the pool with a local linear-quadratic model
//    instead (ridge least squares over the last BNM_SURR_LQ_POINTS simulated
//    points, default max(2 x terms, 2 x lambda)); 2 ranks by the sum of the
//    MLP's and the quadratic model's ranks.
//  * BNM_SURR_ENSEMBLE=k trains k fit MLPs (different initialisation and
//    shuffling seeds) on the same window and ranks by their mean predicted
//    cost; BNM_SURR_UNSCREENED=f (default 0.25) takes round(f x lambda) of
//    each pre-screened generation at random from the pool (an unscreened
//    sample of the CMA-ES distribution) and the rest by predicted cost.
//  * every reported cost is a full simulation; the best point is re-run with
//    the plain method (analytical FIC init) at the end, as in the reference.
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "../../../include/bnm_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

double env_d(const char* k, double d) {
    const char* v = std::getenv(k);
    return v ? std::atof(v) : d;
}
int env_i(const char* k, int d) {
    const char* v = std::getenv(k);
    return v ? std::atoi(v) : d;
}
std::string env_s(const char* k, const char* d) {
    const char* v = std::getenv(k);
    return v ? std::string(v) : std::string(d);
}

[[noreturn]] void die(const std::string& m) {
    std::fprintf(stderr, "%s\n", m.c_str());
    std::exit(1);
}

void check(int st, const char* what) {
    if (st < 0) die(std::string(what) + " failed: " + bnm_last_error());
}

std::vector<double> read_values(const std::string& path, const char* what) {
    std::ifstream f(path);
    if (!f) die(std::string("Error opening ") + what + " file");
    std::vector<double> v;
    double x;
    while (f >> x) v.push_back(x);
    return v;
}

// Spearman rank correlation (average ranks for ties)
double spearman(const std::vector<double>& a, const std::vector<double>& b) {
    const size_t n = a.size();
    if (n < 3) return 0;
    auto ranks = [&](const std::vector<double>& v) {
        std::vector<size_t> idx(n);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t i, size_t j) { return v[i] < v[j]; });
        std::vector<double> r(n);
        for (size_t i = 0; i < n;) {
            size_t j = i;
            while (j + 1 < n && v[idx[j + 1]] == v[idx[i]]) ++j;
            for (size_t k = i; k <= j; ++k) r[idx[k]] = 0.5 * (i + j);
            i = j + 1;
        }
        return r;
    };
    const auto ra = ranks(a), rb = ranks(b);
    const double m = 0.5 * (n - 1);
    double sab = 0, saa = 0, sbb = 0;
    for (size_t i = 0; i < n; ++i) {
        sab += (ra[i] - m) * (rb[i] - m);
        saa += (ra[i] - m) * (ra[i] - m);
        sbb += (rb[i] - m) * (rb[i] - m);
    }
    return saa > 0 && sbb > 0 ? sab / std::sqrt(saa * sbb) : 0;
}

// Pearson correlation of two equally long vectors
double pearson(const double* a, const double* b, size_t n) {
    double ma = 0, mb = 0;
    for (size_t i = 0; i < n; ++i) { ma += a[i]; mb += b[i]; }
    ma /= n; mb /= n;
    double sab = 0, saa = 0, sbb = 0;
    for (size_t i = 0; i < n; ++i) {
        sab += (a[i] - ma) * (b[i] - mb);
        saa += (a[i] - ma) * (a[i] - ma);
        sbb += (b[i] - mb) * (b[i] - mb);
    }
    return sab / std::sqrt(saa * sbb);
}

// printf to the log file and to stdout
__attribute__((format(printf, 2, 3))) void both(FILE* log, const char* fmt, ...) {
    va_list a, b;
    va_start(a, fmt);
    va_copy(b, a);
    std::vfprintf(log, fmt, a);
    std::vprintf(fmt, b);
    va_end(b);
    va_end(a);
}

// Feature / target standardisation for the MLPs
struct Scaler {
    std::vector<double> mu, sd;
    void fit(const std::vector<float>& X, int n, int d) {
        mu.assign(d, 0);
        sd.assign(d, 0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < d; ++k) mu[k] += X[(size_t)i * d + k];
        for (int k = 0; k < d; ++k) mu[k] /= n;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < d; ++k) {
                const double e = X[(size_t)i * d + k] - mu[k];
                sd[k] += e * e;
            }
        for (int k = 0; k < d; ++k) sd[k] = std::sqrt(sd[k] / n) + 1e-8;
    }
    void apply(std::vector<float>& X, int d) const {
        for (size_t i = 0; i < X.size(); ++i) X[i] = (float)((X[i] - mu[i % d]) / sd[i % d]);
    }
};

// Plain FIC restarts that move a node's wIE by d (fic_decide, bnm.cpp:239-257):
// +delta per restart upwards (delta unchanged), delta then delta - 0.001 ...
// (at least 0.001) per restart downwards.
int plain_fic_steps(double d, double delta0) {
    if (d >= 0) return (int)std::ceil(d / delta0 - 1e-12);
    double moved = 0, delta = delta0;
    int n = 0;
    while (moved < -d - 1e-12 && n < 100000) {
        moved += delta;
        delta = std::max(delta - 0.001, 0.001);
        ++n;
    }
    return n;
}

// FIC head file written by bnm_cuda_amd/trainset.py (FICHead.save), little
// endian: "BNMFIC01", int32 in_dim / hidden / depth / out_dim / bf16, float32
// mu[in_dim], sd[in_dim], ysd, int64 n_params, float32 params.
bool load_fic_head(const std::string& path, bnm_mlp** net, Scaler& xs, double* ys, std::string* err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) { *err = "cannot open " + path; return false; }
    char magic[8];
    int32_t h[5];
    if (!f.read(magic, 8) || std::memcmp(magic, "BNMFIC01", 8) != 0 || !f.read((char*)h, sizeof h)) {
        *err = path + " is not a FIC head file";
        return false;
    }
    if (h[0] != 5 || h[3] != 1) { *err = "FIC head must map 5 features to 1 output"; return false; }
    std::vector<float> mu(5), sd(5);
    float ysd = 1;
    int64_t n = 0;
    f.read((char*)mu.data(), 20);
    f.read((char*)sd.data(), 20);
    f.read((char*)&ysd, 4);
    f.read((char*)&n, 8);
    std::vector<float> p(n > 0 && n < (1LL << 28) ? n : 0);
    if (!f || p.empty() || !f.read((char*)p.data(), (std::streamsize)(4 * n))) {
        *err = path + " is truncated";
        return false;
    }
    int st = 0;
    *net = bnm_mlp_create(5, h[1], h[2], 1, h[4], 7, &st);
    if (!*net || bnm_mlp_num_params(*net) != n) { *err = "FIC head shape does not match"; return false; }
    if (bnm_mlp_set_params(*net, p.data())) { *err = bnm_mlp_last_error(); return false; }
    xs.mu.assign(mu.begin(), mu.end());
    xs.sd.assign(sd.begin(), sd.end());
    *ys = ysd;
    return true;
}

// Local linear-quadratic cost model for pre-screening (the surrogate of
// Hansen's lq-CMA-ES, 2019): cost ~ full quadratic in the unit-box
// coordinates, centred and scaled on the fitted points, by ridge least
// squares; linear when there are too few points for the quadratic terms.
struct LqModel {
    int D = 0, nf = 0;
    bool ok = false;
    std::vector<double> mu, sd, beta;
    void features(const double* x, std::vector<double>& f) const {
        f.assign(nf, 0.0);
        f[0] = 1.0;
        std::vector<double> z(D);
        for (int k = 0; k < D; ++k) z[k] = (x[k] - mu[k]) / sd[k];
        for (int k = 0; k < D; ++k) f[1 + k] = z[k];
        if (nf > 1 + D) {
            int q = 1 + D;
            for (int a = 0; a < D; ++a)
                for (int b = a; b < D; ++b) f[q++] = z[a] * z[b];
        }
    }
    // rows x[i*D..], targets y[i]
    void fit(const std::vector<double>& x, const std::vector<double>& y, int d) {
        D = d;
        const int n = (int)y.size();
        const int nq = 1 + D + D * (D + 1) / 2;
        nf = n >= nq + D ? nq : 1 + D;
        ok = n > nf;
        if (!ok) return;
        mu.assign(D, 0.0);
        sd.assign(D, 0.0);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < D; ++k) mu[k] += x[(size_t)i * D + k] / n;
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < D; ++k) sd[k] += std::pow(x[(size_t)i * D + k] - mu[k], 2) / n;
        for (int k = 0; k < D; ++k) sd[k] = std::sqrt(sd[k]) + 1e-12;
        std::vector<double> A((size_t)nf * nf, 0.0), r(nf, 0.0), f;
        for (int i = 0; i < n; ++i) {
            features(&x[(size_t)i * D], f);
            for (int a = 0; a < nf; ++a) {
                r[a] += f[a] * y[i];
                for (int b = 0; b < nf; ++b) A[(size_t)a * nf + b] += f[a] * f[b];
            }
        }
        double tr = 0;
        for (int a = 0; a < nf; ++a) tr += A[(size_t)a * nf + a];
        for (int a = 0; a < nf; ++a) A[(size_t)a * nf + a] += 1e-8 * tr / nf;
        // Cholesky solve
        for (int j = 0; j < nf; ++j) {
            double dsum = A[(size_t)j * nf + j];
            for (int k = 0; k < j; ++k) dsum -= A[(size_t)j * nf + k] * A[(size_t)j * nf + k];
            if (!(dsum > 0)) { ok = false; return; }
            const double l = std::sqrt(dsum);
            A[(size_t)j * nf + j] = l;
            for (int i = j + 1; i < nf; ++i) {
                double v = A[(size_t)i * nf + j];
                for (int k = 0; k < j; ++k) v -= A[(size_t)i * nf + k] * A[(size_t)j * nf + k];
                A[(size_t)i * nf + j] = v / l;
            }
        }
        beta = r;
        for (int i = 0; i < nf; ++i) {
            for (int k = 0; k < i; ++k) beta[i] -= A[(size_t)i * nf + k] * beta[k];
            beta[i] /= A[(size_t)i * nf + i];
        }
        for (int i = nf - 1; i >= 0; --i) {
            for (int k = i + 1; k < nf; ++k) beta[i] -= A[(size_t)k * nf + i] * beta[k];
            beta[i] /= A[(size_t)i * nf + i];
        }
    }
    double predict(const double* x) const {
        std::vector<double> f;
        features(x, f);
        double v = 0;
        for (int a = 0; a < nf; ++a) v += beta[a] * f[a];
        return v;
    }
};

// ranks (0 = smallest) of v
std::vector<double> ranks(const std::vector<double>& v) {
    std::vector<int> idx(v.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return v[a] < v[b]; });
    std::vector<double> r(v.size());
    for (size_t i = 0; i < idx.size(); ++i) r[idx[i]] = (double)i;
    return r;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc != 20) {
        std::printf("Usage: ./run_cmaes <sc_path> <out_prefix|'same'> <fc_tril_path> "
                    "<fcd_tril_path> <maps_path> <n_regions> <G_min[-G_max]> <wee_min[-wee_max]> "
                    "<wei_min[-wei_max]> <wie_min[-wie_max]> <het_params> <time_steps> <BOLD_TR> "
                    "<fcd_step> <fcd_window> <rand_seed> <lambda> <itMax> <SeedMW>\n");
        return 0;
    }
    // ---- arguments (run_CMAES.cpp:192-219) ----
    std::string sc_path = argv[1], out_path = argv[2], fc_tril_path = argv[3], fcd_tril_path = argv[4],
                maps_path = argv[5];
    const int nodes = std::atoi(argv[6]);
    std::map<std::string, std::string> param_strings = {
        {"G", argv[7]}, {"wee", argv[8]}, {"wei", argv[9]}, {"wie", argv[10]}};
    const std::string het_params_str = argv[11];
    const int time_steps = std::atoi(argv[12]), BOLD_TR = std::atoi(argv[13]),
              window_step = std::atoi(argv[14]), window_size = std::atoi(argv[15]),
              rand_seed = std::atoi(argv[16]), lambda = std::atoi(argv[17]), itMax = std::atoi(argv[18]),
              SeedMW = std::atoi(argv[19]);
    if (nodes < 2 || lambda < 2) die("n_regions and lambda must be >= 2");

    // ---- free / fixed parameters (run_CMAES.cpp:226-253): a '-' after the
    // first character marks a range min-max ----
    std::map<std::string, double> params, pmin, pmax, pdiff;
    std::vector<std::string> free_params;
    for (auto& kv : param_strings) {
        const std::string rest = kv.second.substr(1);
        const size_t dash = rest.find('-');
        if (dash != std::string::npos) {
            params[kv.first] = -100;
            pmin[kv.first] = std::atof(kv.second.substr(0, dash + 1).c_str());
            pmax[kv.first] = std::atof(kv.second.substr(dash + 2).c_str());
            pdiff[kv.first] = pmax[kv.first] - pmin[kv.first];
            free_params.push_back(kv.first);
        } else {
            params[kv.first] = std::atof(kv.second.c_str());
        }
    }
    // heterogeneous local parameters (run_CMAES.cpp:255-266)
    const std::vector<std::string> local_params = {"wee", "wie", "wei"};
    std::vector<std::string> het_params, homo_params;
    for (auto& p : local_params)
        (het_params_str.find(p) != std::string::npos ? het_params : homo_params).push_back(p);
    const bool heterogeneous = !het_params.empty();
    // maps and their scale ranges (run_CMAES.cpp:268-310)
    std::vector<std::vector<double>> maps;
    if (heterogeneous) {
        std::ifstream mf(maps_path);
        if (!mf) die("Error opening maps file");
        const double scale_max_minmax = env_d("BNM_SCALE_MAX_MINMAX", 1.0);
        std::string line;
        while (std::getline(mf, line)) {
            std::istringstream iss(line);
            std::vector<double> m;
            double v;
            while (iss >> v) m.push_back(v);
            if (m.empty()) continue;
            if ((int)m.size() != nodes) die("maps file: each map needs n_regions values");
            const double mx = *std::max_element(m.begin(), m.end()), mn = *std::min_element(m.begin(), m.end());
            double smin, smax;
            if (mn == 0 && mx == 1) {
                smin = 0;
                smax = scale_max_minmax;
            } else {
                smin = std::ceil(-1 / mx / 0.01) * 0.01;
                smax = std::floor(-1 / mn / 0.01) * 0.01;
            }
            const int mi = (int)maps.size();
            maps.push_back(m);
            for (auto& hp : het_params) {
                const std::string name = hp + "scale_" + std::to_string(mi);
                params[name] = -100;
                pmin[name] = smin;
                pmax[name] = smax;
                pdiff[name] = smax - smin;
                free_params.push_back(name);
            }
        }
    }
    const int D = (int)free_params.size();
    if (D < 1) die("no free parameter");
    // empirical FC / FCD (run_CMAES.cpp:312-345)
    std::vector<double> emp_fc = read_values(fc_tril_path, "empirical FC");
    const bool no_fcd = fcd_tril_path == "no_fcd";
    std::vector<double> emp_fcd;
    if (!no_fcd) emp_fcd = read_values(fcd_tril_path, "empirical FCD");
    const bool do_fic = params["wie"] == 0;
    if (do_fic && std::find(het_params.begin(), het_params.end(), "wie") != het_params.end())
        die("Error: wIE cannot be a heterogeneous parameter in FIC+ model");
    // SC (run_CMAES.cpp:349-372)
    std::vector<double> SC((size_t)nodes * nodes);
    {
        FILE* fp = std::fopen(sc_path.c_str(), "r");
        if (!fp) die("Error opening SC file");
        for (size_t i = 0; i < SC.size(); ++i)
            if (std::fscanf(fp, "%lf", &SC[i]) != 1) die("SC file: too few values");
        std::fclose(fp);
    }
    std::vector<double> strength(nodes, 0.0);
    for (int j = 0; j < nodes; ++j)
        for (int k = 0; k < nodes; ++k) strength[j] += SC[(size_t)j * nodes + k];

    // ---- output folders and names (run_CMAES.cpp:394-424) ----
    if (out_path == "same") out_path = sc_path.substr(0, sc_path.size() >= 4 ? sc_path.size() - 4 : 0);
    mkdir(out_path.c_str(), 0751);
    std::string sims_dir = out_path + "/sims", cmaes_dir = out_path + "/cmaes";
    if (heterogeneous) { sims_dir += "_multimaps"; cmaes_dir += "_multimaps"; }
    sims_dir += "_gpu";
    cmaes_dir += "_gpu";
    mkdir(sims_dir.c_str(), 0751);
    mkdir(cmaes_dir.c_str(), 0751);
    const size_t pos = fc_tril_path.find_last_of('/');
    const size_t base_start = pos == std::string::npos ? 0 : pos + 1;
    // basename without its "_desc-FCtril.txt" suffix (16 characters); a name
    // shorter than that is kept whole
    std::string out_prefix = fc_tril_path.substr(base_start);
    if (fc_tril_path.size() >= base_start + 16) out_prefix = fc_tril_path.substr(base_start, fc_tril_path.size() - base_start - 16);
    for (auto& kv : param_strings) out_prefix += "_" + kv.first + "_" + kv.second;
    out_prefix += "_het-" + het_params_str + "_SeedMW-" + std::to_string(SeedMW) + "_SeedSim-" +
                  std::to_string(rand_seed) + "_n-" + std::to_string(itMax) + "x" + std::to_string(lambda);
    const std::string log_path = cmaes_dir + "/" + out_prefix + ".txt";
    const std::string sims_prefix = sims_dir + "/" + out_prefix + "_";
    FILE* log = std::fopen(log_path.c_str(), "w");
    if (!log) die("cannot open the log file " + log_path);
    const double sigma0 = env_d("BNM_CMAES_SIGMA", 0.5);
    const int variante = env_i("BNM_CMAES_VARIANTE", 6);
    both(log, "Running %dD CMAES\nSC: %s\nFC: %s\n", D, sc_path.c_str(), fc_tril_path.c_str());
    if (heterogeneous) both(log, "Maps: %s\n", maps_path.c_str());
    both(log, "Free parameters: ");
    for (auto& kv : pmin) both(log, "%s (%f-%f) ", kv.first.c_str(), pmin[kv.first], pmax[kv.first]);
    both(log, "\nFCD step %d, FCD window %d, Simulation random seed %d, CMAES random seed %d, sigma %f, "
         "alphacov %f, gamma_scale %f, bound_soft_edge %f, Variante %d, early stop gens %d, early stop tol %f\n",
         window_step, window_size, rand_seed, SeedMW, sigma0, 2.0, 2.0, 0.0, variante, 30, 1e-3);

    // ---- simulation context ----
    bnm_config cfg;
    check(bnm_default_config(&cfg), "bnm_default_config");
    cfg.nodes = nodes;
    cfg.n_sims = lambda;
    cfg.time_steps = time_steps;
    cfg.bold_tr = BOLD_TR;
    cfg.window_size = window_size;
    cfg.window_step = window_step;
    cfg.do_fic = do_fic;
    cfg.adjust_fic = do_fic && cfg.adjust_fic;
    cfg.calculate_fic_penalty = do_fic;
    cfg.no_fcd = no_fcd;
    cfg.noise_period_ms = (long long)time_steps + 1;
    const std::string prec = env_s("BNM_PRECISION", "f64");
    cfg.precision = prec == "f32" ? BNM_PREC_F32 : prec == "bf16x3" ? BNM_PREC_BF16X3 : BNM_PREC_F64;
    cfg.fc_precision = cfg.precision == BNM_PREC_F64 ? BNM_FC_F64 : BNM_FC_F32;
    check(bnm_finalize_config(&cfg), "bnm_finalize_config");
    const int max_trials = cfg.max_fic_trials;

    // noise: (time_steps + 1) * 10 * N * 2 draws (run_CMAES.cpp:640-657)
    std::printf("Noise precalculation...\n");
    auto t0 = Clock::now();
    const long long noise_len = ((long long)time_steps + 1) * 10 * nodes * 2;
    std::vector<float> noise(noise_len);
    if (env_i("BNM_NOISE_CHUNKED", 0)) check(bnm_make_noise_chunked(rand_seed, noise_len, noise.data(), 0), "noise");
    else check(bnm_make_noise(rand_seed, noise_len, noise.data()), "noise");
    std::printf("took %lf s\n", std::chrono::duration<double>(Clock::now() - t0).count());

    int st = 0;
    bnm_ctx* ctx = bnm_ctx_create(&cfg, &st);
    if (!ctx) die(std::string("bnm_ctx_create: ") + bnm_last_error());
    check(bnm_ctx_set_sc(ctx, SC.data()), "set_sc");
    check(bnm_ctx_set_noise(ctx, noise.data(), noise_len), "set_noise");
    check(bnm_ctx_set_empirical(ctx, emp_fc.data(), (int)emp_fc.size(), no_fcd ? nullptr : emp_fcd.data(),
                                (int)emp_fcd.size()),
          "set_empirical");
    const int n_pairs = bnm_n_pairs(nodes, cfg.exc_interhemispheric);
    if ((int)emp_fc.size() != n_pairs) die("empirical FC size does not match n_regions");
    double emp_fc_mean = 0;
    for (double v : emp_fc) emp_fc_mean += v;
    emp_fc_mean /= emp_fc.size();

    // ---- parameter lists of one candidate (calculate_params, run_CMAES.cpp:61-143) ----
    std::vector<double> G_list(lambda), wEE((size_t)lambda * nodes), wEI((size_t)lambda * nodes),
        wIE((size_t)lambda * nodes);
    auto fill_params = [&](const double* x, int b, double* G_o, double* ee_o, double* ei_o, double* ie_o) {
        std::map<std::string, double> cur = params;
        for (int f = 0; f < D; ++f) cur[free_params[f]] = pdiff[free_params[f]] * x[f] + pmin[free_params[f]];
        G_o[b] = cur["G"];
        std::vector<double> scales(maps.size()), flat;
        for (auto& m : maps) flat.insert(flat.end(), m.begin(), m.end());
        for (auto& hp : het_params) {
            for (size_t m = 0; m < maps.size(); ++m) scales[m] = cur[hp + "scale_" + std::to_string(m)];
            double* o = hp == "wee" ? ee_o : hp == "wei" ? ei_o : ie_o;
            check(bnm_regional_params(nodes, (int)maps.size(), flat.data(), cur[hp], scales.data(),
                                      o + (size_t)b * nodes),
                  "regional params");
        }
        for (auto& hp : homo_params) {
            double* o = hp == "wee" ? ee_o : hp == "wei" ? ei_o : ie_o;
            for (int j = 0; j < nodes; ++j) o[(size_t)b * nodes + j] = cur[hp];
        }
        return cur;
    };

    // ---- surrogate set-up ----
    const bool surrogate = env_i("BNM_SURROGATE", 0) != 0;
    const int pool_factor = std::max(1, env_i("BNM_SURR_POOL", 4));
    const int warmup = std::max(1, env_i("BNM_SURR_WARMUP", 1));
    const double min_rho = env_d("BNM_SURR_MIN_RHO", 0.3);
    const int mlp_hidden = env_i("BNM_SURR_HIDDEN", 256), mlp_depth = env_i("BNM_SURR_DEPTH", 4);
    bnm_mlp* fic_net = nullptr;
    Scaler fic_xs;
    double fic_ys = 1.0;
    double fic_train_sim_s = 0;
    auto an_fic = [&](int B, const double* G_, const double* ee, const double* ei, double* out) {
        std::vector<int> un(B);
        check(bnm_analytical_fic(nodes, B, SC.data(), G_, ee, ei, out, un.data()), "analytical FIC");
    };
    // z_j of the FIC head: the first map (z-scored maps), or 0
    std::vector<double> zmap(nodes, 0.0);
    if (!maps.empty()) zmap = maps[0];
    auto fic_features = [&](int B, const double* G_, const double* ee, const double* ei) {
        std::vector<float> X((size_t)B * nodes * 5);
        for (int b = 0; b < B; ++b)
            for (int j = 0; j < nodes; ++j) {
                float* r = &X[((size_t)b * nodes + j) * 5];
                const size_t q = (size_t)b * nodes + j;
                r[0] = (float)G_[b];
                r[1] = (float)ee[q];
                r[2] = (float)ei[q];
                r[3] = (float)strength[j];
                r[4] = (float)zmap[j];
            }
        return X;
    };
    const std::string fic_model = env_s("BNM_SURR_FIC_MODEL", "");
    const bool guard = env_i("BNM_SURR_GUARD", 1) != 0;
    // online adaptation (BNM_SURR_FIC_ADAPT, on by default): see the loop
    const bool fic_adapt = env_i("BNM_SURR_FIC_ADAPT", 1) != 0;
    const int fic_window = std::max(1, env_i("BNM_SURR_FIC_WINDOW", 3));
    const int fic_steps = env_i("BNM_SURR_FIC_STEPS", 300);
    const double fic_gate = env_d("BNM_SURR_FIC_GATE", 0.9);
    std::vector<float> fa_X, fa_Y;
    std::vector<int> fa_gen;
    if (surrogate && do_fic && cfg.adjust_fic && env_i("BNM_SURR_FIC", 1) && !fic_model.empty()) {
        std::string err;
        if (!load_fic_head(fic_model, &fic_net, fic_xs, &fic_ys, &err)) die("BNM_SURR_FIC_MODEL: " + err);
        both(log, "Surrogate FIC initialiser: head loaded from %s\n", fic_model.c_str());
    } else if (surrogate && do_fic && cfg.adjust_fic && env_i("BNM_SURR_FIC", 1)) {
        // FIC-phase training set: parameter sets drawn uniformly in the search box
        const int n_train = env_i("BNM_SURR_FIC_SIMS", 8 * lambda);
        const int t_max = env_i("BNM_SURR_FIC_TRIALS", 50);
        bnm_config fc = cfg;
        fc.n_sims = n_train;
        fc.time_steps = cfg.i_sampling_end;
        fc.sim_only = 1;
        fc.n_vols_remove = 0;
        fc.noise_period_ms = fc.time_steps + 1;
        fc.max_fic_trials = t_max;
        fc.calculate_fic_penalty = 0;
        bnm_ctx* fctx = bnm_ctx_create(&fc, &st);
        if (!fctx) die(std::string("FIC training context: ") + bnm_last_error());
        check(bnm_ctx_set_sc(fctx, SC.data()), "set_sc");
        check(bnm_ctx_set_noise(fctx, noise.data(), noise_len), "set_noise");
        std::vector<double> g(n_train), ee((size_t)n_train * nodes), ei((size_t)n_train * nodes),
            ie((size_t)n_train * nodes, 0.0), an((size_t)n_train * nodes), wf((size_t)n_train * nodes);
        std::mt19937_64 rng(1000003ull + SeedMW);
        std::uniform_real_distribution<double> u(0, 1);
        std::vector<double> x(D);
        for (int b = 0; b < n_train; ++b) {
            for (int f = 0; f < D; ++f) x[f] = u(rng);
            fill_params(x.data(), b, g.data(), ee.data(), ei.data(), ie.data());
        }
        check(bnm_ctx_set_params(fctx, g.data(), ee.data(), ei.data(), ie.data(), nullptr), "set_params");
        t0 = Clock::now();
        check(bnm_ctx_run(fctx), "FIC training run");
        std::vector<int> failed(n_train), trials(n_train);
        check(bnm_ctx_get_outputs(fctx, nullptr, nullptr, wf.data(), nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, failed.data(), trials.data(), nullptr, nullptr),
              "outputs");
        bnm_ctx_destroy(fctx);
        an_fic(n_train, g.data(), ee.data(), ei.data(), an.data());
        std::vector<float> Xall = fic_features(n_train, g.data(), ee.data(), ei.data()), X, Y;
        int rows = 0;
        double tr_sum = 0;
        for (int b = 0; b < n_train; ++b) {
            tr_sum += trials[b];
            fic_train_sim_s += (fc.time_steps + 1 + (double)trials[b] * (fc.time_steps + 1)) / 1000.0;
            if (failed[b]) continue;
            for (int j = 0; j < nodes; ++j) {
                const size_t q = (size_t)b * nodes + j;
                X.insert(X.end(), &Xall[q * 5], &Xall[q * 5] + 5);
                Y.push_back((float)(wf[q] - an[q]));
                ++rows;
            }
        }
        if (rows < nodes) die("FIC training set: every simulation failed");
        fic_xs.fit(X, rows, 5);
        fic_xs.apply(X, 5);
        double ysd = 0;
        for (float v : Y) ysd += (double)v * v;
        fic_ys = std::sqrt(ysd / rows) + 1e-6;
        for (float& v : Y) v = (float)(v / fic_ys);
        fic_net = bnm_mlp_create(5, mlp_hidden, mlp_depth, 1, 1, 7, &st);
        if (!fic_net) die(std::string("MLP: ") + bnm_mlp_last_error());
        check(bnm_mlp_set_data(fic_net, X.data(), Y.data(), rows), "mlp data");
        const int epochs = env_i("BNM_SURR_FIC_EPOCHS", 60);
        std::vector<float> losses(epochs);
        check(bnm_mlp_train(fic_net, epochs, 1024, 1e-3f, 11, losses.data()), "mlp train");
        both(log, "Surrogate FIC initialiser: %d sets (%d converged), mean trials %.3f, %.1f simulated s; "
             "%d rows, loss %.4g -> %.4g, took %lf s\n",
             n_train, rows / nodes, tr_sum / n_train, fic_train_sim_s, rows, losses.front(), losses.back(),
             std::chrono::duration<double>(Clock::now() - t0).count());
    }
    // fit surrogate: unit-box coordinates -> (FC r, FCD KS, |mean FC diff|,
    // FIC penalty, cost); trained on the last `window` generations, inputs
    // and outputs standardised over that window (so the network resolves
    // the region the search has narrowed to), ranked by the cost output
    constexpr int KF = 5;
    bnm_mlp* fit_net = nullptr;          // first head of the ensemble
    std::vector<bnm_mlp*> fit_nets;
    const int n_heads = std::max(1, env_i("BNM_SURR_ENSEMBLE", 1));
    const int n_unscreened = std::min(lambda, std::max(0, (int)std::lround(env_d("BNM_SURR_UNSCREENED", 0.25) * lambda)));
    std::vector<float> fit_X, fit_Y;
    std::vector<int> fit_gen;            // generation of each row
    Scaler fit_xs, fit_ys;
    const int window = std::max(1, env_i("BNM_SURR_WINDOW", 6));
    // pre-screening model: 0 the MLP, 1 the local linear-quadratic model,
    // 2 both (candidates ranked by the sum of their two ranks)
    const int surr_model = env_i("BNM_SURR_MODEL", 0);
    const int lq_points = env_i("BNM_SURR_LQ_POINTS", 0);   // 0: max(2 * terms, 2 * lambda)
    LqModel lq;
    if (surrogate) {
        for (int h = 0; h < n_heads; ++h) {
            bnm_mlp* m = bnm_mlp_create(D, mlp_hidden, mlp_depth, KF, 1, 13u + 1000u * h, &st);
            if (!m) die(std::string("MLP: ") + bnm_mlp_last_error());
            fit_nets.push_back(m);
        }
        fit_net = fit_nets[0];
    }

    bnm_ctx* sctx = nullptr;
    if (fic_net) {
        bnm_config sc = cfg;
        sc.analytical_fic = 0;
        sctx = bnm_ctx_create(&sc, &st);
        if (!sctx) die(std::string("context: ") + bnm_last_error());
        check(bnm_ctx_set_sc(sctx, SC.data()), "set_sc");
        check(bnm_ctx_set_noise(sctx, noise.data(), noise_len), "set_noise");
        check(bnm_ctx_set_empirical(sctx, emp_fc.data(), (int)emp_fc.size(), no_fcd ? nullptr : emp_fcd.data(),
                                    (int)emp_fcd.size()),
              "set_empirical");
    }

    // ---- the search (run_CMAES.cpp:714-832) ----
    bnm_cmaes* es = bnm_cmaes_create(D, lambda, (unsigned)SeedMW, sigma0, variante);
    if (!es) die("CMA-ES set-up failed (variant must be 6 or 16)");
    std::vector<double> x((size_t)lambda * D), corr(lambda), pen(lambda), raw(lambda), penalized(lambda),
        fc((size_t)lambda * n_pairs), wfic((size_t)lambda * nodes);
    std::vector<int> failed(lambda), trials(lambda);
    FILE* prog = nullptr;
    if (env_i("BNM_PROGRESS_CSV", 1)) {
        prog = std::fopen((cmaes_dir + "/" + out_prefix + "_progress.csv").c_str(), "w");
        if (prog)
            std::fprintf(prog, "iteration,full_sims,simulated_s,gen_best_cost,best_cost,best_plain_cost,"
                               "mean_fic_trials,max_fic_trials,surrogate_used,surrogate_rho,gpu_ms,guard_sims\n");
    }
    // With the learned FIC start every cost the search sees comes from a
    // simulation whose FIC began elsewhere than the reference's analytical
    // start, so the generation's best is re-run with the plain method (one
    // more full simulation per generation, counted) and that cost is the
    // one reported.
    bnm_ctx* jctx = nullptr;
    if (fic_net) {
        bnm_config one = cfg;
        one.n_sims = 1;
        jctx = bnm_ctx_create(&one, &st);
        if (!jctx) die(std::string("context: ") + bnm_last_error());
        check(bnm_ctx_set_sc(jctx, SC.data()), "set_sc");
        check(bnm_ctx_set_noise(jctx, noise.data(), noise_len), "set_noise");
        check(bnm_ctx_set_empirical(jctx, emp_fc.data(), (int)emp_fc.size(), no_fcd ? nullptr : emp_fcd.data(),
                                    (int)emp_fcd.size()),
              "set_empirical");
    }
    long long full_sims = 0;
    double sim_seconds = 0, last_rho = -1, best_so_far = INFINITY, best_plain = INFINITY;
    bool used_surr = false;
    // learned FIC start in the current generation (with adaptation the first
    // generation starts from the analytical wIE and trains the head first)
    bool fic_on = fic_net && !fic_adapt;
    while (bnm_cmaes_iteration(es) < itMax) {
        const int it = bnm_cmaes_iteration(es);
        both(log, "Iteration %d\n", it);
        std::vector<double> pred_cost;
        used_surr = false;
        if (surrogate && it > warmup && last_rho >= min_rho && pool_factor > 1 && (surr_model == 0 || lq.ok)) {
            // pre-screen a pool of pool_factor x lambda candidates
            const int n_pool = pool_factor * lambda;
            std::vector<double> pool((size_t)n_pool * D);
            check(bnm_cmaes_ask_pool(es, n_pool, 7919u * it + (unsigned)SeedMW, pool.data()), "ask_pool");
            std::vector<double> c(n_pool), cl(n_pool);
            if (surr_model != 1) {
                std::vector<float> px(pool.begin(), pool.end()), py((size_t)n_pool * KF);
                fit_xs.apply(px, D);
                std::fill(c.begin(), c.end(), 0.0);
                for (bnm_mlp* m : fit_nets) {
                    check(bnm_mlp_predict(m, px.data(), n_pool, py.data()), "surrogate predict");
                    for (int i = 0; i < n_pool; ++i) c[i] += py[(size_t)i * KF + KF - 1];
                }
                for (int i = 0; i < n_pool; ++i)
                    c[i] = fit_ys.mu[KF - 1] + fit_ys.sd[KF - 1] * c[i] / n_heads;
            }
            if (surr_model != 0) {
                for (int i = 0; i < n_pool; ++i) cl[i] = lq.predict(&pool[(size_t)i * D]);
                if (surr_model == 1) c = cl;
                else {
                    const std::vector<double> ra = ranks(c), rb = ranks(cl);
                    for (int i = 0; i < n_pool; ++i) c[i] = ra[i] + rb[i];
                }
            }
            std::vector<int> idx(n_pool);
            std::iota(idx.begin(), idx.end(), 0);
            // the unscreened part: a seeded random subset of the pool (the
            // pool is an i.i.d. sample of the distribution), moved to the front
            std::mt19937 pick(104729u * it + (unsigned)SeedMW);
            for (int i = 0; i < n_unscreened; ++i)
                std::swap(idx[i], idx[i + (int)(pick() % (unsigned)(n_pool - i))]);
            std::stable_sort(idx.begin() + n_unscreened, idx.end(), [&](int a, int b) { return c[a] < c[b]; });
            idx.resize(lambda);
            check(bnm_cmaes_select(es, idx.data()), "select");
            for (int i = 0; i < lambda; ++i) pred_cost.push_back(c[idx[i]]);
            used_surr = true;
        }
        check(bnm_cmaes_ask(es, x.data()), "ask");
        for (int b = 0; b < lambda; ++b)
            fill_params(&x[(size_t)b * D], b, G_list.data(), wEE.data(), wEI.data(), wIE.data());
        const double* wfic_in = nullptr;
        std::vector<double> an, init;
        std::vector<float> fx_raw;
        if (fic_net) {
            an.resize((size_t)lambda * nodes);
            an_fic(lambda, G_list.data(), wEE.data(), wEI.data(), an.data());
            fx_raw = fic_features(lambda, G_list.data(), wEE.data(), wEI.data());
            std::vector<float> X = fx_raw;
            fic_xs.apply(X, 5);
            std::vector<float> y((size_t)lambda * nodes);
            check(bnm_mlp_predict(fic_net, X.data(), lambda * nodes, y.data()), "FIC predict");
            init.resize(an.size());
            for (size_t q = 0; q < an.size(); ++q) init[q] = an[q] + fic_ys * y[q];
            if (fic_on) wfic_in = init.data();
        }
        // with the learned initialiser the context takes w_IE_fic as given
        const bool learned = fic_net && fic_on;
        bnm_ctx* rctx = learned ? sctx : ctx;
        check(bnm_ctx_set_params(rctx, G_list.data(), wEE.data(), wEI.data(), wIE.data(), wfic_in), "params");
        t0 = Clock::now();
        check(bnm_ctx_run(rctx), "run");
        check(bnm_ctx_get_outputs(rctx, corr.data(), pen.data(), wfic.data(), nullptr, nullptr, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, failed.data(), trials.data(), fc.data(), nullptr),
              "outputs");
        full_sims += lambda;
        double tr_sum = 0;
        int tr_max = 0;
        for (int b = 0; b < lambda; ++b) {
            tr_sum += trials[b];
            tr_max = std::max(tr_max, trials[b]);
            sim_seconds += (time_steps + 1 + (do_fic ? (double)trials[b] * (cfg.i_sampling_end + 1) : 0.0)) / 1000.0;
        }
        // plain-consistency guard: candidates whose plain FIC might have
        // failed are simulated again with the plain method (analytical
        // start) in one batch (the batch's remaining slots repeat them)
        int n_guard = 0;
        std::vector<char> guarded(lambda, 0);
        if (learned && guard) {
            std::vector<int> flag;
            for (int b = 0; b < lambda; ++b) {
                int k = 0;
                for (int j = 0; j < nodes; ++j) {
                    const size_t q = (size_t)b * nodes + j;
                    k = std::max(k, plain_fic_steps(init[q] - an[q], cfg.init_delta));
                }
                if (failed[b] || k + trials[b] >= max_trials) flag.push_back(b);
            }
            n_guard = (int)flag.size();
            if (n_guard > 0) {
                std::vector<double> gG(lambda), gee((size_t)lambda * nodes), gei((size_t)lambda * nodes),
                    gie((size_t)lambda * nodes), gc(lambda), gp(lambda), gfc((size_t)lambda * n_pairs);
                std::vector<int> gf(lambda), gt(lambda);
                std::vector<double> gw((size_t)lambda * nodes);
                for (int s2 = 0; s2 < lambda; ++s2)
                    fill_params(&x[(size_t)flag[s2 % n_guard] * D], s2, gG.data(), gee.data(), gei.data(), gie.data());
                check(bnm_ctx_set_params(ctx, gG.data(), gee.data(), gei.data(), gie.data(), nullptr), "params");
                check(bnm_ctx_run(ctx), "plain-consistency run");
                check(bnm_ctx_get_outputs(ctx, gc.data(), gp.data(), gw.data(), nullptr, nullptr, nullptr, nullptr,
                                          nullptr, nullptr, nullptr, gf.data(), gt.data(), gfc.data(), nullptr),
                      "outputs");
                for (int i = 0; i < n_guard; ++i) {
                    const int b = flag[i];
                    guarded[b] = 1;
                    std::copy(&gw[(size_t)i * nodes], &gw[(size_t)(i + 1) * nodes], &wfic[(size_t)b * nodes]);
                    corr[b] = gc[i];
                    pen[b] = gp[i];
                    failed[b] = gf[i];
                    std::copy(&gfc[(size_t)i * n_pairs], &gfc[(size_t)(i + 1) * n_pairs], &fc[(size_t)b * n_pairs]);
                    tr_sum += gt[i];
                    tr_max = std::max(tr_max, gt[i]);
                    sim_seconds += (time_steps + 1 + (double)gt[i] * (cfg.i_sampling_end + 1)) / 1000.0;
                }
                full_sims += n_guard;
                both(log, "surrogate: plain-consistency re-runs: %d\n", n_guard);
            }
        }
        // online adaptation of the FIC head to this search (SC, regional form):
        // the converged wIE of this generation's converged simulations join a
        // window of the last fic_window generations, the head is trained on
        // it, and the learned start is used in the next generation only if
        // the head is closer to the converged wIE than the analytical start
        if (fic_net && fic_adapt) {
            double e_an = 0, e_lr = 0;
            int n_ok = 0;
            for (int b = 0; b < lambda; ++b) {
                if (failed[b]) continue;
                for (int j = 0; j < nodes; ++j) {
                    const size_t q = (size_t)b * nodes + j;
                    fa_X.insert(fa_X.end(), &fx_raw[q * 5], &fx_raw[q * 5] + 5);
                    fa_Y.push_back((float)((wfic[q] - an[q]) / fic_ys));
                    fa_gen.push_back(it);
                    if (!guarded[b]) {
                        e_an += std::fabs(an[q] - wfic[q]);
                        e_lr += std::fabs(init[q] - wfic[q]);
                    }
                }
                if (!guarded[b]) ++n_ok;
            }
            // the head's error on this generation, before training on it
            const bool better = n_ok > 0 && e_lr < fic_gate * e_an;
            std::vector<float> wx, wy;
            for (size_t i = 0; i < fa_gen.size(); ++i)
                if (fa_gen[i] > it - fic_window) {
                    wx.insert(wx.end(), &fa_X[i * 5], &fa_X[i * 5] + 5);
                    wy.push_back(fa_Y[i]);
                }
            const int nrow = (int)wy.size();
            if (nrow >= 1024) {
                fic_xs.apply(wx, 5);
                check(bnm_mlp_set_data(fic_net, wx.data(), wy.data(), nrow), "FIC data");
                const int epochs = std::max(1, fic_steps / std::max(1, nrow / 1024));
                check(bnm_mlp_train(fic_net, epochs, 1024, 3e-4f, 31u + it, nullptr), "FIC adapt");
            }
            both(log, "surrogate: FIC start %s this generation; mean |start - converged| analytical %.5f, "
                      "head %.5f over %d simulations; next generation %s\n",
                 learned ? "learned" : "analytical", n_ok ? e_an / (n_ok * nodes) : NAN,
                 n_ok ? e_lr / (n_ok * nodes) : NAN, n_ok, better ? "learned" : "analytical");
            fic_on = better;
        }
        for (int b = 0; b < lambda; ++b) raw[b] = corr[b] + (do_fic ? pen[b] : 0.0);   // run_CMAES.cpp:787-793
        const double took = std::chrono::duration<double>(Clock::now() - t0).count();
        std::printf("took %lf s\n", took);
        // surrogate bookkeeping: targets of every simulated candidate
        if (fit_net) {
            std::vector<double> actual(lambda);
            for (int b = 0; b < lambda; ++b) {
                const double* f = &fc[(size_t)b * n_pairs];
                double m = 0;
                for (int i = 0; i < n_pairs; ++i) m += f[i];
                m /= n_pairs;
                const double r = pearson(emp_fc.data(), f, n_pairs);
                const double dm = cfg.use_fc_diff ? std::fabs(emp_fc_mean - m) : 0.0;
                // a simulation whose FC/FCD had a NaN costs max_cost: targets
                // r = 0 and KS = |dmean| = (max_cost - 1) / 2 add up to it
                const bool ok = std::isfinite(r) && corr[b] < cfg.max_cost;
                const double half = 0.5 * (cfg.max_cost - 1);
                const double ks = ok ? corr[b] - (1 - r) - dm : half;
                for (int f2 = 0; f2 < D; ++f2) fit_X.push_back((float)x[(size_t)b * D + f2]);
                fit_Y.push_back((float)(ok ? r : 0.0));
                fit_Y.push_back((float)ks);
                fit_Y.push_back((float)(ok ? dm : half));
                fit_Y.push_back((float)(do_fic ? pen[b] : 0.0));
                fit_Y.push_back((float)raw[b]);
                fit_gen.push_back(it);
                actual[b] = raw[b];
            }
            if (used_surr) last_rho = spearman(pred_cost, actual);
            // training window: the last `window` generations
            std::vector<float> wx, wy;
            for (size_t i = 0; i < fit_gen.size(); ++i)
                if (fit_gen[i] > it - window) {
                    wx.insert(wx.end(), &fit_X[i * D], &fit_X[i * D] + D);
                    wy.insert(wy.end(), &fit_Y[i * KF], &fit_Y[i * KF] + KF);
                }
            const int n = (int)(wy.size() / KF);
            fit_xs.fit(wx, n, D);
            fit_ys.fit(wy, n, KF);
            fit_xs.apply(wx, D);
            fit_ys.apply(wy, KF);
            if (surr_model != 1) {
                const int steps = env_i("BNM_SURR_FIT_STEPS", 1500);
                const int batch = std::min(n, 128);
                const int epochs = std::max(1, steps / std::max(1, n / batch));
                for (int h = 0; h < n_heads; ++h) {
                    check(bnm_mlp_set_data(fit_nets[h], wx.data(), wy.data(), n), "surrogate data");
                    check(bnm_mlp_train(fit_nets[h], epochs, batch, 1e-3f, 17u + it + 7919u * h, nullptr),
                          "surrogate train");
                }
            }
            if (surr_model != 0) {
                // the most recent points whose simulations produced a cost
                const int nq = 1 + D + D * (D + 1) / 2;
                const int want = lq_points > 0 ? lq_points : std::max(2 * nq, 2 * lambda);
                std::vector<double> lx, ly;
                for (size_t i = fit_gen.size(); i-- > 0 && (int)ly.size() < want;) {
                    const double yv = fit_Y[i * KF + KF - 1];
                    if (!(yv < cfg.max_cost)) continue;
                    ly.push_back(yv);
                    lx.insert(lx.end(), &fit_X[i * D], &fit_X[i * D] + D);
                }
                lq.fit(lx, ly, D);
            }
            if (!used_surr) last_rho = 1.0;   // the first fit enables pre-screening
        }
        const int stopped = bnm_cmaes_tell(es, raw.data(), penalized.data());
        if (stopped < 0) die("CMA-ES tell failed");
        double gen_best = INFINITY;
        for (int i = 0; i < lambda; ++i) {
            gen_best = std::min(gen_best, raw[i]);
            both(log, "\t IndPar %d: ", i);
            std::map<std::string, double> cur = params;
            for (int f = 0; f < D; ++f) {
                cur[free_params[f]] = pdiff[free_params[f]] * x[(size_t)i * D + f] + pmin[free_params[f]];
                both(log, "%s=%f,", free_params[f].c_str(), cur[free_params[f]]);
            }
            if (do_fic) both(log, "gof=%f,fic_penalty=%f,penalized_cost=%f\n", -corr[i], pen[i], penalized[i]);
            else both(log, "gof=%f,penalized_cost=%f\n", -corr[i], penalized[i]);
        }
        best_so_far = std::min(best_so_far, gen_best);
        if (jctx) {
            int ib = 0;
            for (int i = 1; i < lambda; ++i)
                if (raw[i] < raw[ib]) ib = i;
            std::vector<double> g1(1), ee1(nodes), ei1(nodes), ie1(nodes), c1(1), p1(1);
            int t1 = 0;
            fill_params(&x[(size_t)ib * D], 0, g1.data(), ee1.data(), ei1.data(), ie1.data());
            check(bnm_ctx_set_params(jctx, g1.data(), ee1.data(), ei1.data(), ie1.data(), nullptr), "params");
            check(bnm_ctx_run(jctx), "plain re-run");
            check(bnm_ctx_get_outputs(jctx, c1.data(), p1.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, nullptr, nullptr, &t1, nullptr, nullptr),
                  "outputs");
            const double plain_cost = c1[0] + (do_fic ? p1[0] : 0.0);
            ++full_sims;
            sim_seconds += (time_steps + 1 + (do_fic ? (double)t1 * (cfg.i_sampling_end + 1) : 0.0)) / 1000.0;
            best_plain = std::min(best_plain, plain_cost);
            both(log, "surrogate: plain re-run of the generation best (IndPar %d): cost %f (search saw %f)\n", ib,
                 plain_cost, raw[ib]);
        } else {
            best_plain = best_so_far;
        }
        if (prog)
            std::fprintf(prog, "%d,%lld,%.3f,%.9g,%.9g,%.9g,%.4f,%d,%d,%.4f,%.3f,%d\n", it, full_sims, sim_seconds,
                         gen_best, best_so_far, best_plain, tr_sum / (lambda + n_guard), tr_max, used_surr ? 1 : 0,
                         used_surr ? last_rho : NAN, took * 1e3, n_guard);
        if (surrogate && used_surr) both(log, "surrogate: rank correlation of predicted and simulated cost %.3f\n", last_rho);
        if (stopped) break;   // early stop, or the covariance lost definiteness
    }
    if (prog) std::fclose(prog);

    // ---- best feasible point, re-run with the plain method (run_CMAES.cpp:835-906) ----
    std::vector<double> bx(D);
    double best_raw = 0;
    check(bnm_cmaes_best(es, bx.data(), &best_raw), "best");
    both(log, "Best feasible parameters: ");
    {
        std::map<std::string, double> cur = params;
        for (int f = 0; f < D; ++f) {
            cur[free_params[f]] = pdiff[free_params[f]] * bx[f] + pmin[free_params[f]];
            both(log, "%s=%f,", free_params[f].c_str(), cur[free_params[f]]);
        }
    }
    bnm_config one = cfg;
    one.n_sims = 1;
    bnm_ctx* bctx = bnm_ctx_create(&one, &st);
    if (!bctx) die(std::string("context: ") + bnm_last_error());
    check(bnm_ctx_set_sc(bctx, SC.data()), "set_sc");
    check(bnm_ctx_set_noise(bctx, noise.data(), noise_len), "set_noise");
    check(bnm_ctx_set_empirical(bctx, emp_fc.data(), (int)emp_fc.size(), no_fcd ? nullptr : emp_fcd.data(),
                                (int)emp_fcd.size()),
          "set_empirical");
    std::vector<double> g1(1), ee1(nodes), ei1(nodes), ie1(nodes);
    fill_params(bx.data(), 0, g1.data(), ee1.data(), ei1.data(), ie1.data());
    check(bnm_ctx_set_params(bctx, g1.data(), ee1.data(), ei1.data(), ie1.data(), nullptr), "params");
    check(bnm_ctx_run(bctx), "best run");
    const int ots = bnm_output_ts(time_steps, BOLD_TR);
    std::vector<double> c1(1), p1(1), w1(nodes), bold((size_t)ots * nodes), ext(6 * (size_t)nodes);
    int f1 = 0, t1 = 0;
    check(bnm_ctx_get_outputs(bctx, c1.data(), p1.data(), w1.data(), bold.data(), &ext[0], &ext[nodes],
                              &ext[2 * nodes], &ext[3 * nodes], &ext[4 * nodes], &ext[5 * nodes], &f1, &t1, nullptr,
                              nullptr),
          "best outputs");
    bnm_ctx_destroy(bctx);
    if (do_fic) both(log, "gof=%f,fic_penalty=%f,gof_fic_penalty=%f\n", -c1[0], p1[0], -best_raw);
    else both(log, "gof=%f\n", -best_raw);
    // best-simulation files (bnm.cu:771-833)
    auto wfile = [&](const std::string& name, const double* v, int n) {
        FILE* f = std::fopen((sims_prefix + name).c_str(), "w");
        if (!f) die("cannot write " + sims_prefix + name);
        for (int j = 0; j < n; ++j) std::fprintf(f, "%f ", v[j]);
        std::fclose(f);
    };
    {
        FILE* f = std::fopen((sims_prefix + "bold.txt").c_str(), "w");
        if (!f) die("cannot write " + sims_prefix + "bold.txt");
        for (int i = 0; i < ots; ++i) {
            for (int j = 0; j < nodes; ++j) std::fprintf(f, "%f ", bold[(size_t)i * nodes + j]);
            std::fprintf(f, "\n");
        }
        std::fclose(f);
    }
    const char* ext_names[6] = {"S_E.txt", "I_E.txt", "r_E.txt", "S_I.txt", "I_I.txt", "r_I.txt"};
    for (int k = 0; k < 6; ++k) wfile(ext_names[k], &ext[(size_t)k * nodes], nodes);
    wfile("w_EE.txt", ee1.data(), nodes);
    wfile("w_EI.txt", ei1.data(), nodes);
    if (do_fic) {
        wfile("w_IE.txt", w1.data(), nodes);
        if (cfg.adjust_fic) {
            FILE* f = std::fopen((sims_prefix + "fic_failed.txt").c_str(), "w");
            if (f) { std::fprintf(f, "%d", f1); std::fclose(f); }
            f = std::fopen((sims_prefix + "fic_ntrials.txt").c_str(), "w");
            if (f) { std::fprintf(f, "%d of max %d", t1, max_trials); std::fclose(f); }
        }
    }
    if (surrogate)
        both(log, "Surrogate summary: full simulations %lld, simulated s %.1f (+%.1f for the FIC training set)\n",
             full_sims, sim_seconds, fic_train_sim_s);
    std::fclose(log);
    bnm_cmaes_destroy(es);
    if (fic_net) bnm_mlp_destroy(fic_net);
    for (bnm_mlp* m : fit_nets) bnm_mlp_destroy(m);
    if (sctx) bnm_ctx_destroy(sctx);
    if (jctx) bnm_ctx_destroy(jctx);
    bnm_ctx_destroy(ctx);
    return 0;
}
