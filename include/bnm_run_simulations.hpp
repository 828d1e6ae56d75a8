// bnm_run_simulations.hpp -- C++ drop-in for the reference's GPU
// run_simulations (bnm.cu:594-617) on top of the C ABI in bnm_hip.h.
//
// The parameter list keeps the reference's names, order and meaning; the types
// change where the reference used GSL, CUDA or array-of-pointer buffers:
//   gsl_vector*      -> std::vector<double>
//   u_real**         -> flat sim-major double arrays ([N_SIMS*nodes] etc.)
//   output_ts        -> noise_len (output_ts follows from time_steps / BOLD_TR)
//   cudaDeviceProp, window/pair tables, scratch buffers -> derived inside
// The reference's callee-visible environment constants (BNM_*) are read by
// bnm_default_config.  Returns a bnm status code instead of calling exit().
//
// Defaults follow the reference's default build (u_real = double unless
// USE_FLOATS, constants.h:8-22): FP64 integration and FP64 FC/FCD unless
// RunOptions asks for another precision.
//
// extended_output: the per-node means S_E .. r_I are copied to the caller only
// when it is true (bnm.cu:145, 305-314); the FIC penalty uses r_E either way.
// save_output: with N_SIMS == 1 writes the reference's text files
// (bnm.cu:771-833): <prefix>bold.txt (one "%f " per node, one line per
// volume), S_E / I_E / r_E / S_I / I_I / r_I / w_EE / w_EI .txt (one line of
// "%f " per node), and with do_fic w_IE.txt (adjusted wIE) plus, with
// adjust_fic, fic_failed.txt ("%d") and fic_ntrials.txt ("%d of max %d").
// With N_SIMS > 1 it returns BNM_E_UNSUP before running (the reference prints
// "Saving multiple simulations output not implemented" and exits).
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "bnm_hip.h"

namespace bnm {

struct RunOptions {
    int window_step = 5;          // FCD window step (the reference caller builds
                                  // the window table with it, run_CMAES.cpp:529-594)
    long long noise_period_ms = 0;  // 0: noise covers the whole run (T+1 ms)
    int precision = BNM_PREC_F64;
    int fc_precision = BNM_FC_F64;
    int kernel = BNM_KERNEL_AUTO;
    int analytical_fic = 1;       // compute the analytical wIE on the device
};

namespace detail {

// one "%f " per value, then an optional newline (bnm.cu:797-817)
inline bool write_row(FILE* f, const double* v, int n) {
    for (int j = 0; j < n; ++j)
        if (std::fprintf(f, "%f ", v[j]) < 0) return false;
    return true;
}

inline int save_outputs(const std::string& prefix, int nodes, int output_ts, int max_fic_trials,
                        bool do_fic, bool adjust_fic, const double* bold, const double* const ext[6],
                        const double* w_EE, const double* w_EI, const double* w_IE, int failed,
                        int n_trials) {
    static const char* names[6] = {"S_E", "I_E", "r_E", "S_I", "I_I", "r_I"};
    auto open = [&](const char* name) { return std::fopen((prefix + name).c_str(), "w"); };
    FILE* f = open("bold.txt");
    if (!f) return BNM_E_ARG;
    bool ok = true;
    for (int i = 0; i < output_ts && ok; ++i)
        ok = write_row(f, bold + (size_t)i * nodes, nodes) && std::fprintf(f, "\n") >= 0;
    std::fclose(f);
    for (int k = 0; k < 6 && ok; ++k) {
        f = open((std::string(names[k]) + ".txt").c_str());
        if (!f) return BNM_E_ARG;
        ok = write_row(f, ext[k], nodes);
        std::fclose(f);
    }
    const double* w[3] = {w_EE, w_EI, w_IE};
    const char* wn[3] = {"w_EE.txt", "w_EI.txt", "w_IE.txt"};
    for (int k = 0; k < (do_fic ? 3 : 2) && ok; ++k) {
        f = open(wn[k]);
        if (!f) return BNM_E_ARG;
        ok = write_row(f, w[k], nodes);
        std::fclose(f);
    }
    if (do_fic && adjust_fic && ok) {
        f = open("fic_failed.txt");
        if (!f) return BNM_E_ARG;
        ok = std::fprintf(f, "%d", failed) >= 0;
        std::fclose(f);
        f = open("fic_ntrials.txt");
        if (!f) return BNM_E_ARG;
        ok = ok && std::fprintf(f, "%d of max %d", n_trials, max_fic_trials) >= 0;
        std::fclose(f);
    }
    return ok ? BNM_OK : BNM_E_ARG;
}

}  // namespace detail

inline int run_simulations(
    double* corr, double* fic_penalties,
    const double* G_list, const double* w_EE_list, const double* w_EI_list,
    const double* w_IE_list, double* w_IE_fic,
    const double* SC, int nodes, bool calculate_fic_penalty,
    int time_steps, int BOLD_TR, int _max_fic_trials, bool sim_only,
    const std::vector<double>& emp_FC_tril, const std::vector<double>& emp_FCD_tril,
    int window_size,
    bool save_output, bool extended_output, const std::string& sims_out_prefix,
    int N_SIMS, const float* noise, long long noise_len,
    double* BOLD_ex,
    double* S_E, double* I_E, double* r_E, double* S_I, double* I_I, double* r_I,
    bool do_fic, bool adjust_fic, int* FIC_failed, int* fic_n_trials,
    int n_vols_remove,
    double* fc_trils, double* fcd_trils,
    const RunOptions& opt = RunOptions()) {
    if (save_output && N_SIMS != 1) return BNM_E_UNSUP;
    bnm_config c;
    int st = bnm_default_config(&c);
    if (st) return st;
    c.nodes = nodes;
    c.n_sims = N_SIMS;
    c.time_steps = time_steps;
    c.bold_tr = BOLD_TR;
    c.n_vols_remove = n_vols_remove;
    c.window_size = window_size;
    c.window_step = opt.window_step;
    c.max_fic_trials = _max_fic_trials;
    c.sim_only = sim_only;
    c.calculate_fic_penalty = calculate_fic_penalty;
    c.do_fic = do_fic;
    c.adjust_fic = adjust_fic;
    c.noise_period_ms = opt.noise_period_ms > 0 ? opt.noise_period_ms : (long long)time_steps + 1;
    c.precision = opt.precision;
    c.fc_precision = opt.fc_precision;
    c.kernel = opt.kernel;
    c.analytical_fic = opt.analytical_fic;
    // what the output files need when the caller did not ask for it
    const int ots = bnm_output_ts(time_steps, BOLD_TR);
    if (ots < 0) return ots;
    std::vector<double> bold_tmp, ext_tmp, wfic_tmp;
    std::vector<int> failed_tmp, trials_tmp;
    double* ext[6] = {S_E, I_E, r_E, S_I, I_I, r_I};
    if (!extended_output)
        for (double*& p : ext) p = nullptr;
    if (save_output) {
        if (!BOLD_ex) {
            bold_tmp.resize((size_t)ots * nodes);
            BOLD_ex = bold_tmp.data();
        }
        ext_tmp.resize((size_t)6 * nodes);
        for (int k = 0; k < 6; ++k)
            if (!ext[k]) ext[k] = ext_tmp.data() + (size_t)k * nodes;
        if (do_fic && !w_IE_fic && !opt.analytical_fic) return BNM_E_ARG;
        if (do_fic && !w_IE_fic) {
            wfic_tmp.resize(nodes);
            w_IE_fic = wfic_tmp.data();
        }
        if (!FIC_failed) {
            failed_tmp.resize(1);
            FIC_failed = failed_tmp.data();
        }
        if (!fic_n_trials) {
            trials_tmp.resize(1);
            fic_n_trials = trials_tmp.data();
        }
    }
    st = bnm_run_simulations(&c, corr, fic_penalties, G_list, w_EE_list, w_EI_list, w_IE_list,
                             w_IE_fic, SC, noise, noise_len, emp_FC_tril.data(),
                             (int)emp_FC_tril.size(), emp_FCD_tril.data(),
                             (int)emp_FCD_tril.size(), BOLD_ex, ext[0], ext[1], ext[2], ext[3],
                             ext[4], ext[5], FIC_failed, fic_n_trials, fc_trils, fcd_trils);
    if (st || !save_output) return st;
    const double* cext[6] = {ext[0], ext[1], ext[2], ext[3], ext[4], ext[5]};
    return detail::save_outputs(sims_out_prefix, nodes, ots, _max_fic_trials, do_fic, adjust_fic,
                                BOLD_ex, cext, w_EE_list, w_EI_list, w_IE_fic, FIC_failed[0],
                                fic_n_trials[0]);
}

}  // namespace bnm
