// bnm_run_simulations.hpp -- C++ drop-in for the reference's GPU
// run_simulations (bnm.cu:594-617) on top of the C ABI in bnm_hip.h.
//
// The parameter list keeps the reference's names, order and meaning; the types
// change where the reference used GSL, CUDA or array-of-pointer buffers:
//   gsl_vector*      -> std::vector<double>
//   u_real**         -> flat sim-major double arrays ([N_SIMS*nodes] etc.)
//   cudaDeviceProp, window/pair tables, scratch buffers -> derived inside
// The reference's callee-visible environment constants (BNM_*) are read by
// bnm_default_config.  Returns a bnm status code instead of calling exit().
#pragma once
#include <stdexcept>
#include <string>
#include <vector>

#include "bnm_hip.h"

namespace bnm {

struct RunOptions {
    int window_step = 5;          // FCD window step (the reference caller builds
                                  // the window table with it, run_CMAES.cpp:529-594)
    long long noise_period_ms = 0;  // 0: noise covers the whole run (T+1 ms)
    int precision = BNM_PREC_F32;
    int fc_precision = BNM_FC_F32;
    int kernel = BNM_KERNEL_AUTO;
    int analytical_fic = 1;       // compute the analytical wIE on the device
};

inline int run_simulations(
    double* corr, double* fic_penalties,
    const double* G_list, const double* w_EE_list, const double* w_EI_list,
    const double* w_IE_list, double* w_IE_fic,
    const double* SC, int nodes, bool calculate_fic_penalty,
    int time_steps, int BOLD_TR, int _max_fic_trials, bool sim_only,
    const std::vector<double>& emp_FC_tril, const std::vector<double>& emp_FCD_tril,
    int window_size, bool extended_output,
    int N_SIMS, const float* noise, long long noise_len,
    double* BOLD_ex,
    double* S_E, double* I_E, double* r_E, double* S_I, double* I_I, double* r_I,
    bool do_fic, bool adjust_fic, int* FIC_failed, int* fic_n_trials,
    int n_vols_remove,
    double* fc_trils, double* fcd_trils,
    const RunOptions& opt = RunOptions()) {
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
    (void)extended_output;  // the means are produced whenever an output pointer is given
    return bnm_run_simulations(&c, corr, fic_penalties, G_list, w_EE_list, w_EI_list, w_IE_list,
                               w_IE_fic, SC, noise, noise_len, emp_FC_tril.data(),
                               (int)emp_FC_tril.size(), emp_FCD_tril.data(),
                               (int)emp_FCD_tril.size(), BOLD_ex, S_E, I_E, r_E, S_I, I_I, r_I,
                               FIC_failed, fic_n_trials, fc_trils, fcd_trils);
}

}  // namespace bnm
