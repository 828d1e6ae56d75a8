// drop_in.cpp -- a run_CMAES-style caller of the C++ drop-in
// (include/bnm_run_simulations.hpp): one generation of a small population on
// the GPU, printing the cost, FIC penalty and FIC trial count per particle.
// Usage: drop_in <nodes> <n_sims> <time_steps_ms> [precision: 0=f64, 1=f32]
//                [sims_out_prefix: with n_sims = 1, save_output writes the
//                 reference's bold.txt, S_E.txt ... fic_ntrials.txt there]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../../include/bnm_run_simulations.hpp"

int main(int argc, char** argv) {
    const int N = argc > 1 ? std::atoi(argv[1]) : 20;
    const int B = argc > 2 ? std::atoi(argv[2]) : 4;
    const int T = argc > 3 ? std::atoi(argv[3]) : 60000;
    const int prec = argc > 4 ? std::atoi(argv[4]) : BNM_PREC_F64;
    const std::string prefix = argc > 5 ? argv[5] : "";
    const bool save = !prefix.empty();
    const int TR = 1000, ws = 10, step = 2, nvr = 10;
    // ring SC with unit row sums
    std::vector<double> SC((size_t)N * N, 0.0);
    for (int j = 0; j < N; ++j) {
        SC[(size_t)j * N + (j + 1) % N] = 0.5;
        SC[(size_t)j * N + (j + N - 1) % N] = 0.5;
    }
    std::vector<double> G(B), wEE((size_t)B * N), wEI((size_t)B * N), wIE((size_t)B * N, 0.0),
        wfic((size_t)B * N, 0.0);
    for (int b = 0; b < B; ++b) {
        G[b] = 0.5 + 0.25 * b;
        for (int j = 0; j < N; ++j) {
            wEE[(size_t)b * N + j] = 0.21;
            wEI[(size_t)b * N + j] = 0.15;
        }
    }
    const long long nlen = (long long)(T + 1) * 10 * N * 2;
    std::vector<float> noise(nlen);
    if (bnm_make_noise(1, nlen, noise.data())) return 1;
    bnm_config c;
    bnm_default_config(&c);
    c.nodes = N; c.n_sims = B; c.time_steps = T; c.bold_tr = TR; c.n_vols_remove = nvr;
    c.window_size = ws; c.window_step = step; c.exc_interhemispheric = N % 2 == 0;
    const int np = bnm_n_pairs(N, c.exc_interhemispheric);
    const int nw = bnm_n_windows(&c);
    std::vector<double> emp_fc(np), emp_fcd(nw * (nw - 1) / 2);
    for (int i = 0; i < np; ++i) emp_fc[i] = std::sin(0.37 * i);
    for (size_t i = 0; i < emp_fcd.size(); ++i) emp_fcd[i] = std::cos(0.11 * i);
    const int ots = bnm_output_ts(T, TR);
    std::vector<double> corr(B), pen(B), bold((size_t)B * ots * N), rE((size_t)B * N),
        fc((size_t)B * np), fcd((size_t)B * emp_fcd.size());
    std::vector<int> failed(B), trials(B);
    bnm::RunOptions opt;
    opt.window_step = step;
    opt.precision = prec;
    opt.fc_precision = prec == BNM_PREC_F64 ? BNM_FC_F64 : BNM_FC_F32;
    const int st = bnm::run_simulations(
        corr.data(), pen.data(), G.data(), wEE.data(), wEI.data(), wIE.data(), wfic.data(),
        SC.data(), N, true, T, TR, 10, false, emp_fc, emp_fcd, ws, save, true, prefix, B, noise.data(), nlen,
        bold.data(), nullptr, nullptr, rE.data(), nullptr, nullptr, nullptr, true, true,
        failed.data(), trials.data(), nvr, fc.data(), fcd.data(), opt);
    if (st) {
        std::fprintf(stderr, "run_simulations failed: %d %s\n", st, bnm_last_error());
        return 2;
    }
    for (int b = 0; b < B; ++b)
        std::printf("%d %.12f %.12f %d %d\n", b, corr[b], pen[b], trials[b], failed[b]);
    return 0;
}
