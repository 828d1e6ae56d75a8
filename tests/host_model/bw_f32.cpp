// FP32 Balloon-Windkessel (BW<float>: FP64 deviations f - 1, nu - 1, q - 1
// with FP32 increments, series expm1 / log1p) against the FP64 step in the
// reference's order (BW<double>), on the host (model.hpp built with
// tests/host_model/hip/hip_runtime.h).  Both are driven by the same S_E
// trajectory, one step per ms, BOLD sampled every TR = 1000 steps over the
// first 60 s; each trajectory is an Ornstein-Uhlenbeck S_E in [0, 1] around
// a random level, as the integration produces.  Prints the largest per-TR
// error relative to the trajectory's largest |BOLD|, over all trajectories,
// and the same with the weights of the Friston 2003 variant.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>

#include "../../bnm_cuda_amd/csrc/model.hpp"

using namespace bnm;

int main(int argc, char** argv) {
    const int n_traj = argc > 1 ? std::atoi(argv[1]) : 200;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> Z(0.0, 1.0);
    double worst = 0.0, worst_f = 0.0;
    for (int k = 0; k < n_traj; ++k) {
        const double level = 0.05 + 0.4 * U(rng), amp = 0.02 + 0.2 * U(rng), tau = 50.0 + 2000.0 * U(rng);
        BW<double> d;
        BW<float> f;
        double S = level, bmax = 0.0, bmax_f = 0.0, err = 0.0, err_f = 0.0;
        for (int ms = 1; ms <= 60000; ++ms) {
            S += (level - S) / tau + amp * std::sqrt(2.0 / tau) * Z(rng);
            S = std::min(1.0, std::max(0.0, S));
            const float Sf = (float)S;   // the kernels keep S_E in FP32
            d.step((double)Sf);
            f.step(Sf);
            if (ms % 1000 == 0) {
                const double b = d.bold(C64::k1, C64::k2, C64::k3), bf = f.bold(C64::k1, C64::k2, C64::k3);
                const double c = d.bold(C64::k1_F, C64::k2_F, C64::k3_F), cf = f.bold(C64::k1_F, C64::k2_F, C64::k3_F);
                bmax = std::max(bmax, std::fabs(b));
                bmax_f = std::max(bmax_f, std::fabs(c));
                err = std::max(err, std::fabs(bf - b));
                err_f = std::max(err_f, std::fabs(cf - c));
            }
        }
        worst = std::max(worst, err / bmax);
        worst_f = std::max(worst_f, err_f / bmax_f);
    }
    std::printf("per-TR BOLD rel %.3e friston %.3e over %d trajectories\n", worst, worst_f, n_traj);
    return 0;
}
