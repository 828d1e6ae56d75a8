// Strict FP32 node update + FP32 Balloon-Windkessel (NodeStrict, BW<float>)
// against the FP64 update in the reference's order (Node<double>, BW<double>)
// for one node coupled to itself (c = s * S_E of the previous substep, s the
// row sum of a normalised SC row; FP32 path: s and the product in FP32, as
// the kernels form the dot), the same FP32 noise draws, ten substeps per ms,
// BW per ms, BOLD every TR = 1000 ms over the first 60 s.  Built on the host
// (tests/host_model/hip/hip_runtime.h: IEEE exp2 and division for the
// device's exp2 / reciprocal).  Prints the largest per-TR BOLD error
// relative to the trajectory's largest |BOLD| and the largest difference of
// the mean r_E over the 60 s, over all parameter sets.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>

#include "../../bnm_cuda_amd/csrc/model.hpp"

using namespace bnm;

int main(int argc, char** argv) {
    const int n_sets = argc > 1 ? std::atoi(argv[1]) : 64;
    std::mt19937_64 rng(11);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<float> Z(0.0f, 1.0f);
    double worst = 0.0, worst_r = 0.0;
    for (int k = 0; k < n_sets; ++k) {
        const double G = 0.5 + 2.5 * U(rng), wEE = 0.1 + 0.5 * U(rng), wEI = 0.1 + 0.5 * U(rng);
        const double wIE = 0.8 + 1.2 * U(rng), s = 0.5 + 1.5 * U(rng);
        const float s32 = (float)s;
        NodeStrict ns;
        ns.set_G(G);
        ns.set_w(wEE, wEI, wIE);
        float SE = 0.001f, SEl = NodeStrict::kS0l, SI = 0.001f, SIl = NodeStrict::kS0l;
        double SEd = 0.001, SId = 0.001;
        BW<float> bf;
        BW<double> bd;
        double bmax = 0.0, err = 0.0, sr = 0.0, srd = 0.0;
        for (int ms = 1; ms <= 60000; ++ms) {
            for (int sub = 0; sub < 10; ++sub) {
                const float xiE = Z(rng), xiI = Z(rng);
                float rE, rI, IE, II;
                ns.step(s32 * SE, xiE, xiI, SE, SEl, SI, SIl, rE, rI, &IE, &II);
                double IEd, IId, rEd, rId;
                Node<double>::step(s * SEd, G, wEE, wEI, wIE, (double)xiE, (double)xiI, SEd, SId, IEd, IId, rEd, rId);
                if (sub == 9) { sr += rE; srd += rEd; }
            }
            bf.step(SE);
            bd.step(SEd);
            if (ms % 1000 == 0) {
                const double b = bd.bold(C64::k1, C64::k2, C64::k3);
                bmax = std::max(bmax, std::fabs(b));
                err = std::max(err, std::fabs((double)bf.bold(C64::k1, C64::k2, C64::k3) - b));
            }
        }
        worst = std::max(worst, err / bmax);
        worst_r = std::max(worst_r, std::fabs(sr - srd) / 60000.0);
    }
    std::printf("per-TR BOLD rel %.3e mean rE diff %.3e Hz over %d sets\n", worst, worst_r, n_sets);
    return 0;
}
