// One simulation of N nodes on the host with the strict FP32 pieces of
// model.hpp that the FP32 kernels use by default -- NodeStrict (hi / lo
// coefficients, S + Sl), BW<float> (FP64 deviations, FP32 increments) and
// FIC in FP64 (fic_decide<double> on wIE kept in FP64, then split into the
// update's hi / lo pair) -- in the kernels' integration loop: the coupling
// dot as one FP32 FMA chain over k of the FP32 SC and the previous
// substep's S_E, ten substeps per ms, BW per ms, BOLD at TR, restart of
// every node when one is out of the FIC band.  Built on the host without
// contraction (tests/host_model/hip/hip_runtime.h: IEEE exp2 and division
// for the device's exp2 / reciprocal).
//
// Input file: as node_f64_loop.cpp (n_vols_remove is not used).  Output
// file: float64 bold[ots*N]; int32 trials, failed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../bnm_cuda_amd/csrc/model.hpp"

using namespace bnm;

template <class T>
static void rd(FILE* f, T* p, size_t n) {
    if (std::fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "short input\n");
        std::exit(2);
    }
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) return 2;
    int hdr[8];
    rd(in, hdr, 8);
    const int N = hdr[0], T = hdr[1], TR = hdr[2], i_start = hdr[4], i_end = hdr[5], max_trials = hdr[6];
    const long long period = hdr[7];
    double gd[2];
    rd(in, gd, 2);
    const double G = gd[0], init_delta = gd[1];
    std::vector<double> SC((size_t)N * N), wEE(N), wEI(N), wIE(N);
    rd(in, SC.data(), SC.size());
    rd(in, wEE.data(), N);
    rd(in, wEI.data(), N);
    rd(in, wIE.data(), N);
    std::vector<float> noise((size_t)period * 20 * N);
    rd(in, noise.data(), noise.size());
    std::fclose(in);

    std::vector<float> SC32(SC.size());
    for (size_t i = 0; i < SC.size(); ++i) SC32[i] = (float)SC[i];
    const int ots = T / TR + 1;
    std::vector<double> bold((size_t)ots * N, 0.0), mIE(N, 0.0), wIEd(wIE), delta(N, init_delta);
    std::vector<NodeStrict> ns(N);
    for (int j = 0; j < N; ++j) {
        ns[j].set_G(G);
        ns[j].set_w(wEE[j], wEI[j], wIEd[j]);
    }
    std::vector<float> SE(N, 0.001f), SEl(N, NodeStrict::kS0l), SI(N, 0.001f), SIl(N, NodeStrict::kS0l), Sp(N);
    std::vector<float> IE(N), II(N), rE(N), rI(N);
    std::vector<BW<float>> bw(N);
    const unsigned dur = (unsigned)(i_end - i_start + 1);
    int ts = 0, vol = 0, trial = 0;
    bool adjusting = true, failed = false;
    while (ts <= T) {
        const long long ms = ts % period;
        for (int s = 0; s < 10; ++s) {
            Sp = SE;
            for (int j = 0; j < N; ++j) {
                float c = 0.0f;
                for (int k = 0; k < N; ++k) c = fmaf(SC32[(size_t)j * N + k], Sp[k], c);
                const size_t b = (((size_t)ms * 10 + s) * N + j) * 2;
                ns[j].step(c, noise[b], noise[b + 1], SE[j], SEl[j], SI[j], SIl[j], rE[j], rI[j], &IE[j], &II[j]);
            }
        }
        for (int j = 0; j < N; ++j) bw[j].step(SE[j]);
        if (ts % TR == 0) {
            for (int j = 0; j < N; ++j) bold[(size_t)vol * N + j] = bw[j].bold(C64::k1, C64::k2, C64::k3);
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (ts >= i_start && ts <= i_end)
                for (int j = 0; j < N; ++j) mIE[j] += (double)IE[j];
            if (ts == i_end) {
                bool out = false;
                for (int j = 0; j < N; ++j) {
                    out |= fic_decide<double>(mIE[j], dur, trial, max_trials, wIEd[j], delta[j]);
                    ns[j].set_wIE(wIEd[j]);
                }
                if (out) {
                    if (trial < max_trials) {
                        for (int j = 0; j < N; ++j) {
                            SE[j] = SI[j] = 0.001f;
                            SEl[j] = SIl[j] = NodeStrict::kS0l;
                            bw[j].reset();
                            mIE[j] = 0.0;
                        }
                        ts = 0; vol = 0; ++trial;
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
    }
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(bold.data(), sizeof(double), bold.size(), out);
    const int tail[2] = {trial, failed ? 1 : 0};
    std::fwrite(tail, sizeof(int), 2, out);
    std::fclose(out);
    return 0;
}
