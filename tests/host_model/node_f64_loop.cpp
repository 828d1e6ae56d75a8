// One simulation of N nodes on the host with the FP64 pieces of model.hpp
// that the FP64 kernels use -- Node<double>::step (bnm.cpp:181-200),
// BW<double> (bnm.cpp:205-211) and fic_decide<double> (bnm.cpp:239-257) --
// in the integration loop every kernel runs: ten substeps per ms from the
// previous substep's S_E, BW per ms, BOLD and extended-output sums at
// ts % TR == 0, mean I_E over [i_start, i_end], restart of every node when
// any node is out of the FIC band.  The coupling dot runs over k in order.
// Built on the host without contraction (model.hpp through
// tests/host_model/hip/hip_runtime.h; exp and pow from the host's libm).
//
// Input file (native layout): int32 N, time_steps, bold_tr, n_vols_remove,
// i_start, i_end, max_fic_trials, noise_period_ms; float64 G, init_delta;
// float64 SC[N*N], wEE[N], wEI[N], wIE[N]; float32 noise[period*10*N*2]
// ([ms][substep][node][E, I]).  Output file: float64 bold[ots*N],
// ext[6*N] (means), wIE[N]; int32 trials, failed.
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
    const int N = hdr[0], T = hdr[1], TR = hdr[2], nrm = hdr[3], i_start = hdr[4], i_end = hdr[5];
    const int max_trials = hdr[6];
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

    const int ots = T / TR + 1;
    std::vector<double> bold((size_t)ots * N, 0.0), ext(6 * (size_t)N, 0.0);
    std::vector<double> SE(N, 0.001), SI(N, 0.001), Sp(N), IE(N), II(N), rE(N), rI(N), mIE(N, 0.0), delta(N, init_delta);
    std::vector<BW<double>> bw(N);
    const unsigned dur = (unsigned)(i_end - i_start + 1);
    int ts = 0, vol = 0, trial = 0;
    bool adjusting = true, failed = false;
    while (ts <= T) {
        const long long ms = ts % period;
        for (int s = 0; s < 10; ++s) {
            Sp = SE;
            for (int j = 0; j < N; ++j) {
                double c = 0.0;
                for (int k = 0; k < N; ++k) c += SC[(size_t)j * N + k] * Sp[k];
                const size_t b = (((size_t)ms * 10 + s) * N + j) * 2;
                Node<double>::step(c, G, wEE[j], wEI[j], wIE[j], (double)noise[b], (double)noise[b + 1], SE[j], SI[j],
                                   IE[j], II[j], rE[j], rI[j]);
            }
        }
        for (int j = 0; j < N; ++j) bw[j].step(SE[j]);
        if (ts % TR == 0) {
            for (int j = 0; j < N; ++j) {
                bold[(size_t)vol * N + j] = bw[j].bold(C64::k1, C64::k2, C64::k3);
                if (vol >= nrm) {
                    const double v[6] = {SE[j], IE[j], rE[j], SI[j], II[j], rI[j]};
                    for (int q = 0; q < 6; ++q) ext[(size_t)q * N + j] += v[q];
                }
            }
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (ts >= i_start && ts <= i_end)
                for (int j = 0; j < N; ++j) mIE[j] += IE[j];
            if (ts == i_end) {
                bool out = false;
                for (int j = 0; j < N; ++j) out |= fic_decide<double>(mIE[j], dur, trial, max_trials, wIE[j], delta[j]);
                if (out) {
                    if (trial < max_trials) {
                        for (int j = 0; j < N; ++j) {
                            SE[j] = SI[j] = 0.001;
                            bw[j].reset();
                            mIE[j] = 0.0;
                        }
                        for (double& e : ext) e = 0.0;
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
    const int n = vol - nrm;
    for (double& e : ext) e = n > 0 ? e / n : 0.0;
    FILE* out = std::fopen(argv[2], "wb");
    if (!out) return 2;
    std::fwrite(bold.data(), sizeof(double), bold.size(), out);
    std::fwrite(ext.data(), sizeof(double), ext.size(), out);
    std::fwrite(wIE.data(), sizeof(double), N, out);
    const int tail[2] = {trial, failed ? 1 : 0};
    std::fwrite(tail, sizeof(int), 2, out);
    std::fclose(out);
    return 0;
}
