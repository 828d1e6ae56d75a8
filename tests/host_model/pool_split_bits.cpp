// Bitwise check of the pool-split node updates against the two-pool ones,
// on the host (model.hpp built with tests/host_model/hip/hip_runtime.h):
//   PoolStrict (one pool per lane, constants per lane)  vs NodeStrict::step
//   ei_step_f64 (one pool per lane, reference order)     vs Node<double>::step
// over random parameter sets and trajectories, FIC-like wIE changes, forced
// clamping at S = 0 and S = 1, and sweeps through x = 0 (the rate's series
// branch).  Prints the number of compared values and of mismatches.
#include <cstdio>
#include <cstring>
#include <random>

#include "../../bnm_cuda_amd/csrc/model.hpp"

using namespace bnm;

static long n_cmp = 0, n_bad = 0;

template <class T>
static void same(T a, T b, const char* what, long step) {
    ++n_cmp;
    if (std::memcmp(&a, &b, sizeof(T)) != 0) {
        if (n_bad < 10) std::printf("mismatch %s at %ld: %.17g vs %.17g\n", what, step, (double)a, (double)b);
        ++n_bad;
    }
}

struct F32Pair {
    NodeStrict ns;
    PoolStrict pe, pi;
    float SE = 0.001f, SEl = NodeStrict::kS0l, SI = 0.001f, SIl = NodeStrict::kS0l;   // two-pool state
    float SEo = 0.001f, SElo = NodeStrict::kS0l, SIo = 0.001f, SIlo = NodeStrict::kS0l;  // split state
    double G = 0, wEE = 0, wEI = 0;
    void init(double G_, double wEE_, double wEI_, double wIE) {
        G = G_; wEE = wEE_; wEI = wEI_;
        ns.set_G(G);
        ns.set_w(wEE, wEI, wIE);
        pe.init(0, G, wEE, wEI, wIE);
        pi.init(1, G, wEE, wEI, wIE);
    }
    void set_wIE(double wIE) { ns.set_wIE(wIE); pe.set_wIE(wIE); }
    void set_state(float se, float si) {
        SE = SEo = se; SI = SIo = si;
        SEl = SElo = SIl = SIlo = 0.0f;
    }
    void step(float c, float xiE, float xiI, long t) {
        float rE, rI, IE, II, rEo, rIo, IEo, IIo;
        ns.step(c, xiE, xiI, SE, SEl, SI, SIl, rE, rI, &IE, &II);
        const float SEv = SEo, SIv = SIo;   // both lanes read the previous S
        pe.step(c, xiE, SEv, SIv, SEo, SElo, rEo, IEo);
        pi.step(c, xiI, SEv, SIv, SIo, SIlo, rIo, IIo);
        same(SE, SEo, "f32 S_E", t); same(SEl, SElo, "f32 S_E lo", t);
        same(SI, SIo, "f32 S_I", t); same(SIl, SIlo, "f32 S_I lo", t);
        same(rE, rEo, "f32 r_E", t); same(rI, rIo, "f32 r_I", t);
        same(IE, IEo, "f32 I_E", t); same(II, IIo, "f32 I_I", t);
    }
};

struct F64Pair {
    double G = 0, wEE = 0, wEI = 0, wIE = 0;
    double SE = 0.001, SI = 0.001, SEo = 0.001, SIo = 0.001;
    void step(double c, double xiE, double xiI, long t) {
        double IE, II, rE, rI, IEo, IIo, rEo, rIo;
        Node<double>::step(c, G, wEE, wEI, wIE, xiE, xiI, SE, SI, IE, II, rE, rI);
        const double SEv = SEo, SIv = SIo;
        ei_step_f64(c, G, C64::wE_I0, wEE, wIE, C64::a_E, C64::b_E, C64::d_E, C64::gamma_E, 1.0, C64::tau_E, xiE,
                    SEv, SIv, SEo, IEo, rEo);
        ei_step_f64(c, 0.0, C64::wI_I0, wEI, 1.0, C64::a_I, C64::b_I, C64::d_I, C64::gamma_I, 0.0, C64::tau_I, xiI,
                    SEv, SIv, SIo, IIo, rIo);
        same(SE, SEo, "f64 S_E", t); same(SI, SIo, "f64 S_I", t);
        same(rE, rEo, "f64 r_E", t); same(rI, rIo, "f64 r_I", t);
        same(IE, IEo, "f64 I_E", t); same(II, IIo, "f64 I_I", t);
    }
};

int main() {
    std::mt19937_64 rng(20261203);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<float> Z(0.0f, 1.0f);
    long t = 0;
    // trajectories: coupling c = cs * S_E (a node whose neighbours follow
    // it), FIC-like wIE changes every 2000 substeps
    for (int set = 0; set < 300; ++set) {
        const double G = 0.5 + 3.0 * U(rng), wEE = 0.05 + 0.75 * U(rng), wEI = 0.05 + 0.75 * U(rng);
        double wIE = 0.5 + 2.0 * U(rng);
        const float cs = (float)(0.2 + 6.0 * U(rng));
        F32Pair f;
        f.init(G, wEE, wEI, wIE);
        F64Pair d;
        d.G = G; d.wEE = wEE; d.wEI = wEI; d.wIE = wIE;
        for (int s = 0; s < 6000; ++s, ++t) {
            if (s % 2000 == 1999) {
                wIE += (U(rng) < 0.5 ? -1 : 1) * 0.05 * U(rng);
                f.set_wIE(wIE);
                d.wIE = wIE;
            }
            const float xiE = Z(rng), xiI = Z(rng);
            f.step(cs * f.SE, xiE, xiI, t);
            d.step((double)cs * d.SE, (double)xiE, (double)xiI, t);
        }
    }
    // clamping: large draws drive S to 0 and 1
    for (int set = 0; set < 200; ++set) {
        const double G = 0.5 + 3.0 * U(rng), wEE = 0.05 + 0.75 * U(rng), wEI = 0.05 + 0.75 * U(rng);
        const double wIE = 0.5 + 2.0 * U(rng);
        F32Pair f;
        f.init(G, wEE, wEI, wIE);
        F64Pair d;
        d.G = G; d.wEE = wEE; d.wEI = wEI; d.wIE = wIE;
        for (int s = 0; s < 200; ++s, ++t) {
            const float big = (float)(s % 2 ? 5000.0 : -5000.0) * (float)U(rng);
            const float c = (float)(2.0 * U(rng));
            f.step(c, big, -big, t);
            d.step((double)c, (double)big, (double)-big, t);
        }
    }
    // x = 0 sweeps: the excitatory pool by the coupling, the inhibitory pool
    // by S_I, from either side of the series branch's |x| < 0.1
    for (int set = 0; set < 200; ++set) {
        const double G = 0.5 + 3.0 * U(rng), wEE = 0.05 + 0.75 * U(rng), wEI = 0.2 + 0.6 * U(rng);
        const double wIE = 0.5 + 2.0 * U(rng);
        const float se = (float)(0.1 + 0.8 * U(rng));
        const double si_e = 0.5 * U(rng);
        const double cE0 = (C64::b_E / C64::a_E - C64::wE_I0 - wEE * se + wIE * si_e) / (G * C64::J_NMDA);
        const double siI0 = C64::wI_I0 + wEI * se - C64::b_I / C64::a_I;
        for (int s = 0; s < 400; ++s, ++t) {
            const double off = (s - 200) * 2e-5;
            const double c = cE0 + off / (C64::a_E * G * C64::J_NMDA) * 100.0;
            const double si = siI0 - off / C64::a_I * 100.0;
            F32Pair f;
            f.init(G, wEE, wEI, wIE);
            f.set_state(se, (float)(c >= 0 ? si_e : (si > 0 && si < 1 ? si : si_e)));
            if (c >= 0) f.step((float)c, 0.3f, -0.2f, t);
            if (si > 0 && si < 1) {
                f.set_state(se, (float)si);
                f.step(0.5f, -0.1f, 0.4f, t);
            }
            F64Pair d;
            d.G = G; d.wEE = wEE; d.wEI = wEI; d.wIE = wIE;
            d.SE = d.SEo = se;
            d.SI = d.SIo = si_e;
            if (c >= 0) d.step(c, 0.3, -0.2, t);
            if (si > 0 && si < 1) {
                d.SE = d.SEo = se;
                d.SI = d.SIo = si;
                d.step(0.5, -0.1, 0.4, t);
            }
        }
    }
    std::printf("compared %ld values, mismatches %ld\n", n_cmp, n_bad);
    return n_bad == 0 ? 0 : 1;
}
