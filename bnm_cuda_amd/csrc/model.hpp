// model.hpp -- reduced Wong-Wang (Deco 2014) + Balloon-Windkessel constants and
// per-node update steps, shared by every integration kernel.
//
// Constants follow constants.h:27-90 of the reference.  Two numeric flavours:
//  * Exact<double>: the reference CPU expression order (bnm.cpp:181-211) with
//    libm-equivalent exp/pow, used by the FP64 kernels so that their BOLD can be
//    compared against the FP64 CPU oracle at 1e-6 relative.
//  * Fast<float>: hardware exp2/log2/rcp (v_exp_f32, v_log_f32, v_rcp_f32) with
//    a series branch for the rate function's removable singularity.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

namespace bnm {

// constants.h:27-90
struct C64 {
    static constexpr double dt = 0.1;
    static constexpr double model_dt = 0.001;
    static constexpr double sqrt_dt = 0.316227;       // truncated, as in constants.h:29
    static constexpr double J_NMDA = 0.15;
    static constexpr double a_E = 310, b_E = 125, d_E = 0.16;
    static constexpr double a_I = 615, b_I = 177, d_I = 0.087;
    static constexpr double b_a_ratio_E = 125.0 / 310.0;
    static constexpr double gamma_E = 0.641 / 1000.0;
    static constexpr double gamma_I = 1.0 / 1000.0;
    static constexpr double tau_E = 100, tau_I = 10;
    static constexpr double itau_E = 1.0 / 100, itau_I = 1.0 / 10;
    static constexpr double sigma = 0.01;
    static constexpr double wE_I0 = 1.0 * 0.382;
    static constexpr double wI_I0 = 0.7 * 0.382;
    static constexpr double rho = 0.34, alpha = 0.32, tau = 0.98;
    static constexpr double y = 1.0 / 0.41, kappa = 1.0 / 0.65;
    static constexpr double V0 = 0.02 * 100;
    static constexpr double ialpha = 1.0 / 0.32, itau = 1.0 / 0.98, oneminrho = 1.0 - 0.34;
    // BOLD weights: 3T values of Demirtas 2019 (the reference's default) and,
    // with BW_FRISTON2003, Friston 2003's 7 rho, 2, 2 rho - 0.2
    static constexpr double k1 = 3.72, k2 = 0.527, k3 = 0.53;
    static constexpr double k1_F = 7 * 0.34, k2_F = 2.0, k3_F = 2 * 0.34 - 0.2;
    // FIC steady state of the isolated node (constants.h:57-68)
    static constexpr double tau_I_s = 0.01, gamma_I_s = 1.0;
    static constexpr double I_E_ss = 0.3773805650, S_E_ss = 0.1647572075;
};

// log2(e) and log2(1 - rho) for the exp2/log2 forms
constexpr double kLog2e = 1.4426950408889634;
constexpr double kLog2OneMinRho = -0.5994620704162712;  // log2(0.66)

// Per-node Euler-Maruyama substep.  In: coupling c = sum_k SC[j,k] S_E[k]
// (previous substep), noise draws xiE/xiI.  Updates S_E/S_I, returns I_E, I_I,
// r_E, r_I of this substep (the values the reference records at TR samples).
template <class R> struct Node;

// FP64, reference expression order of bnm.cpp:181-200 (IEEE ops; the kernels
// that use it compile with contraction off).
template <> struct Node<double> {
    static __device__ __forceinline__ void step(double c, double G, double wEE, double wEI,
                                                double wIE, double xiE, double xiI, double& SE,
                                                double& SI, double& IE, double& II, double& rE,
                                                double& rI) {
#pragma clang fp contract(off)
        IE = C64::wE_I0 + wEE * SE + c * G * C64::J_NMDA - wIE * SI;
        II = C64::wI_I0 + wEI * SE - SI;
        const double eE = exp(-C64::d_E * (C64::a_E * IE - C64::b_E));
        const double eI = exp(-C64::d_I * (C64::a_I * II - C64::b_I));
        rE = (C64::a_E * IE - C64::b_E) / (1 - eE);
        rI = (C64::a_I * II - C64::b_I) / (1 - eI);
        const double dSE = xiE * C64::sigma * C64::sqrt_dt +
                           C64::dt * ((1 - SE) * C64::gamma_E * rE - (SE / C64::tau_E));
        const double dSI = xiI * C64::sigma * C64::sqrt_dt +
                           C64::dt * (C64::gamma_I * rI - (SI / C64::tau_I));
        SE = fmax(0.0, fmin(SE + dSE, 1.0));
        SI = fmax(0.0, fmin(SI + dSI, 1.0));
    }
};

// FP64 update of one pool with the pool's constants as arguments (the FP64
// E/I kernels: lanes that each step one pool).  Excitatory: k0 = wE_I0,
// kSE = wEE, Gl = G, kSI = wIE, m = 1; inhibitory: k0 = wI_I0, kSE = wEI,
// Gl = 0, kSI = 1, m = 0, whose extra terms are exact ((c*0)*J = +0 for
// c >= 0, 1*S_I, (1 - 0*S)*gamma = gamma), so each pool evaluates
// Node<double>'s expression for it.
__device__ __forceinline__ void ei_step_f64(double c, double Gl, double k0, double kSE, double kSI, double aa,
                                            double bb, double dd, double g, double m, double tau, double xi,
                                            double SEv, double SIv, double& S, double& I, double& r) {
#pragma clang fp contract(off)
    I = k0 + kSE * SEv + c * Gl * C64::J_NMDA - kSI * SIv;
    const double x = aa * I - bb;
    const double e = exp(-dd * x);
    r = x / (1 - e);
    const double dS = xi * C64::sigma * C64::sqrt_dt + C64::dt * ((1 - m * S) * g * r - (S / tau));
    S = fmax(0.0, fmin(S + dS, 1.0));
}

// x / (1 - exp(-d x)), FP32.  Near x = 0 the quotient is 1/d + x/2 + d x^2/12
// (series of the removable singularity); the series is used for |x| < 0.1,
// where the direct form would lose digits to cancellation.
template <int which>
__device__ __forceinline__ float rate_f32(float x) {
    constexpr float nd_log2e = (float)(-(which == 0 ? C64::d_E : C64::d_I) * kLog2e);
    constexpr float inv_d = (float)(1.0 / (which == 0 ? C64::d_E : C64::d_I));
    constexpr float d12 = (float)((which == 0 ? C64::d_E : C64::d_I) / 12.0);
    const float e = __builtin_amdgcn_exp2f(nd_log2e * x);
    const float direct = x * __builtin_amdgcn_rcpf(1.0f - e);
    const float series = inv_d + x * (0.5f + d12 * x);
    return fabsf(x) < 0.1f ? series : direct;
}

template <> struct Node<float> {
    static __device__ __forceinline__ void step(float c, float GJ, float wEE, float wEI, float wIE,
                                                float xiE, float xiI, float& SE, float& SI,
                                                float& IE, float& II, float& rE, float& rI) {
        constexpr float nz = (float)(C64::sigma * C64::sqrt_dt);
        constexpr float dt_gE = (float)(C64::dt * C64::gamma_E);
        constexpr float dt_itE = (float)(C64::dt / C64::tau_E);
        constexpr float dt_gI = (float)(C64::dt * C64::gamma_I);
        constexpr float dt_itI = (float)(C64::dt / C64::tau_I);
        IE = (float)C64::wE_I0 + wEE * SE + c * GJ - wIE * SI;
        II = (float)C64::wI_I0 + wEI * SE - SI;
        rE = rate_f32<0>((float)C64::a_E * IE - (float)C64::b_E);
        rI = rate_f32<1>((float)C64::a_I * II - (float)C64::b_I);
        const float dSE = xiE * nz + (dt_gE * rE - dt_gE * rE * SE - dt_itE * SE);
        const float dSI = xiI * nz + (dt_gI * rI - dt_itI * SI);
        SE = __builtin_amdgcn_fmed3f(SE + dSE, 0.0f, 1.0f);
        SI = __builtin_amdgcn_fmed3f(SI + dSI, 0.0f, 1.0f);
    }
};

// Strict FP32 node update, both pools in one lane: per pool the expressions of
// the E/I kernel's strict short-chain form (k_sim_ei, FORM 3 | 64).  Every
// coefficient that multiplies a state value is an FP32 hi + lo pair of its
// FP64 value (w I_0, wEE / wEI, wIE, a G J_NMDA, dt gamma, dt / tau,
// sigma sqrt_dt, -d log2 e), the exponent is taken from the rounded x, and
// each S is an unevaluated sum S + Sl (the increment's rounding carried in
// Sl by Fast2Sum, starting from the exact 0.001: kS0l).  What is left is the
// FP32 rounding of x, exp2 and the reciprocal in the rate.
struct NodeStrict {
    float wEEh = 0, wEEl = 0, wEIh = 0, wEIl = 0, wIEh = 0, wIEl = 0;   // per node
    float aGJh = 0, aGJl = 0, GJ = 0;                                    // per simulation
    static constexpr float kS0l = (float)(0.001 - (double)0.001f);

    static __device__ __forceinline__ void split(double v, float& h, float& l) {
        h = (float)v;
        l = (float)(v - (double)h);
    }
    // per simulation (the same in every lane of a workgroup)
    __device__ __forceinline__ void set_G(double G) {
        split(C64::a_E * (G * C64::J_NMDA), aGJh, aGJl);
        GJ = (float)(G * C64::J_NMDA);
    }
    __device__ __forceinline__ void set_w(double wEE, double wEI, double wIE) {
        split(wEE, wEEh, wEEl);
        split(wEI, wEIh, wEIl);
        split(wIE, wIEh, wIEl);
    }
    __device__ __forceinline__ void set_wIE(double wIE) { split(wIE, wIEh, wIEl); }

    // one pool: x from pre / prel and the coupling, the rate, and S + Sl
    // advanced by g r (1 - m S) - S dt / tau + noise
    template <int P>
    static __device__ __forceinline__ float pool(float pre, float prel, float c, float ch, float cl, float xi,
                                                 float& S, float& Sl) {
        constexpr double a64 = P == 0 ? C64::a_E : C64::a_I, b64 = P == 0 ? C64::b_E : C64::b_I;
        constexpr double d64 = P == 0 ? C64::d_E : C64::d_I;
        constexpr float aa = (float)a64, bb = (float)b64;
        constexpr float ndlh = (float)(-d64 * kLog2e), ndll = (float)(-d64 * kLog2e - (double)ndlh);
        constexpr float invd = (float)(1.0 / d64), d12 = (float)(d64 / 12.0);
        constexpr double g64 = C64::dt * (P == 0 ? C64::gamma_E : C64::gamma_I);
        constexpr double it64 = C64::dt / (P == 0 ? C64::tau_E : C64::tau_I);
        constexpr float g = (float)g64, gl = (float)(g64 - (double)g);
        constexpr float gm = P == 0 ? g : 0.0f, gml = P == 0 ? gl : 0.0f;
        constexpr float it = (float)it64, itl = (float)(it64 - (double)it);
        constexpr double nzs64 = C64::sigma * C64::sqrt_dt;
        constexpr float nzs = (float)nzs64, nzsl = (float)(nzs64 - (double)nzs);
        const float X0s = fmaf(aa, prel, fmaf(aa, pre, -bb));
        float x = X0s;   // the inhibitory pool has no long-range coupling
        if constexpr (P == 0) x = fmaf(ch, c, fmaf(cl, c, X0s));
        const float y = fmaf(ndlh, x, ndll * x);
        const float direct = x * __builtin_amdgcn_rcpf(1.0f - __builtin_amdgcn_exp2f(y));
        const float series = invd + x * (0.5f + d12 * x);
        const float r = fabsf(x) < 0.1f ? series : direct;
        const float Ac = fmaf(-gm, S, g), Acl = fmaf(-gml, S, gl);
        const float D = fmaf(xi, nzsl, fmaf(-itl, S, fmaf(-it, S, xi * nzs))) + Sl;
        const float u = fmaf(r, Ac, fmaf(r, Acl, D));
        const float t = S + u;
        const float tc = __builtin_amdgcn_fmed3f(t, 0.0f, 1.0f);
        Sl = tc == t ? u - (t - S) : 0.0f;
        S = tc;
        return r;
    }

    // one substep from the previous S values; I_E / I_I (hi + lo, rounded)
    // only when asked (their uses are once per ms)
    __device__ __forceinline__ void step(float c, float xiE, float xiI, float& SE, float& SEl, float& SI, float& SIl,
                                         float& rE, float& rI, float* IE, float* II) const {
        constexpr float kE0 = (float)C64::wE_I0, kE0l = (float)(C64::wE_I0 - (double)kE0);
        constexpr float kI0 = (float)C64::wI_I0, kI0l = (float)(C64::wI_I0 - (double)kI0);
        const float SEv = SE, SIv = SI;
        const float preE = fmaf(-wIEh, SIv, fmaf(wEEh, SEv, kE0));
        const float prelE = fmaf(-wIEl, SIv, fmaf(wEEl, SEv, kE0l));
        const float preI = fmaf(-1.0f, SIv, fmaf(wEIh, SEv, kI0));
        const float prelI = fmaf(wEIl, SEv, kI0l);
        rE = pool<0>(preE, prelE, c, aGJh, aGJl, xiE, SE, SEl);
        rI = pool<1>(preI, prelI, c, 0.0f, 0.0f, xiI, SI, SIl);
        if (IE) *IE = fmaf(GJ, c, preE) + prelE;
        if (II) *II = preI + prelI;
    }
};

// NodeStrict's update of one pool, with the pool chosen per lane at run time
// (excitatory or inhibitory constants held in registers), so that the two
// lanes of a node run one instruction stream, one pool each.  The same
// expressions as NodeStrict::pool<P>: for the inhibitory pool the coupling
// and wIE terms have zero coefficients (c and S finite), which adds exact
// zeros, so both give the same bits.
struct PoolStrict {
    float k0h = 0, k0l = 0, kSEh = 0, kSEl = 0, kSIh = 0, kSIl = 0;   // pre = k0 + kSE S_E - kSI S_I
    float ch = 0, cl = 0, GJp = 0;                                     // coupling (excitatory only)
    float aa = 0, bb = 0, ndlh = 0, ndll = 0, invd = 0, d12 = 0;
    float g = 0, gl = 0, gm = 0, gml = 0, it = 0, itl = 0;
    __device__ __forceinline__ void init(int pop, double G, double wEE, double wEI, double wIE) {
        const bool e = pop == 0;
        const double k0 = e ? C64::wE_I0 : C64::wI_I0, a = e ? C64::a_E : C64::a_I, b = e ? C64::b_E : C64::b_I;
        const double d = e ? C64::d_E : C64::d_I;
        const double g64 = C64::dt * (e ? C64::gamma_E : C64::gamma_I), it64 = C64::dt / (e ? C64::tau_E : C64::tau_I);
        NodeStrict::split(k0, k0h, k0l);
        NodeStrict::split(e ? wEE : wEI, kSEh, kSEl);
        NodeStrict::split(e ? wIE : 1.0, kSIh, kSIl);
        NodeStrict::split(e ? a * (G * C64::J_NMDA) : 0.0, ch, cl);
        GJp = e ? (float)(G * C64::J_NMDA) : 0.0f;
        aa = (float)a;
        bb = (float)b;
        NodeStrict::split(-d * kLog2e, ndlh, ndll);
        invd = (float)(1.0 / d);
        d12 = (float)(d / 12.0);
        NodeStrict::split(g64, g, gl);
        gm = e ? g : 0.0f;
        gml = e ? gl : 0.0f;
        NodeStrict::split(it64, it, itl);
    }
    __device__ __forceinline__ void set_wIE(double wIE) { NodeStrict::split(wIE, kSIh, kSIl); }

    // one substep of this lane's pool from the previous S_E, S_I; returns r
    // and the pool's input current I
    __device__ __forceinline__ void step(float c, float xi, float SEv, float SIv, float& S, float& Sl, float& r,
                                         float& I) const {
        constexpr double nzs64 = C64::sigma * C64::sqrt_dt;
        constexpr float nzs = (float)nzs64, nzsl = (float)(nzs64 - (double)nzs);
        const float pre = fmaf(-kSIh, SIv, fmaf(kSEh, SEv, k0h));
        const float prel = fmaf(-kSIl, SIv, fmaf(kSEl, SEv, k0l));
        const float X0s = fmaf(aa, prel, fmaf(aa, pre, -bb));
        const float x = fmaf(ch, c, fmaf(cl, c, X0s));
        const float y = fmaf(ndlh, x, ndll * x);
        const float direct = x * __builtin_amdgcn_rcpf(1.0f - __builtin_amdgcn_exp2f(y));
        const float series = invd + x * (0.5f + d12 * x);
        r = fabsf(x) < 0.1f ? series : direct;
        const float Ac = fmaf(-gm, S, g), Acl = fmaf(-gml, S, gl);
        const float D = fmaf(xi, nzsl, fmaf(-itl, S, fmaf(-it, S, xi * nzs))) + Sl;
        const float u = fmaf(r, Ac, fmaf(r, Acl, D));
        const float t = S + u;
        const float tc = __builtin_amdgcn_fmed3f(t, 0.0f, 1.0f);
        Sl = tc == t ? u - (t - S) : 0.0f;
        S = tc;
        I = fmaf(GJp, c, pre) + prel;
    }
};

// Balloon-Windkessel state; one Euler step of 1 ms (bnm.cpp:205-211), update
// order: x from old f, f from new x, nu and q from old f and (q) new nu.
template <class R> struct BW;

template <> struct BW<double> {
    double x = 0.0, f = 1.0, nu = 1.0, q = 1.0;
    __device__ __forceinline__ void reset() { x = 0.0; f = 1.0; nu = 1.0; q = 1.0; }
    __device__ __forceinline__ void step(double SE) {
#pragma clang fp contract(off)
        x = x + C64::model_dt * (SE - C64::kappa * x - C64::y * (f - 1.0));
        const double f_new = f + C64::model_dt * x;
        nu = nu + C64::model_dt * C64::itau * (f - pow(nu, C64::ialpha));
        q = q + C64::model_dt * C64::itau *
                    (f * (1.0 - pow(C64::oneminrho, (1.0 / f))) / C64::rho -
                     pow(nu, C64::ialpha) * q / nu);
        f = f_new;
    }
    __device__ __forceinline__ double bold(double k1, double k2, double k3) const {
#pragma clang fp contract(off)
        return C64::V0 * (k1 * (1 - q) + k2 * (1 - q / nu) + k3 * (1 - nu));
    }
};

// FP32 flavour.  f, nu and q sit near 1 and BOLD is formed from 1 - q,
// 1 - q/nu and 1 - nu, so the step works on the deviations f - 1, nu - 1,
// q - 1 (kept in FP64; the increments are formed in FP32):
//   nu^(1/alpha) - 1 = expm1(ln(1 + u) / alpha),        u = nu - 1
//   (1-rho)^(1/f) = (1-rho) (1 + expm1(-ln(1-rho) (f-1)/f))
//   f (1 - (1-rho)^(1/f)) / rho - 1 = (f-1) - f (1-rho)/rho expm1(..)
//   nu^(1/alpha) q / nu - 1 = (E + (q-1) + E (q-1) - u) / (1 + u)
// with ln(1+u) = 2 atanh(u / (2+u)) (odd series) and expm1 as a degree-9
// Taylor polynomial: each term then carries FP32 relative error of a small
// quantity instead of the ~1e-7 absolute error of v_log_f32 / v_exp_f32 near
// 1, which the 1 - q cancellation in BOLD would turn into ~1e-6 relative.
// nu^(1/alpha) at the updated nu (the q step uses the new nu, bnm.cpp:209)
// follows by a first-order step from the old one (|dnu| ~ 1e-6: the
// second-order term is ~1e-12).  Arguments outside the series' range fall
// back to the hardware functions.
__device__ __forceinline__ float bw_expm1(float z) {
    if (fabsf(z) > 0.5f) return __builtin_amdgcn_exp2f(z * (float)kLog2e) - 1.0f;
    float p = 1.0f / 362880.0f;
    p = fmaf(p, z, 1.0f / 40320.0f);
    p = fmaf(p, z, 1.0f / 5040.0f);
    p = fmaf(p, z, 1.0f / 720.0f);
    p = fmaf(p, z, 1.0f / 120.0f);
    p = fmaf(p, z, 1.0f / 24.0f);
    p = fmaf(p, z, 1.0f / 6.0f);
    p = fmaf(p, z, 0.5f);
    p = fmaf(p, z, 1.0f);
    return p * z;
}

__device__ __forceinline__ float bw_log1p(float u) {
    const float s = u * __builtin_amdgcn_rcpf(2.0f + u);
    if (fabsf(s) > 0.15f) return __builtin_amdgcn_logf(1.0f + u) * (float)(1.0 / kLog2e);
    const float s2 = s * s;
    float p = 1.0f / 11.0f;
    p = fmaf(p, s2, 1.0f / 9.0f);
    p = fmaf(p, s2, 1.0f / 7.0f);
    p = fmaf(p, s2, 1.0f / 5.0f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    p = fmaf(p, s2, 1.0f);
    return 2.0f * s * p;
}

template <> struct BW<float> {
    double x = 0.0, fm = 0.0, um = 0.0, qm = 0.0;   // x, f - 1, nu - 1, q - 1
    __device__ __forceinline__ void reset() { x = 0.0; fm = 0.0; um = 0.0; qm = 0.0; }
    __device__ __forceinline__ void step(float SE) {
        constexpr float mdt = (float)C64::model_dt;
        constexpr float mdt_itau = (float)(C64::model_dt * C64::itau);
        constexpr float kap = (float)C64::kappa, yy = (float)C64::y;
        constexpr float ia = (float)C64::ialpha;
        constexpr float c1 = (float)(C64::oneminrho / C64::rho);
        constexpr float c2 = (float)(-kLog2OneMinRho / kLog2e);   // -ln(1 - rho)
        const float xf = (float)x, f1 = (float)fm, u = (float)um, q1 = (float)qm;
        x += (double)(mdt * (SE - kap * xf - yy * f1));
        const double fm_new = fm + (double)(mdt * (float)x);
        const float ru = __builtin_amdgcn_rcpf(1.0f + u);
        const float E = bw_expm1(ia * bw_log1p(u));          // nu^(1/alpha) - 1
        const float dnu = mdt_itau * (f1 - E);
        um += (double)dnu;
        const float un = (float)um;
        const float En = fmaf(ia * (1.0f + E) * ru, dnu, E);  // at the new nu
        const float em = bw_expm1(c2 * f1 * __builtin_amdgcn_rcpf(1.0f + f1));
        const float t1 = f1 - (1.0f + f1) * c1 * em;
        const float t2 = (En + q1 + En * q1 - un) * __builtin_amdgcn_rcpf(1.0f + un);
        qm += (double)(mdt_itau * (t1 - t2));
        fm = fm_new;
    }
    __device__ __forceinline__ float bold(double k1, double k2, double k3) const {
        // V0 (k1 (1 - q) + k2 (1 - q/nu) + k3 (1 - nu)), 1 - q/nu = (u - q1)/(1 + u)
        return (float)(C64::V0 * (-k1 * qm + k2 * (um - qm) / (1.0 + um) - k3 * um));
    }
};

// Numerical FIC decision for one node (bnm.cpp:239-257): returns true when the
// node's mean I_E is out of band; adjusts wIE/delta when a trial is left.
// FP64 delta follows the CPU path's float rounding of delta after a down-step
// (bnm.cpp:250 fmaxf).
template <class R>
__device__ __forceinline__ bool fic_decide(double mean_IE_sum, unsigned duration, int trial,
                                           int max_trials, R& wIE, R& delta) {
    const double mean = mean_IE_sum / duration;
    const double diff = mean - C64::b_a_ratio_E;
    if (fabs(diff + 0.026) > 0.005) {
        if (trial < max_trials) {
            if (diff < -0.026) {
                wIE -= delta;
                delta -= (R)0.001;
                delta = (R)fmaxf((float)delta, 0.001f);
            } else {
                wIE += delta;
            }
        }
        return true;
    }
    return false;
}

// Per-launch integration settings (host fills, passed by value).
struct SimArgs {
    int N;                 // nodes
    int B;                 // simulations in this launch
    int time_steps;        // ms, loop runs ts = 0..time_steps
    int bold_tr;           // ms
    int output_ts;         // BOLD samples per simulation
    int n_vols_remove;
    int adjust_fic;        // numerical FIC on
    int max_fic_trials;
    int i_start, i_end;    // FIC sampling window (ms, on the post-increment clock)
    double init_delta;
    long long noise_period;   // ms of draws in the noise array
    const float* noise;       // [period][10][N][2]
    const double* G;          // [B]
    const double* wEE;        // [B*N]
    const double* wEI;        // [B*N]
    const double* wIE_in;     // [B*N] initial wIE
    double* wIE_out;          // [B*N] adjusted wIE (may alias wIE_in)
    void* bold;               // [B][output_ts][N] in the kernel precision
    double* ext;              // [B][6][N] means of S_E, I_E, r_E, S_I, I_I, r_I
    int* fic_failed;          // [B]
    int* fic_trials;          // [B]
    double bw_k1, bw_k2, bw_k3;   // BOLD weights (C64::k1..k3 or k1_F..k3_F)
};

}  // namespace bnm
