// fic_kernel.hip -- analytical feedback-inhibition control on the device
// (analytical_fic of fic.cpp:130-207, which the reference's driver runs on the
// host per particle, run_CMAES.cpp:144-174).
//
// For node j: solve I_I = w_I*I_0 + wEI_j*S_E_ss - gamma_I*tau_I*phi_I(I_I) on
// [0, 2] by Brent's method (R. P. Brent 1973, ch. 4) with the interval test
// |hi - lo| < 1e-3 * min(|lo|, |hi|) and at most 100 iterations (fic.cpp:71-100);
// then S_I = phi_I(I_I)*tau_I*gamma_I and
//   wIE_j = -(1/S_I) * (I_E_ss - I_ext - w_E*I_0 - sum_k K_EE[j,k]*S_E_ss),
//   K_EE = G*J_NMDA*SC + diag(wEE).
// A failed root or any wIE_j < 0 marks the simulation unstable and sets wIE = 1
// on every node (fic.cpp:183-186, 199-202; run_CMAES.cpp:162-166).
// One workgroup per simulation, FP64, contraction off.
#include "kernels.hpp"

namespace bnm {

namespace {

__device__ __forceinline__ double phi_I(double II) {
#pragma clang fp contract(off)
    return ((C64::a_I * II) - C64::b_I) / (1 - exp(-1 * C64::d_I * ((C64::a_I * II) - C64::b_I)));
}

__device__ __forceinline__ double inh_fixed_point(double x, double wEI) {
#pragma clang fp contract(off)
    return C64::wI_I0 + wEI * C64::S_E_ss - 1.0 * C64::gamma_I_s * C64::tau_I_s * phi_I(x) - x;
}

__device__ double brent_inh_root(double wEI) {
#pragma clang fp contract(off)
    const double epsrel = 1e-3, eps = 2.220446049250313e-16;
    double a = 0.0, b = 2.0;
    double fa = inh_fixed_point(a, wEI), fb = inh_fixed_point(b, wEI);
    if ((fa < 0.0 && fb < 0.0) || (fa > 0.0 && fb > 0.0)) return -1;
    double c = b, fc = fb, d = b - a, e = b - a;
    for (int iter = 1; iter <= 100; ++iter) {
        if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
            c = a; fc = fa; d = b - a; e = b - a;
        }
        if (fabs(fc) < fabs(fb)) {
            a = b; b = c; c = a;
            fa = fb; fb = fc; fc = fa;
        }
        const double tol = 0.5 * eps * fabs(b);
        const double m = 0.5 * (c - b);
        if (fb == 0 || fabs(m) <= tol) return b;
        if (fabs(e) < tol || fabs(fa) <= fabs(fb)) {
            d = m; e = m;
        } else {
            double p, q, r;
            const double s = fb / fa;
            if (a == c) {
                p = 2 * m * s;
                q = 1 - s;
            } else {
                q = fa / fc;
                r = fb / fc;
                p = s * (2 * m * q * (q - r) - (b - a) * (r - 1));
                q = (q - 1) * (r - 1) * (s - 1);
            }
            if (p > 0) q = -q; else p = -p;
            if (2 * p < fmin(3 * m * q - fabs(tol * q), fabs(e * q))) {
                e = d; d = p / q;
            } else {
                d = m; e = m;
            }
        }
        a = b; fa = fb;
        if (fabs(d) > tol) b += d; else b += (m > 0 ? tol : -tol);
        fb = inh_fixed_point(b, wEI);
        if ((fb < 0 && fc < 0) || (fb > 0 && fc > 0)) {
            c = a; fc = fa; d = b - a; e = b - a;
        }
        const double lo = fmin(b, c), hi = fmax(b, c);
        const double min_abs = ((lo > 0 && hi > 0) || (lo < 0 && hi < 0)) ? fmin(fabs(lo), fabs(hi)) : 0;
        if (fabs(hi - lo) < epsrel * min_abs) return b;
    }
    return -1;
}

}  // namespace

__global__ __launch_bounds__(256) void k_fic_analytical(const double* __restrict__ SC, int N,
                                                        const double* __restrict__ G,
                                                        const double* __restrict__ wEE,
                                                        const double* __restrict__ wEI,
                                                        double* __restrict__ wIE_out,
                                                        int* __restrict__ unstable) {
#pragma clang fp contract(off)
    const int sim = blockIdx.x;
    const double gJ = G[sim] * C64::J_NMDA;
    double* out = wIE_out + (size_t)sim * N;
    bool bad = false;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const double II = brent_inh_root(wEI[(size_t)sim * N + j]);
        if (II == -1) { bad = true; continue; }
        const double SI = phi_I(II) * C64::tau_I_s * C64::gamma_I_s;
        double dot = 0;
        const double* row = SC + (size_t)j * N;
        for (int k = 0; k < N; ++k) {
            double kee = row[k] * gJ;
            if (k == j) kee += wEE[(size_t)sim * N + j];
            dot += kee * C64::S_E_ss;
        }
        const double J = (-1 / SI) * (C64::I_E_ss - 0.0 - C64::wE_I0 - dot);
        if (J < 0) bad = true;
        out[j] = J;
    }
    const int any = __syncthreads_or(bad);
    if (any)
        for (int j = threadIdx.x; j < N; j += blockDim.x) out[j] = 1.0;
    if (threadIdx.x == 0 && unstable) unstable[sim] = any;
}

hipError_t launch_fic_analytical(const double* SC, int N, int B, const double* G,
                                 const double* wEE, const double* wEI, double* wIE_out,
                                 int* unstable, hipStream_t st) {
    hipLaunchKernelGGL(k_fic_analytical, dim3(B), dim3(256), 0, st, SC, N, G, wEE, wEI, wIE_out,
                       unstable);
    return hipGetLastError();
}

}  // namespace bnm
