// sim_kernels.hip -- integration kernels: rWW substeps, Balloon-Windkessel,
// BOLD sampling at TR, numerical FIC (bnm.cu:72-315 / bnm.cpp:143-290 of the
// reference), one simulation per workgroup.
//
//  k_sim_plain<R>  one lane per node, coupling dot read from SC in global
//                  memory (row-major, lane j walks row j), S_E of the previous
//                  substep in LDS, two barriers per substep: the reference's
//                  GPU design, kept as the "plain" denominator.
//  k_sim_split<R,KC,LS> SC row slices register-resident; the coupling dot
//                  of a node split over LS lanes of a wave and KS wave groups
//                  (see the kernel's comment); one barrier per substep when
//                  KS == 1.
#include <algorithm>
#include <type_traits>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "model.hpp"

namespace bnm {

namespace {

// buffer offset of lanes without a node: beyond every SC resource (at most
// (1024 + kSCtPad) * 1024 * 8 bytes), so their loads are dropped
constexpr int kIdleLaneOffset = 0x40000000;

template <class R> struct Vec4;
template <> struct Vec4<float> { using T = float4; };
template <> struct Vec4<double> { using T = double2; };

// Per-lane state of one node.
template <class R>
struct Lane {
    R SE = R(0.001), SI = R(0.001), IE = 0, II = 0, rE = 0, rI = 0;
    BW<R> bw;
    R wEE = 0, wEI = 0, wIE = 0, delta = 0;
    double mIE = 0;
    double ext[6] = {0, 0, 0, 0, 0, 0};
    __device__ void reset() {
        SE = R(0.001); SI = R(0.001);
        bw.reset();
        mIE = 0;
        for (int i = 0; i < 6; ++i) ext[i] = 0;
    }
};

// Coupling gain argument of Node<R>::step: FP64 keeps c*G*J_NMDA in the
// reference order, FP32 folds G*J_NMDA.
template <class R> __device__ __forceinline__ R gain(double G);
template <> __device__ __forceinline__ double gain<double>(double G) { return G; }
template <> __device__ __forceinline__ float gain<float>(double G) { return (float)(G * C64::J_NMDA); }

template <class R>
__device__ __forceinline__ void load_params(const SimArgs& a, int sim, int j, Lane<R>& L) {
    const size_t o = (size_t)sim * a.N + j;
    L.wEE = (R)a.wEE[o];
    L.wEI = (R)a.wEI[o];
    L.wIE = (R)a.wIE_in[o];
    L.delta = (R)a.init_delta;
}

// BOLD sample + extended-output sums at ts % TR == 0 (bnm.cpp:212-226).
template <class R>
__device__ __forceinline__ void sample(const SimArgs& a, int sim, int j, int vol, Lane<R>& L) {
    R* bold = reinterpret_cast<R*>(a.bold);
    bold[((size_t)sim * a.output_ts + vol) * a.N + j] = L.bw.bold(a.bw_k1, a.bw_k2, a.bw_k3);
    if (vol >= a.n_vols_remove) {
        L.ext[0] += (double)L.SE;
        L.ext[1] += (double)L.IE;
        L.ext[2] += (double)L.rE;
        L.ext[3] += (double)L.SI;
        L.ext[4] += (double)L.II;
        L.ext[5] += (double)L.rI;
    }
}

template <class R>
__device__ __forceinline__ void finish(const SimArgs& a, int sim, int j, bool owner, int vol,
                                       int trial, bool failed, const Lane<R>& L) {
    if (owner) {
        const size_t o = (size_t)sim * a.N + j;
        a.wIE_out[o] = (double)L.wIE;
        const int n = vol - a.n_vols_remove;
        for (int i = 0; i < 6; ++i)
            a.ext[((size_t)sim * 6 + i) * a.N + j] = n > 0 ? L.ext[i] / n : 0.0;
    }
    if (threadIdx.x == 0) {
        a.fic_failed[sim] = failed ? 1 : 0;
        a.fic_trials[sim] = trial;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// plain: one lane per node, SC from global memory (bnm.cu:177-217 design)
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(1024) void k_sim_plain(const R* __restrict__ SC, SimArgs a) {
#pragma clang fp contract(off)
    const int sim = blockIdx.x;
    const int j = threadIdx.x;
    const int N = a.N;
    const bool act = j < N;
    extern __shared__ unsigned char smem_raw[];
    R* S_prev = reinterpret_cast<R*>(smem_raw);

    Lane<R> L;
    if (act) load_params(a, sim, j, L);
    const R gp = gain<R>(a.G[sim]);
    if (act) S_prev[j] = R(0.001);
    int ts = 0, vol = 0, trial = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const float* noise = a.noise;
    while (ts <= a.time_steps) {
        const long long ms = ts % a.noise_period;
        for (int s = 0; s < 10; ++s) {
            __syncthreads();
            if (act) {
                R c = 0;
                const R* row = SC + (size_t)j * N;
                for (int k = 0; k < N; ++k) c += row[k] * S_prev[k];
                const size_t nb = (((size_t)ms * 10 + s) * N + j) * 2;
                Node<R>::step(c, gp, L.wEE, L.wEI, L.wIE, (R)noise[nb], (R)noise[nb + 1], L.SE,
                              L.SI, L.IE, L.II, L.rE, L.rI);
            }
            __syncthreads();
            if (act) S_prev[j] = L.SE;
        }
        if (act) L.bw.step(L.SE);
        if (ts % a.bold_tr == 0) {
            if (act) sample(a, sim, j, vol, L);
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (act && ts >= a.i_start && ts <= a.i_end) L.mIE += (double)L.IE;
            if (ts == a.i_end) {
                const bool out = act && fic_decide<R>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        if (act) { L.reset(); S_prev[j] = R(0.001); }
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
    finish(a, sim, j, act, vol, trial, failed, L);
}

// ---------------------------------------------------------------------------
// split: register-resident SC slices.  The coupling dot of node j is split
// into LS*KS k-slices of KC values: LS slices live in LS lanes of one wave
// (summed by v_permlane32/16_swap, no LDS, no barrier) and KS slices in KS
// wave groups (summed through LDS, one extra barrier).  S_E is double-buffered
// in LDS, slices padded (stride SS) so the LS concurrent 16-byte broadcast
// reads of a wave fall in distinct banks.
// ---------------------------------------------------------------------------
template <class R> __device__ __forceinline__ R lane_sum32(R v);
template <class R> __device__ __forceinline__ R lane_sum16(R v);
template <> __device__ __forceinline__ float lane_sum32<float>(float v) {
    auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
template <> __device__ __forceinline__ float lane_sum16<float>(float v) {
    auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ double swap_f64(double v, bool s32) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    auto rl = s32 ? __builtin_amdgcn_permlane32_swap(lo, lo, false, false)
                  : __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    auto rh = s32 ? __builtin_amdgcn_permlane32_swap(hi, hi, false, false)
                  : __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    // element 0: the lower-row lane's value, element 1: the upper-row lane's
    const double a = __longlong_as_double(((unsigned long long)rh[0] << 32) | rl[0]);
    const double b = __longlong_as_double(((unsigned long long)rh[1] << 32) | rl[1]);
    return a + b;
}
template <> __device__ __forceinline__ double lane_sum32<double>(double v) { return swap_f64(v, true); }
template <> __device__ __forceinline__ double lane_sum16<double>(double v) { return swap_f64(v, false); }
// both values of a permlane32 swap pair in every lane: the lower lane's
// (lanes 0-31) and the upper lane's (32-63)
__device__ __forceinline__ void pair32(float v, float& lower, float& upper) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    lower = __uint_as_float(r[0]);
    upper = __uint_as_float(r[1]);
}
__device__ __forceinline__ void pair32(double v, double& lower, double& upper) {
    const unsigned long long u = __double_as_longlong(v);
    const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
    const auto rl = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto rh = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    lower = __longlong_as_double(((unsigned long long)rh[0] << 32) | rl[0]);
    upper = __longlong_as_double(((unsigned long long)rh[1] << 32) | rl[1]);
}

template <class R, int KC, int LS, int MAXT>
__global__ __launch_bounds__(MAXT) void k_sim_split(const R* __restrict__ SC, SimArgs a, int NW,
                                                     int KS, int SS) {
    using V = typename Vec4<R>::T;
    constexpr int VW = sizeof(V) / sizeof(R);
    constexpr int GN = 64 / LS;  // nodes per wave
    static_assert(KC % 4 == 0, "slice length must be a multiple of 4");
    const int sim = blockIdx.x;
    const int N = a.N;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = wave % NW, ks = wave / NW;
    const int ln = lane % GN, sl = lane / GN;
    const int j = nw * GN + ln;
    const int slice = ks * LS + sl;
    const int k0 = slice * KC;
    const bool comp = ks == 0 && j < N;  // computes the node update
    const bool writer = comp && sl == 0; // publishes S_E, BOLD, outputs
    const int ldS = KS * LS * SS;
    const int posj = (j / KC) * SS + (j % KC);  // LDS slot of node j
    extern __shared__ unsigned char smem_raw[];
    R* Sb = reinterpret_cast<R*>(smem_raw);   // [2][ldS]
    R* part = Sb + 2 * ldS;                   // [KS][NW*GN]

    R sc[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
        const int k = k0 + q;
        sc[q] = (j < N && k < N) ? SC[(size_t)j * N + k] : R(0);
    }
    for (int i = threadIdx.x; i < 2 * ldS; i += blockDim.x) {
        const int slot = i % ldS, qq = slot % SS, k = (slot / SS) * KC + qq;
        Sb[i] = (qq < KC && k < N) ? R(0.001) : R(0);
    }

    Lane<R> L;
    if (comp) load_params(a, sim, j, L);
    const R gp = gain<R>(a.G[sim]);
    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const int P = (int)a.noise_period;
    const float2* noise2 = reinterpret_cast<const float2*>(a.noise) + j;
    float2 nz[10], nzn[10];
    int ms = 0;
    if (comp) {
#pragma unroll
        for (int s = 0; s < 10; ++s) nz[s] = noise2[(size_t)s * N];
    }
    __syncthreads();
    while (ts <= a.time_steps) {
        const int ms_next = ms + 1 == P ? 0 : ms + 1;
        if (comp) {  // prefetch the next ms of noise behind this ms's work
            const float2* p = noise2 + (size_t)ms_next * 10 * N;
#pragma unroll
            for (int s = 0; s < 10; ++s) nzn[s] = p[(size_t)s * N];
        }
#pragma unroll
        for (int s = 0; s < 10; ++s) {
            const V* Sr = reinterpret_cast<const V*>(Sb + buf * ldS + slice * SS);
            R acc[4] = {0, 0, 0, 0};
            V sv[KC / VW];  // all slice reads in flight before the FMAs
#pragma unroll
            for (int q = 0; q < KC / VW; ++q) sv[q] = Sr[q];
#pragma unroll
            for (int q = 0; q < KC / VW; ++q) {
                const V v = sv[q];
                if constexpr (VW == 4) {
                    acc[0] = fma(sc[4 * q + 0], v.x, acc[0]);
                    acc[1] = fma(sc[4 * q + 1], v.y, acc[1]);
                    acc[2] = fma(sc[4 * q + 2], v.z, acc[2]);
                    acc[3] = fma(sc[4 * q + 3], v.w, acc[3]);
                } else {
                    acc[(2 * q) & 3] = fma(sc[2 * q + 0], v.x, acc[(2 * q) & 3]);
                    acc[(2 * q + 1) & 3] = fma(sc[2 * q + 1], v.y, acc[(2 * q + 1) & 3]);
                }
            }
            R c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if constexpr (LS == 4) c = lane_sum16<R>(c);
            if constexpr (LS >= 2) c = lane_sum32<R>(c);
            if (KS > 1) {
                if (ks > 0 && sl == 0) part[ks * NW * GN + j] = c;
                __syncthreads();
                if (ks == 0)
                    for (int t = 1; t < KS; ++t) c += part[t * NW * GN + j];
            }
            if (comp) {
                Node<R>::step(c, gp, L.wEE, L.wEI, L.wIE, (R)nz[s].x, (R)nz[s].y, L.SE, L.SI, L.IE,
                              L.II, L.rE, L.rI);
                if (writer) Sb[(buf ^ 1) * ldS + posj] = L.SE;
            }
            __syncthreads();
            buf ^= 1;
        }
        if (comp) L.bw.step(L.SE);
        if (ts % a.bold_tr == 0) {
            if (writer) sample(a, sim, j, vol, L);
            ++vol;
        }
        ++ts;
        bool restarted = false;
        if (adjusting) {
            if (comp && ts >= a.i_start && ts <= a.i_end) L.mIE += (double)L.IE;
            if (ts == a.i_end) {
                // every lane of a node holds the same state and decides alike
                const bool out = comp && fic_decide<R>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        if (comp) L.reset();
                        if (writer) Sb[buf * ldS + posj] = R(0.001);
                        ts = 0; vol = 0; ++trial;
                        restarted = true;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
        if (restarted) {
            ms = 0;
            if (comp) {
#pragma unroll
                for (int s = 0; s < 10; ++s) nz[s] = noise2[(size_t)s * N];
            }
        } else {
            ms = ms_next;
#pragma unroll
            for (int s = 0; s < 10; ++s) nz[s] = nzn[s];
        }
    }
    finish(a, sim, j, writer, vol, trial, failed, L);
}

// ---------------------------------------------------------------------------
// split_ei (FP32): the split kernel with the node update divided between the
// lanes of a node by population.  After the coupling reduction every lane of
// node j holds c_j; lane slice 0 advances the excitatory pool (S_E, I_E, r_E,
// BW, FIC) and slice 1 the inhibitory pool (S_I, I_I, r_I), running the same
// instruction stream on lane-specific constants:
//   I  = k0 + kSE*S_E + kc*c - kSI*S_I      E: wE*I0, wEE, G*J_NMDA, wIE
//                                           I: wI*I0, wEI, 0,        1
//   r  = (aI - b) / (1 - exp(-d (aI - b)))  (a, b, d of the pool)
//   dS = xi*sigma*sqrt(dt) + (g*r - g*r*S*m - S*dt/tau)   m = 1 (E), 0 (I)
// Each lane swaps S with its partner (v_permlane16/32_swap) before the
// update, so both see the previous substep's S_E and S_I.  A wave issues
// one pool's update instead of both, which is what bounds the split kernel
// (two waves per SIMD at N = 100).  LS = 2 or 4, KS = 1; slices 2 and 3 of
// LS = 4 mirror slices 0 and 1.
//
// BC (LS = 4, KC <= 32): S_E reaches the coupling dot by DPP row broadcast
// instead of LDS vector reads.  With LS = 4 a slice is one 16-lane DPP row;
// lane q of the row reads only S_E[slice base + q] (and + 16) from LDS, and
// v_fmac_f32_dpp row_newbcast:q multiplies lane q's value into every lane's
// accumulator.  Every lane of a slice otherwise reads the whole slice, so
// the LDS returns N^2 * 4 bytes per substep (40 KB at N = 100) whatever the
// slicing; the broadcast form returns 2 * 4 bytes per lane.
// ---------------------------------------------------------------------------

// acc[i] += (lane 4G + i of this DPP row's src) * sc[i], i = 0..3.  The
// leading s_nop 1 provides the two wait states a DPP read needs after a VALU
// write of its source register.
template <int G>
__device__ __forceinline__ void bcast_fma4(float (&acc)[4], float src, float s0, float s1, float s2, float s3) {
#define BNM_BC4(Q0, Q1, Q2, Q3)                                                                  \
    asm("s_nop 1\n\t"                                                                            \
        "v_fmac_f32_dpp %0, %4, %5 row_newbcast:" #Q0 " row_mask:0xf bank_mask:0xf\n\t"           \
        "v_fmac_f32_dpp %1, %4, %6 row_newbcast:" #Q1 " row_mask:0xf bank_mask:0xf\n\t"           \
        "v_fmac_f32_dpp %2, %4, %7 row_newbcast:" #Q2 " row_mask:0xf bank_mask:0xf\n\t"           \
        "v_fmac_f32_dpp %3, %4, %8 row_newbcast:" #Q3 " row_mask:0xf bank_mask:0xf"                 \
        : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3])                                  \
        : "v"(src), "v"(s0), "v"(s1), "v"(s2), "v"(s3))
    if constexpr (G == 0) BNM_BC4(0, 1, 2, 3);
    else if constexpr (G == 1) BNM_BC4(4, 5, 6, 7);
    else if constexpr (G == 2) BNM_BC4(8, 9, 10, 11);
    else BNM_BC4(12, 13, 14, 15);
#undef BNM_BC4
}

// sum over q < NQ of sc[Q0 + q] * (lane q of src's row), into 4 accumulators
template <int NQ, int Q0, int KC>
__device__ __forceinline__ void bcast_dot(float (&acc)[4], float src, const float (&sc)[KC]) {
    static_assert(NQ % 4 == 0 && NQ <= 16, "groups of 4 lanes of one row");
    if constexpr (NQ >= 4) bcast_fma4<0>(acc, src, sc[Q0], sc[Q0 + 1], sc[Q0 + 2], sc[Q0 + 3]);
    if constexpr (NQ >= 8) bcast_fma4<1>(acc, src, sc[Q0 + 4], sc[Q0 + 5], sc[Q0 + 6], sc[Q0 + 7]);
    if constexpr (NQ >= 12) bcast_fma4<2>(acc, src, sc[Q0 + 8], sc[Q0 + 9], sc[Q0 + 10], sc[Q0 + 11]);
    if constexpr (NQ >= 16) bcast_fma4<3>(acc, src, sc[Q0 + 12], sc[Q0 + 13], sc[Q0 + 14], sc[Q0 + 15]);
}

// a + byte B of d in one VALU instruction (SDWA source select), so that the
// packed steps stay packed: unpacked, the compiler keeps one register per
// slot and runs out of them.
template <int B>
__device__ __forceinline__ unsigned add_step(unsigned a, unsigned d) {
    unsigned r;
    if constexpr (B == 0)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_0"
            : "=v"(r) : "v"(a), "v"(d));
    else if constexpr (B == 1)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_1"
            : "=v"(r) : "v"(a), "v"(d));
    else if constexpr (B == 2)
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_2"
            : "=v"(r) : "v"(a), "v"(d));
    else
        asm("v_add_u32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:DWORD src1_sel:BYTE_3"
            : "=v"(r) : "v"(a), "v"(d));
    return r;
}

// 32-bit LDS byte address of a pointer into shared memory, and a 4-byte LDS
// read at such an address (one ds_read_b32 on the address register itself)
__device__ __forceinline__ unsigned lds_byte_addr(const void* p) {
    return (unsigned)(size_t)(const __attribute__((address_space(3))) void*)p;
}
__device__ __forceinline__ float lds_f32(unsigned addr) {
    return *(const __attribute__((address_space(3))) float*)(size_t)addr;
}
template <class R>
__device__ __forceinline__ R lds_load(unsigned addr) {
    return *(const __attribute__((address_space(3))) R*)(size_t)addr;
}

// FORM: 0 the update as written above; 1 the same with the DPP row broadcast;
// 2 / 3 the short-chain update: everything that does not depend on the
// coupling c is formed before it arrives, so that after the reduction
//   x = aa*kc*c + (aa*pre - bb),  y = ndl*aa*kc*c + ndl*(aa*pre - bb)
//   r = rate(x, exp2(y)),         S = clamp(fma(r, g - gm*S, S + xi*nzs - it*S))
// (eleven instructions instead of seventeen; the same expressions up to FP32
// rounding), I_E is formed only on the last substep of a millisecond (its
// only use), and S_E / S_I come straight from the two halves of the lane
// swap.  FORM 2 keeps four partial sums in the dot, FORM 3 one chain.
// Precision bits on top of the form (FORM & 3):
//   64 strict (short-chain forms): every coefficient that multiplies a state
//      value (wEE / wEI, wIE, a*G*J_NMDA, w*I_0, dt*gamma, dt/tau,
//      sigma*sqrt_dt, -d*log2(e)) is an FP32 hi + lo pair formed from its FP64
//      value; the exponent is taken from the rounded x (y = -d*log2(e)*x,
//      where the plain short chain forms it separately from rounded
//      products); S is kept as an unevaluated sum hi + lo (the increment is
//      added to lo and renormalised with Fast2Sum, starting from the
//      reference's exact 0.001), and wIE / delta of the FIC stay in FP64.  A
//      rounded coefficient shifts the fixed point by its relative rounding
//      (~3e-8) times the loop gain, a systematic error that does not average
//      out; the splits remove it at a few FMAs per substep.  What is left is
//      the FP32 rounding of x, exp2 and the reciprocal in the rate.
//   Diagnostics (N = 100 build only), measuring single error terms:
//   8  Balloon-Windkessel step and BOLD in FP64 (libm pow)
//   16 node update (currents, rate, S) in FP64 with only the coupling dot in
//      FP32 (implies 8)
template <int KC, int LS, int MAXT, int FORM>
__global__ __launch_bounds__(MAXT) void k_sim_ei(const float* __restrict__ SC, SimArgs a, int NW, int SS) {
    constexpr int F = FORM & 3;
    constexpr bool STRICT = (FORM & 64) != 0, BW64 = (FORM & 8) != 0, UPD64 = (FORM & 16) != 0;
    constexpr bool BC = F == 1, LEAN = F >= 2;
    static_assert(!UPD64 || LEAN, "FP64 update in the short-chain forms only");
    static_assert(!STRICT || LEAN, "strict coefficients in the short-chain forms only");
    static_assert(LS == 2 || LS == 4, "E/I split needs two lanes per node");
    static_assert(KC % 4 == 0, "slice length must be a multiple of 4");
    static_assert(!BC || (LS == 4 && KC <= 32), "row broadcast: 16-lane slices of at most 32 values");
    constexpr int GN = 64 / LS;
    const int sim = blockIdx.x;
    const int N = a.N;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ln = lane % GN, sl = lane / GN;
    const int j = wave * GN + ln;
    const int pop = sl & 1;               // 0: excitatory pool, 1: inhibitory
    const int k0i = sl * KC;
    const bool comp = j < N;
    const bool writer = comp && sl == 0;  // publishes S_E, BOLD, wIE
    const bool owner = comp && sl < 2;    // writes its pool's means
    const int ldS = LS * SS;
    const int posj = (j / KC) * SS + (j % KC);
    extern __shared__ unsigned char smem_raw[];
    float* Sb = reinterpret_cast<float*>(smem_raw);   // [2][ldS]

    float sc[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
        const int k = k0i + q;
        sc[q] = (j < N && k < N) ? SC[(size_t)j * N + k] : 0.0f;
    }
    for (int i = threadIdx.x; i < 2 * ldS; i += blockDim.x) {
        const int slot = i % ldS, qq = slot % SS, k = (slot / SS) * KC + qq;
        Sb[i] = (qq < KC && k < N) ? 0.001f : 0.0f;
    }
    // pool constants
    const size_t o = (size_t)sim * N + (comp ? j : 0);
    const float GJ = gain<float>(a.G[sim]);
    const float k0 = pop ? (float)C64::wI_I0 : (float)C64::wE_I0;
    const float kSE = comp ? (float)(pop ? a.wEI[o] : a.wEE[o]) : 0.0f;
    const float kc = pop ? 0.0f : GJ;
    const float aa = pop ? (float)C64::a_I : (float)C64::a_E;
    const float bb = pop ? (float)C64::b_I : (float)C64::b_E;
    const float ndl = pop ? (float)(-C64::d_I * kLog2e) : (float)(-C64::d_E * kLog2e);
    const float invd = pop ? (float)(1.0 / C64::d_I) : (float)(1.0 / C64::d_E);
    const double ndl64_ = pop ? -C64::d_I * kLog2e : -C64::d_E * kLog2e;
    const float ndlh = (float)ndl64_, ndll = (float)(ndl64_ - (double)(float)ndl64_);
    const float d12 = pop ? (float)(C64::d_I / 12.0) : (float)(C64::d_E / 12.0);
    const float g = pop ? (float)(C64::dt * C64::gamma_I) : (float)(C64::dt * C64::gamma_E);
    const float gm = pop ? 0.0f : g;
    const float it = pop ? (float)(C64::dt / C64::tau_I) : (float)(C64::dt / C64::tau_E);
    constexpr float nzs = (float)(C64::sigma * C64::sqrt_dt);
    // short-chain form: products of the pool constants, formed in FP64
    const double aa64 = pop ? C64::a_I : C64::a_E, bb64 = pop ? C64::b_I : C64::b_E;
    const double ndl64 = pop ? -C64::d_I * kLog2e : -C64::d_E * kLog2e;
    const double kc64 = pop ? 0.0 : (double)GJ;
    const float aakc = (float)(aa64 * kc64), ndlaakc = (float)(ndl64 * aa64 * kc64);
    const float ndlaa = (float)(ndl64 * aa64), ndlbb = (float)(ndl64 * bb64);
    // strict form: lo parts of the coefficients (value - (double)hi)
    const double kSE64 = comp ? (pop ? a.wEI[o] : a.wEE[o]) : 0.0;
    const float kSEl = (float)(kSE64 - (double)kSE);
    const float k0l = (float)((pop ? C64::wI_I0 : C64::wE_I0) - (double)k0);
    const double aakc64 = aa64 * (pop ? 0.0 : a.G[sim] * C64::J_NMDA);
    const float aakch = (float)aakc64, aakcl = (float)(aakc64 - (double)(float)aakc64);
    const double g64 = pop ? C64::dt * C64::gamma_I : C64::dt * C64::gamma_E;
    const double it64 = pop ? C64::dt / C64::tau_I : C64::dt / C64::tau_E;
    const float gl = (float)(g64 - (double)g), gml = pop ? 0.0f : gl;
    const float itl = (float)(it64 - (double)it);
    constexpr double nzs64 = C64::sigma * C64::sqrt_dt;
    constexpr float nzsl = (float)(nzs64 - (double)(float)nzs64);
    float wIEh = comp ? (float)a.wIE_in[o] : 0.0f;
    float wIEl = comp ? (float)(a.wIE_in[o] - (double)wIEh) : 0.0f;
    float wIE = comp ? (float)a.wIE_in[o] : 0.0f, delta = (float)a.init_delta;
    float S = 0.001f, I = 0.0f, r = 0.0f;
    // low part of S (strict): the reference starts from S = 0.001 exactly,
    // float(0.001) is 4.7e-11 above it
    constexpr float kS0l = (float)(0.001 - (double)0.001f);
    float Sl = STRICT ? kS0l : 0.0f;
    double Sd = 0.001, Id = 0.0, rd = 0.0;   // FP64 node state (bit 16)
    double wIEd = comp ? a.wIE_in[o] : 0.0, deltad = a.init_delta;
    const double Gd = pop ? 0.0 : a.G[sim];
    const double k0d = pop ? C64::wI_I0 : C64::wE_I0;
    const double kSEd = comp ? (pop ? a.wEI[o] : a.wEE[o]) : 0.0;
    const double ddd = pop ? C64::d_I : C64::d_E;
    const double gd = pop ? C64::gamma_I : C64::gamma_E, md = pop ? 0.0 : 1.0;
    const double taud = pop ? C64::tau_I : C64::tau_E;
    BW<float> bw;
    BW<double> bw64;
    double mI = 0, ext[3] = {0, 0, 0};

    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const int P = (int)a.noise_period;
    const float* noise1 = a.noise + 2 * (size_t)(comp ? j : 0) + pop;   // this pool's draw
    float nz[10], nzn[10];
    int ms = 0;
#pragma unroll
    for (int s = 0; s < 10; ++s) nz[s] = comp ? noise1[(size_t)s * 2 * N] : 0.0f;
    __syncthreads();
    while (ts <= a.time_steps) {
        const int ms_next = ms + 1 == P ? 0 : ms + 1;
        if (comp) {
            const float* p = noise1 + (size_t)ms_next * 20 * N;
#pragma unroll
            for (int s = 0; s < 10; ++s) nzn[s] = p[(size_t)s * 2 * N];
        }
        const float kSI = pop ? 1.0f : STRICT ? wIEh : wIE;
        const float kSIl = pop ? 0.0f : wIEl;
#pragma unroll
        for (int s = 0; s < 10; ++s) {
            const float4* Sr = reinterpret_cast<const float4*>(Sb + buf * ldS + sl * SS);
            float4 sv[BC ? 1 : KC / 4];
            float v0 = 0.0f, v1 = 0.0f;
            if constexpr (BC) {
                v0 = Sb[buf * ldS + sl * SS + ln];
                if constexpr (KC > 16) v1 = Sb[buf * ldS + sl * SS + 16 + ln];
            } else {
#pragma unroll
                for (int q = 0; q < KC / 4; ++q) sv[q] = Sr[q];
            }
            // previous-substep S of the partner pool, and the part of I that
            // does not wait for the coupling
            float So;
            {
                auto w = LS == 4 ? __builtin_amdgcn_permlane16_swap(__float_as_uint(S), __float_as_uint(S), false, false)
                                 : __builtin_amdgcn_permlane32_swap(__float_as_uint(S), __float_as_uint(S), false, false);
                // element 0: lower-row lane's value, element 1: upper-row lane's
                So = __uint_as_float(pop ? w[0] : w[1]);
            }
            if constexpr (LEAN) {
                // S_E is the lower row's value of the swap pair, S_I the upper's
                // (slices 2 and 3 mirror 0 and 1), in every lane
                const auto w = LS == 4 ? __builtin_amdgcn_permlane16_swap(__float_as_uint(S), __float_as_uint(S), false, false)
                                       : __builtin_amdgcn_permlane32_swap(__float_as_uint(S), __float_as_uint(S), false, false);
                const float pre = fmaf(-kSI, __uint_as_float(w[1]), fmaf(kSE, __uint_as_float(w[0]), k0));
                const float X0 = fmaf(aa, pre, -bb);
                const float Y0 = fmaf(ndlaa, pre, -ndlbb);
                const float Ac = fmaf(-gm, S, g);
                const float SB = fmaf(-it, S, fmaf(nz[s], nzs, S));
                // strict form: the lo parts, all formed before c arrives
                float X0s = 0.0f, Acl = 0.0f, D = 0.0f, prel = 0.0f;
                if constexpr (STRICT) {
                    prel = fmaf(-kSIl, __uint_as_float(w[1]), fmaf(kSEl, __uint_as_float(w[0]), k0l));
                    X0s = fmaf(aa, prel, fmaf(aa, pre, -bb));
                    Acl = fmaf(-gml, S, gl);
                    D = fmaf(nz[s], nzsl, fmaf(-itl, S, fmaf(-it, S, nz[s] * nzs))) + Sl;
                }
                float c;
                if constexpr (F == 3) {
                    c = sc[0] * sv[0].x;
                    c = fmaf(sc[1], sv[0].y, c);
                    c = fmaf(sc[2], sv[0].z, c);
                    c = fmaf(sc[3], sv[0].w, c);
#pragma unroll
                    for (int q = 1; q < KC / 4; ++q) {
                        c = fmaf(sc[4 * q + 0], sv[q].x, c);
                        c = fmaf(sc[4 * q + 1], sv[q].y, c);
                        c = fmaf(sc[4 * q + 2], sv[q].z, c);
                        c = fmaf(sc[4 * q + 3], sv[q].w, c);
                    }
                } else {
                    float acc[4] = {sc[0] * sv[0].x, sc[1] * sv[0].y, sc[2] * sv[0].z, sc[3] * sv[0].w};
#pragma unroll
                    for (int q = 1; q < KC / 4; ++q) {
                        acc[0] = fmaf(sc[4 * q + 0], sv[q].x, acc[0]);
                        acc[1] = fmaf(sc[4 * q + 1], sv[q].y, acc[1]);
                        acc[2] = fmaf(sc[4 * q + 2], sv[q].z, acc[2]);
                        acc[3] = fmaf(sc[4 * q + 3], sv[q].w, acc[3]);
                    }
                    c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
                }
                if constexpr (LS == 4) c = lane_sum16<float>(c);
                c = lane_sum32<float>(c);
                if constexpr (UPD64) {
                    double SEv, SIv;
                    {
                        const unsigned long long u = __double_as_longlong(Sd);
                        const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
                        auto rl = LS == 4 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                                          : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
                        auto rh = LS == 4 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                                          : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
                        SEv = __longlong_as_double(((unsigned long long)rh[0] << 32) | rl[0]);
                        SIv = __longlong_as_double(((unsigned long long)rh[1] << 32) | rl[1]);
                    }
                    ei_step_f64((double)c, Gd, k0d, kSEd, pop ? 1.0 : wIEd, aa64, bb64, ddd, gd, md, taud,
                                (double)nz[s], SEv, SIv, Sd, Id, rd);
                    S = (float)Sd;
                    r = (float)rd;
                    if (s == 9) I = (float)Id;
                } else if constexpr (STRICT) {
                    const float x = fmaf(aakch, c, fmaf(aakcl, c, X0s));
                    const float y = fmaf(ndlh, x, ndll * x);
                    const float direct = x * __builtin_amdgcn_rcpf(1.0f - __builtin_amdgcn_exp2f(y));
                    const float series = invd + x * (0.5f + d12 * x);
                    r = fabsf(x) < 0.1f ? series : direct;
                    // S + lo + increment, the rounding of the sum carried in lo
                    const float u = fmaf(r, Ac, fmaf(r, Acl, D));
                    const float t = S + u;
                    const float tc = __builtin_amdgcn_fmed3f(t, 0.0f, 1.0f);
                    Sl = tc == t ? u - (t - S) : 0.0f;
                    S = tc;
                    if (s == 9) I = fmaf(kc, c, pre) + prel;
                } else {
                    const float x = fmaf(aakc, c, X0);
                    const float e = __builtin_amdgcn_exp2f(fmaf(ndlaakc, c, Y0));
                    const float direct = x * __builtin_amdgcn_rcpf(1.0f - e);
                    const float series = invd + x * (0.5f + d12 * x);
                    r = fabsf(x) < 0.1f ? series : direct;
                    S = __builtin_amdgcn_fmed3f(fmaf(r, Ac, SB), 0.0f, 1.0f);
                    if (s == 9) I = fmaf(kc, c, pre);
                }
                if (writer) Sb[(buf ^ 1) * ldS + posj] = S;
                __syncthreads();
                buf ^= 1;
                continue;
            }
            const float SEv = pop ? So : S, SIv = pop ? S : So;
            const float pre = k0 + kSE * SEv - kSI * SIv;
            float acc[4] = {0, 0, 0, 0};
            if constexpr (BC) {
                bcast_dot<(KC < 16 ? KC : 16), 0>(acc, v0, sc);
                if constexpr (KC > 16) bcast_dot<KC - 16, 16>(acc, v1, sc);
            } else {
#pragma unroll
                for (int q = 0; q < KC / 4; ++q) {
                    acc[0] = fmaf(sc[4 * q + 0], sv[q].x, acc[0]);
                    acc[1] = fmaf(sc[4 * q + 1], sv[q].y, acc[1]);
                    acc[2] = fmaf(sc[4 * q + 2], sv[q].z, acc[2]);
                    acc[3] = fmaf(sc[4 * q + 3], sv[q].w, acc[3]);
                }
            }
            float c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if constexpr (LS == 4) c = lane_sum16<float>(c);
            c = lane_sum32<float>(c);
            I = pre + kc * c;
            const float x = aa * I - bb;
            const float e = __builtin_amdgcn_exp2f(ndl * x);
            const float direct = x * __builtin_amdgcn_rcpf(1.0f - e);
            const float series = invd + x * (0.5f + d12 * x);
            r = fabsf(x) < 0.1f ? series : direct;
            const float dS = nz[s] * nzs + (g * r - gm * r * S - it * S);
            S = __builtin_amdgcn_fmed3f(S + dS, 0.0f, 1.0f);
            if (writer) Sb[(buf ^ 1) * ldS + posj] = S;
            __syncthreads();
            buf ^= 1;
        }
        if (comp && pop == 0) {
            if constexpr (UPD64) bw64.step(Sd);
            else if constexpr (BW64) bw64.step((double)S + (double)Sl);
            else bw.step(S);
        }
        if (ts % a.bold_tr == 0) {
            if (writer) {
                float* bold = reinterpret_cast<float*>(a.bold);
                bold[((size_t)sim * a.output_ts + vol) * N + j] = BW64 || UPD64 ? (float)bw64.bold(a.bw_k1, a.bw_k2, a.bw_k3) : bw.bold(a.bw_k1, a.bw_k2, a.bw_k3);
            }
            if (vol >= a.n_vols_remove) {
                ext[0] += UPD64 ? Sd : (double)S + (double)Sl;
                ext[1] += UPD64 ? Id : (double)I;
                ext[2] += UPD64 ? rd : (double)r;
            }
            ++vol;
        }
        ++ts;
        bool restarted = false;
        if (adjusting) {
            if (ts >= a.i_start && ts <= a.i_end) mI += UPD64 ? Id : (double)I;
            if (ts == a.i_end) {
                // the excitatory lanes of a node decide alike; inhibitory lanes never do
                bool out = false;
                if (comp && pop == 0) {
                    if constexpr (UPD64) {
                        out = fic_decide<double>(mI, dur, trial, a.max_fic_trials, wIEd, deltad);
                    } else if constexpr (STRICT) {
                        out = fic_decide<double>(mI, dur, trial, a.max_fic_trials, wIEd, deltad);
                        wIEh = (float)wIEd;
                        wIEl = (float)(wIEd - (double)wIEh);
                    } else {
                        out = fic_decide<float>(mI, dur, trial, a.max_fic_trials, wIE, delta);
                    }
                }
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        S = 0.001f;
                        Sl = STRICT ? kS0l : 0.0f;
                        Sd = 0.001;
                        bw.reset();
                        bw64.reset();
                        mI = 0;
                        ext[0] = ext[1] = ext[2] = 0;
                        if (writer) Sb[buf * ldS + posj] = 0.001f;
                        ts = 0; vol = 0; ++trial;
                        restarted = true;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
        if (restarted) {
            ms = 0;
#pragma unroll
            for (int s = 0; s < 10; ++s) nz[s] = comp ? noise1[(size_t)s * 2 * N] : 0.0f;
        } else {
            ms = ms_next;
#pragma unroll
            for (int s = 0; s < 10; ++s) nz[s] = nzn[s];
        }
    }
    if (owner) {
        if (pop == 0) a.wIE_out[o] = UPD64 || STRICT ? wIEd : (double)wIE;
        const int n = vol - a.n_vols_remove;
        for (int i = 0; i < 3; ++i)
            a.ext[((size_t)sim * 6 + 3 * pop + i) * N + j] = n > 0 ? ext[i] / n : 0.0;
    }
    if (threadIdx.x == 0) {
        a.fic_failed[sim] = failed ? 1 : 0;
        a.fic_trials[sim] = trial;
    }
}

// ---------------------------------------------------------------------------
// split_ei (FP64): the E/I lane division of the FP32 kernel above with the
// reference's expression order per pool (bnm.cpp:181-200, contraction off):
//   I  = ((k0 + kSE*S_E) + (c*Gl)*J_NMDA) - kSI*S_I   E: wE_I0, wEE, G, wIE
//                                                     I: wI_I0, wEI, 0, 1
//   x  = a*I - b,  r = x / (1 - exp(-d*x))
//   dS = (xi*sigma)*sqrt_dt + dt*(((1 - m*S)*gamma)*r - S/tau)   m = 1 (E), 0 (I)
// For the inhibitory lanes the extra terms are exact: (c*0)*J = +0 (c >= 0),
// 1*S_I = S_I and (1 - 0*S)*gamma = gamma, so each pool gets the reference's
// own expression.  The coupling dot and its reduction are those of the FP64
// split kernel with the same LS, so the results are bit-identical to it; each
// lane evaluates one FP64 exp per substep instead of two.
// ---------------------------------------------------------------------------

template <int KC, int LS, int MAXT>
__global__ __launch_bounds__(MAXT) void k_sim_ei64(const double* __restrict__ SC, SimArgs a, int NW, int SS) {
    static_assert(LS == 2 || LS == 4, "E/I split needs two lanes per node");
    static_assert(KC % 4 == 0, "slice length must be a multiple of 4");
    constexpr int GN = 64 / LS;
    const int sim = blockIdx.x;
    const int N = a.N;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int ln = lane % GN, sl = lane / GN;
    const int j = wave * GN + ln;
    const int pop = sl & 1;               // 0: excitatory pool, 1: inhibitory
    const int k0i = sl * KC;
    const bool comp = j < N;
    const bool writer = comp && sl == 0;  // publishes S_E, BOLD, wIE
    const bool owner = comp && sl < 2;    // writes its pool's means
    const int ldS = LS * SS;
    const int posj = (j / KC) * SS + (j % KC);
    extern __shared__ unsigned char smem_raw[];
    double* Sb = reinterpret_cast<double*>(smem_raw);   // [2][ldS]

    double sc[KC];
#pragma unroll
    for (int q = 0; q < KC; ++q) {
        const int k = k0i + q;
        sc[q] = (j < N && k < N) ? SC[(size_t)j * N + k] : 0.0;
    }
    for (int i = threadIdx.x; i < 2 * ldS; i += blockDim.x) {
        const int slot = i % ldS, qq = slot % SS, k = (slot / SS) * KC + qq;
        Sb[i] = (qq < KC && k < N) ? 0.001 : 0.0;
    }
    const size_t o = (size_t)sim * N + (comp ? j : 0);
    const double Gl = pop ? 0.0 : a.G[sim];
    const double k0 = pop ? C64::wI_I0 : C64::wE_I0;
    const double kSE = comp ? (pop ? a.wEI[o] : a.wEE[o]) : 0.0;
    const double aa = pop ? C64::a_I : C64::a_E, bb = pop ? C64::b_I : C64::b_E, dd = pop ? C64::d_I : C64::d_E;
    const double g = pop ? C64::gamma_I : C64::gamma_E, m = pop ? 0.0 : 1.0, tau = pop ? C64::tau_I : C64::tau_E;
    double wIE = comp ? a.wIE_in[o] : 0.0, delta = a.init_delta;
    double S = 0.001, I = 0.0, r = 0.0;
    BW<double> bw;
    double mI = 0, ext[3] = {0, 0, 0};

    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const int P = (int)a.noise_period;
    const float* noise1 = a.noise + 2 * (size_t)(comp ? j : 0) + pop;   // this pool's draw
    float nz[10], nzn[10];
    int ms = 0;
#pragma unroll
    for (int s = 0; s < 10; ++s) nz[s] = comp ? noise1[(size_t)s * 2 * N] : 0.0f;
    __syncthreads();
    while (ts <= a.time_steps) {
        const int ms_next = ms + 1 == P ? 0 : ms + 1;
        if (comp) {
            const float* p = noise1 + (size_t)ms_next * 20 * N;
#pragma unroll
            for (int s = 0; s < 10; ++s) nzn[s] = p[(size_t)s * 2 * N];
        }
        const double kSI = pop ? 1.0 : wIE;
#pragma unroll
        for (int s = 0; s < 10; ++s) {
            const double2* Sr = reinterpret_cast<const double2*>(Sb + buf * ldS + sl * SS);
            // previous-substep S of both pools: the lower row holds S_E, the
            // upper S_I (slices 2 and 3 mirror 0 and 1)
            double SEv, SIv;
            {
                const unsigned long long u = __double_as_longlong(S);
                const unsigned lo = (unsigned)u, hi = (unsigned)(u >> 32);
                auto rl = LS == 4 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                                  : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
                auto rh = LS == 4 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                                  : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
                SEv = __longlong_as_double(((unsigned long long)rh[0] << 32) | rl[0]);
                SIv = __longlong_as_double(((unsigned long long)rh[1] << 32) | rl[1]);
            }
            double acc[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q = 0; q < KC / 2; ++q) {
                // reads issued by the scheduler as registers allow (all 26 of
                // KC = 52 in flight at once would not fit beside the SC values)
                const double2 v = Sr[q];
                acc[(2 * q) & 3] = fma(sc[2 * q + 0], v.x, acc[(2 * q) & 3]);
                acc[(2 * q + 1) & 3] = fma(sc[2 * q + 1], v.y, acc[(2 * q + 1) & 3]);
            }
            double c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if constexpr (LS == 4) c = lane_sum16<double>(c);
            c = lane_sum32<double>(c);
            ei_step_f64(c, Gl, k0, kSE, kSI, aa, bb, dd, g, m, tau, (double)nz[s], SEv, SIv, S, I, r);
            if (writer) Sb[(buf ^ 1) * ldS + posj] = S;
            __syncthreads();
            buf ^= 1;
        }
        if (comp && pop == 0) bw.step(S);
        if (ts % a.bold_tr == 0) {
            if (writer) {
                double* bold = reinterpret_cast<double*>(a.bold);
                bold[((size_t)sim * a.output_ts + vol) * N + j] = bw.bold(a.bw_k1, a.bw_k2, a.bw_k3);
            }
            if (vol >= a.n_vols_remove) {
                ext[0] += S;
                ext[1] += I;
                ext[2] += r;
            }
            ++vol;
        }
        ++ts;
        bool restarted = false;
        if (adjusting) {
            if (ts >= a.i_start && ts <= a.i_end) mI += I;
            if (ts == a.i_end) {
                // the excitatory lanes of a node decide alike; inhibitory lanes never do
                const bool out = comp && pop == 0 && fic_decide<double>(mI, dur, trial, a.max_fic_trials, wIE, delta);
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        S = 0.001;
                        bw.reset();
                        mI = 0;
                        ext[0] = ext[1] = ext[2] = 0;
                        if (writer) Sb[buf * ldS + posj] = 0.001;
                        ts = 0; vol = 0; ++trial;
                        restarted = true;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
        if (restarted) {
            ms = 0;
#pragma unroll
            for (int s = 0; s < 10; ++s) nz[s] = comp ? noise1[(size_t)s * 2 * N] : 0.0f;
        } else {
            ms = ms_next;
#pragma unroll
            for (int s = 0; s < 10; ++s) nz[s] = nzn[s];
        }
    }
    if (owner) {
        if (pop == 0) a.wIE_out[o] = wIE;
        const int n = vol - a.n_vols_remove;
        for (int i = 0; i < 3; ++i)
            a.ext[((size_t)sim * 6 + 3 * pop + i) * N + j] = n > 0 ? ext[i] / n : 0.0;
    }
    if (threadIdx.x == 0) {
        a.fic_failed[sim] = failed ? 1 : 0;
        a.fic_trials[sim] = trial;
    }
}

// ---------------------------------------------------------------------------
// hyb64 (FP64, N > 128): the hybrid kernel's residency scheme for the FP64
// path.  Lane (j, ks) owns SC[j][k0 .. k0+kc); the first RR values stay in
// VGPRs, the next LL in LDS ([kk/2][thread] double2 slots), the rest is
// streamed from L2, 8 values (64 bytes, four 16-byte buffer loads) per chunk,
// double-buffered, from a chunked copy SCc64[((k/8)*N + j)*8 + k%8].  S_E of
// the previous substep is read from LDS by 16-byte broadcast loads; the node
// update is the FP64 split kernel's (reference order, Node<double>).
// ---------------------------------------------------------------------------
template <int RR, int MAXT>
__global__ __launch_bounds__(MAXT) void k_sim_hyb64(const double* __restrict__ SCc, SimArgs a, int NW, int KS,
                                                    int kc, int LL) {
    constexpr int CH = 8;
    static_assert(RR % CH == 0, "register part must be whole chunks");
    const int sim = blockIdx.x;
    const int N = a.N;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int nw = wave % NW, ks = wave / NW;
    const int j = nw * 64 + lane;
    const bool comp = ks == 0 && j < N;
    const int k0 = ks * kc;
    const int ldS = KS * kc;
    const int jc = j < N ? j : N - 1;
    extern __shared__ unsigned char smem_raw[];
    double* Sb = reinterpret_cast<double*>(smem_raw);   // [2][ldS]
    double* part = Sb + 2 * ldS;                          // [KS][NW*64]
    double2* SCl = reinterpret_cast<double2*>(part + KS * NW * 64);   // [LL/2][T]
    for (int i = tid; i < 2 * ldS; i += T) Sb[i] = (i % ldS) < N ? 0.001 : 0.0;
    // chunks k/8 >= N/8 hold zeros (padding); the resource's size makes any
    // read past the array return 0
    const __amdgpu_buffer_rsrc_t scr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<double*>(SCc), (short)0, (int)((N + (int)kSCtPad) * N * sizeof(double)), 0x00020000);
    // lanes past the last node take an offset beyond the resource: their
    // loads return 0 without a memory request (the range check covers the
    // lane offset, not the chunk base in soffset)
    const int voff = j < N ? jc * CH * (int)sizeof(double) : kIdleLaneOffset;
    auto chunk = [&](int kb, double* w) {     // the lane's 8 values of chunk kb
        const int so = kb * N * CH * (int)sizeof(double);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(scr, voff + 16 * q, so, 0);
            w[2 * q + 0] = __longlong_as_double(((long long)v[1] << 32) | v[0]);
            w[2 * q + 1] = __longlong_as_double(((long long)v[3] << 32) | v[2]);
        }
    };
    double wr[RR > 0 ? RR : 1];
#pragma unroll
    for (int c = 0; c < RR / CH; ++c) chunk(k0 / CH + c, wr + c * CH);
    for (int c = 0; c < LL / CH; ++c) {
        double w[CH];
        chunk((k0 + RR) / CH + c, w);
#pragma unroll
        for (int q = 0; q < CH / 2; ++q) SCl[(size_t)(c * (CH / 2) + q) * T + tid] = double2{w[2 * q], w[2 * q + 1]};
    }
    const int gb = (k0 + RR + LL) / CH;   // first streamed chunk
    const int nch = (kc - RR - LL) / CH;  // streamed chunks

    Lane<double> L;
    if (comp) load_params(a, sim, j, L);
    const double gp = gain<double>(a.G[sim]);
    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const float2* noise2 = reinterpret_cast<const float2*>(a.noise);
    __syncthreads();
    while (ts <= a.time_steps) {
        const long long ms = ts % a.noise_period;
        const float2* nzp = noise2 + (size_t)ms * 10 * N + j;
#pragma unroll 1
        for (int s = 0; s < 10; ++s) {
            const float2 nz = comp ? nzp[(size_t)s * N] : float2{0.0f, 0.0f};
            double w0[CH], w1[CH];
            if (nch > 0) chunk(gb, w0);
            const double2* Sr = reinterpret_cast<const double2*>(Sb + buf * ldS + k0);
            double acc[4] = {0, 0, 0, 0};
#pragma unroll
            for (int q2 = 0; q2 < RR / 2; ++q2) {
                const double2 v = Sr[q2];
                acc[(2 * q2) & 3] = fma(wr[2 * q2 + 0], v.x, acc[(2 * q2) & 3]);
                acc[(2 * q2 + 1) & 3] = fma(wr[2 * q2 + 1], v.y, acc[(2 * q2 + 1) & 3]);
            }
            for (int q2 = 0; q2 < LL / 2; ++q2) {
                const double2 v = Sr[RR / 2 + q2];
                const double2 w = SCl[(size_t)q2 * T + tid];
                acc[(2 * q2) & 3] = fma(w.x, v.x, acc[(2 * q2) & 3]);
                acc[(2 * q2 + 1) & 3] = fma(w.y, v.y, acc[(2 * q2 + 1) & 3]);
            }
            const double2* Sg = Sr + (RR + LL) / 2;
            for (int c = 0; c < nch; c += 2) {
                if (c + 1 < nch) chunk(gb + c + 1, w1);
#pragma unroll
                for (int q2 = 0; q2 < CH / 2; ++q2) {
                    const double2 v = Sg[c * (CH / 2) + q2];
                    acc[(2 * q2) & 3] = fma(w0[2 * q2 + 0], v.x, acc[(2 * q2) & 3]);
                    acc[(2 * q2 + 1) & 3] = fma(w0[2 * q2 + 1], v.y, acc[(2 * q2 + 1) & 3]);
                }
                if (c + 2 < nch) chunk(gb + c + 2, w0);
                if (c + 1 < nch) {
#pragma unroll
                    for (int q2 = 0; q2 < CH / 2; ++q2) {
                        const double2 v = Sg[(c + 1) * (CH / 2) + q2];
                        acc[(2 * q2) & 3] = fma(w1[2 * q2 + 0], v.x, acc[(2 * q2) & 3]);
                        acc[(2 * q2 + 1) & 3] = fma(w1[2 * q2 + 1], v.y, acc[(2 * q2 + 1) & 3]);
                    }
                }
            }
            double c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if (KS > 1) {
                if (ks > 0) part[ks * NW * 64 + j] = c;
                __syncthreads();
                if (ks == 0)
                    for (int t = 1; t < KS; ++t) c += part[t * NW * 64 + j];
            }
            if (comp) {
                Node<double>::step(c, gp, L.wEE, L.wEI, L.wIE, (double)nz.x, (double)nz.y, L.SE, L.SI, L.IE,
                                   L.II, L.rE, L.rI);
                Sb[(buf ^ 1) * ldS + j] = L.SE;
            }
            __syncthreads();
            buf ^= 1;
        }
        if (comp) L.bw.step(L.SE);
        if (ts % a.bold_tr == 0) {
            if (comp) sample(a, sim, j, vol, L);
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (comp && ts >= a.i_start && ts <= a.i_end) L.mIE += L.IE;
            if (ts == a.i_end) {
                const bool out = comp && fic_decide<double>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        if (comp) { L.reset(); Sb[buf * ldS + j] = 0.001; }
                        ts = 0; vol = 0; ++trial;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
    }
    finish(a, sim, j, comp, vol, trial, failed, L);
}

// ---------------------------------------------------------------------------
// split_g: the split kernel for node counts whose SC does not fit in
// registers (N > 256).  Lane j streams its SC column slice from a transposed
// copy SCt[k][j] in global memory (L2-resident; consecutive lanes read
// consecutive nodes, one 256-byte line pair per wave and k), 16 k at a time
// with the loads issued ahead of the FMAs.  LS = 1; KS wave groups split k.
// ---------------------------------------------------------------------------
template <class R>
__global__ __launch_bounds__(1024) void k_sim_split_g(const R* __restrict__ SCt, SimArgs a, int NW,
                                                      int KS, int kc) {
    using V = typename Vec4<R>::T;
    constexpr int VW = sizeof(V) / sizeof(R);
    constexpr int CH = 16;  // k per chunk
    const int sim = blockIdx.x;
    const int N = a.N;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nw = wave % NW, ks = wave / NW;
    const int j = nw * 64 + lane;
    const bool comp = ks == 0 && j < N;
    const int k0 = ks * kc;
    const int ldS = KS * kc;  // kc multiple of 16
    const int jc = j < N ? j : N - 1;  // clamped column for the streamed loads
    extern __shared__ unsigned char smem_raw[];
    R* Sb = reinterpret_cast<R*>(smem_raw);
    R* part = Sb + 2 * ldS;
    for (int i = threadIdx.x; i < 2 * ldS; i += blockDim.x) Sb[i] = (i % ldS) < N ? R(0.001) : R(0);

    Lane<R> L;
    if (comp) load_params(a, sim, j, L);
    const R gp = gain<R>(a.G[sim]);
    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const float2* noise2 = reinterpret_cast<const float2*>(a.noise);
    const int nch = kc / CH;
    __syncthreads();
    while (ts <= a.time_steps) {
        const long long ms = ts % a.noise_period;
        const float2* nzp = noise2 + (size_t)ms * 10 * N + j;
#pragma unroll 1
        for (int s = 0; s < 10; ++s) {
            // this substep's noise, loaded ahead of the streamed dot
            const float2 nz = comp ? nzp[(size_t)s * N] : float2{0.0f, 0.0f};
            const R* Sr = Sb + buf * ldS + k0;
            R acc[4] = {0, 0, 0, 0};
            for (int c = 0; c < nch; ++c) {
                const int kb = k0 + c * CH;
                R w[CH];
#pragma unroll
                for (int q = 0; q < CH; ++q) {
                    const int k = kb + q;
                    w[q] = k < N ? SCt[(size_t)k * N + jc] : R(0);
                }
                const V* Sv = reinterpret_cast<const V*>(Sr + c * CH);
#pragma unroll
                for (int q = 0; q < CH / VW; ++q) {
                    const V v = Sv[q];
                    if constexpr (VW == 4) {
                        acc[0] = fma(w[4 * q + 0], v.x, acc[0]);
                        acc[1] = fma(w[4 * q + 1], v.y, acc[1]);
                        acc[2] = fma(w[4 * q + 2], v.z, acc[2]);
                        acc[3] = fma(w[4 * q + 3], v.w, acc[3]);
                    } else {
                        acc[(2 * q) & 3] = fma(w[2 * q + 0], v.x, acc[(2 * q) & 3]);
                        acc[(2 * q + 1) & 3] = fma(w[2 * q + 1], v.y, acc[(2 * q + 1) & 3]);
                    }
                }
            }
            R c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if (KS > 1) {
                if (ks > 0) part[ks * NW * 64 + j] = c;
                __syncthreads();
                if (ks == 0)
                    for (int t = 1; t < KS; ++t) c += part[t * NW * 64 + j];
            }
            if (comp) {
                Node<R>::step(c, gp, L.wEE, L.wEI, L.wIE, (R)nz.x, (R)nz.y, L.SE, L.SI, L.IE,
                              L.II, L.rE, L.rI);
                Sb[(buf ^ 1) * ldS + j] = L.SE;
            }
            __syncthreads();
            buf ^= 1;
        }
        if (comp) L.bw.step(L.SE);
        if (ts % a.bold_tr == 0) {
            if (comp) sample(a, sim, j, vol, L);
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (comp && ts >= a.i_start && ts <= a.i_end) L.mIE += (double)L.IE;
            if (ts == a.i_end) {
                const bool out = comp && fic_decide<R>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        if (comp) { L.reset(); Sb[buf * ldS + j] = R(0.001); }
                        ts = 0; vol = 0; ++trial;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
    }
    finish(a, sim, j, comp, vol, trial, failed, L);
}

// ---------------------------------------------------------------------------
// hyb (FP32, N > 128): the streaming kernel with most of SC resident.  Lane
// (j, ks) owns SC[j][k0 .. k0+kc); the first RR values stay in VGPRs, the
// next LL in LDS (float4 per lane per 4 k, laid out [kk/4][thread] so a
// wave reads 1 KB contiguous), and only the rest is streamed from L2 each
// substep, 16 k per chunk, double-buffered, the first chunk issued before
// the resident FMAs.  SC comes in a chunked layout SCc[k/16][j][k%16], so a
// lane's 16 values of a chunk are 64 contiguous bytes (four 16-byte buffer
// loads; a wave reads 4 KB contiguous), addressed by one VGPR column offset
// and one SGPR chunk offset.  Per-CU L1 bandwidth (64 B/clk) is what bounds
// the plain streaming kernel at N = 400 (640 KB of SC per substep);
// residency cuts the streamed part to what LDS and registers cannot hold.
// ---------------------------------------------------------------------------
//
// BC: S_E reaches the dot by DPP row broadcast.  Every lane of a wave needs
// the same S_E[k]; for each 16-k chunk lane q of every 16-lane row reads
// S_E[k + q] (one 4-byte LDS read, 256 bytes per wave instead of 4 KB of
// 16-byte broadcast reads) and v_fmac_f32_dpp row_newbcast:q hands it to the
// whole row.  Same accumulation order as the LDS form.
//
// FORM 0: S_E by LDS broadcast; 1: BC; 2: BC with three streamed chunks in
// flight, the first three issued before the register-resident products and
// the LDS-resident chunks interleaved between streamed ones, so that each
// load has about three chunks of work to hide its L2 latency behind (with
// one chunk ahead, each of the ~12 streamed chunks at N = 400 waits out most
// of it).  FORM 2 accumulates in another order (LDS chunks between streamed
// ones).
template <int RR, int MAXT, int FORM>
__global__ __launch_bounds__(MAXT) void k_sim_hyb(const float* __restrict__ SCc, SimArgs a, int NW, int KS,
                                                  int kc, int LL) {
    constexpr bool BC = FORM >= 1;
    constexpr int CH = 16;
    static_assert(RR % CH == 0, "register part must be whole chunks");
    const int sim = blockIdx.x;
    const int N = a.N;
    const int T = blockDim.x;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int nw = wave % NW, ks = wave / NW;
    const int j = nw * 64 + lane;
    const bool comp = ks == 0 && j < N;
    const int k0 = ks * kc;
    const int ldS = KS * kc;
    const int jc = j < N ? j : N - 1;
    extern __shared__ unsigned char smem_raw[];
    float* Sb = reinterpret_cast<float*>(smem_raw);   // [2][ldS]
    float* part = Sb + 2 * ldS;                        // [KS][NW*64]
    float4* SCl = reinterpret_cast<float4*>(part + KS * NW * 64);   // [LL/4][T]
    for (int i = tid; i < 2 * ldS; i += T) Sb[i] = (i % ldS) < N ? 0.001f : 0.0f;
    // chunks k/16 >= N/16 hold zeros (padding); the resource's size makes any
    // read past the array return 0
    const __amdgpu_buffer_rsrc_t scr = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float*>(SCc), (short)0, (int)((N + (int)kSCtPad) * N * sizeof(float)), 0x00020000);
    const int voff = j < N ? jc * CH * (int)sizeof(float) : kIdleLaneOffset;   // as in k_sim_hyb64
    auto chunk = [&](int kb, float* w) {     // the lane's 16 values of chunk kb
        const int so = kb * N * CH * (int)sizeof(float);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const auto v = __builtin_amdgcn_raw_buffer_load_b128(scr, voff + 16 * q, so, 0);
            w[4 * q + 0] = __uint_as_float(v[0]);
            w[4 * q + 1] = __uint_as_float(v[1]);
            w[4 * q + 2] = __uint_as_float(v[2]);
            w[4 * q + 3] = __uint_as_float(v[3]);
        }
    };
    float wr[RR > 0 ? RR : 1];
#pragma unroll
    for (int c = 0; c < RR / CH; ++c) chunk(k0 / CH + c, wr + c * CH);
    for (int c = 0; c < LL / CH; ++c) {
        float w[CH];
        chunk((k0 + RR) / CH + c, w);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            SCl[(size_t)(c * 4 + q) * T + tid] = float4{w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]};
    }
    const int gb = (k0 + RR + LL) / CH;   // first streamed chunk
    const int nch = (kc - RR - LL) / CH;  // streamed chunks

    Lane<float> L;
    if (comp) load_params(a, sim, j, L);
    const float gp = gain<float>(a.G[sim]);
    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const float2* noise2 = reinterpret_cast<const float2*>(a.noise);
    __syncthreads();
    while (ts <= a.time_steps) {
        const long long ms = ts % a.noise_period;
        const float2* nzp = noise2 + (size_t)ms * 10 * N + j;
#pragma unroll 1
        for (int s = 0; s < 10; ++s) {
            const float2 nz = comp ? nzp[(size_t)s * N] : float2{0.0f, 0.0f};
            float w0[CH], w1[CH], w2[FORM == 2 ? CH : 1];
            if (nch > 0) chunk(gb, w0);
            if constexpr (FORM == 2) {
                if (nch > 1) chunk(gb + 1, w1);
                if (nch > 2) chunk(gb + 2, w2);
            }
            const float4* Sr = reinterpret_cast<const float4*>(Sb + buf * ldS + k0);
            const float* Sq = Sb + buf * ldS + k0 + (lane & 15);   // BC: this lane's S_E of each chunk
            float acc[4] = {0, 0, 0, 0};
            // the 16 products of one chunk: w[0..15] times S_E[kq .. kq+15]
            auto dot16 = [&](const float* w, int kq) {
                if constexpr (BC) {
                    const float src = Sq[kq];
                    bcast_fma4<0>(acc, src, w[0], w[1], w[2], w[3]);
                    bcast_fma4<1>(acc, src, w[4], w[5], w[6], w[7]);
                    bcast_fma4<2>(acc, src, w[8], w[9], w[10], w[11]);
                    bcast_fma4<3>(acc, src, w[12], w[13], w[14], w[15]);
                } else {
#pragma unroll
                    for (int q4 = 0; q4 < CH / 4; ++q4) {
                        const float4 v = Sr[kq / 4 + q4];
                        acc[0] = fmaf(w[4 * q4 + 0], v.x, acc[0]);
                        acc[1] = fmaf(w[4 * q4 + 1], v.y, acc[1]);
                        acc[2] = fmaf(w[4 * q4 + 2], v.z, acc[2]);
                        acc[3] = fmaf(w[4 * q4 + 3], v.w, acc[3]);
                    }
                }
            };
            // LDS-resident chunk c
            auto lds16 = [&](int c) {
                float w[CH];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 v = SCl[(size_t)(c * 4 + q) * T + tid];
                    w[4 * q + 0] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
                }
                dot16(w, RR + c * CH);
            };
            if constexpr (FORM == 2) {
#pragma unroll
                for (int c = 0; c < RR / CH; ++c) dot16(wr + c * CH, c * CH);
                const int kg = RR + LL, nl = LL / CH;
                int l = 0;
                for (int c = 0; c < nch; c += 3) {
                    dot16(w0, kg + c * CH);
                    if (c + 3 < nch) chunk(gb + c + 3, w0);
                    if (l < nl) lds16(l++);
                    if (c + 1 < nch) {
                        dot16(w1, kg + (c + 1) * CH);
                        if (c + 4 < nch) chunk(gb + c + 4, w1);
                        if (l < nl) lds16(l++);
                    }
                    if (c + 2 < nch) {
                        dot16(w2, kg + (c + 2) * CH);
                        if (c + 5 < nch) chunk(gb + c + 5, w2);
                        if (l < nl) lds16(l++);
                    }
                }
                while (l < nl) lds16(l++);
            } else if constexpr (BC) {
#pragma unroll
                for (int c = 0; c < RR / CH; ++c) dot16(wr + c * CH, c * CH);
                for (int c = 0; c < LL / CH; ++c) {
                    float w[CH];
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const float4 v = SCl[(size_t)(c * 4 + q) * T + tid];
                        w[4 * q + 0] = v.x; w[4 * q + 1] = v.y; w[4 * q + 2] = v.z; w[4 * q + 3] = v.w;
                    }
                    dot16(w, RR + c * CH);
                }
            } else {
#pragma unroll
                for (int q4 = 0; q4 < RR / 4; ++q4) {
                    const float4 v = Sr[q4];
                    acc[0] = fmaf(wr[4 * q4 + 0], v.x, acc[0]);
                    acc[1] = fmaf(wr[4 * q4 + 1], v.y, acc[1]);
                    acc[2] = fmaf(wr[4 * q4 + 2], v.z, acc[2]);
                    acc[3] = fmaf(wr[4 * q4 + 3], v.w, acc[3]);
                }
                for (int q4 = 0; q4 < LL / 4; ++q4) {
                    const float4 v = Sr[RR / 4 + q4];
                    const float4 w = SCl[(size_t)q4 * T + tid];
                    acc[0] = fmaf(w.x, v.x, acc[0]);
                    acc[1] = fmaf(w.y, v.y, acc[1]);
                    acc[2] = fmaf(w.z, v.z, acc[2]);
                    acc[3] = fmaf(w.w, v.w, acc[3]);
                }
            }
            if constexpr (FORM != 2) {
                const int kg = RR + LL;
                for (int c = 0; c < nch; c += 2) {
                    if (c + 1 < nch) chunk(gb + c + 1, w1);
                    dot16(w0, kg + c * CH);
                    if (c + 2 < nch) chunk(gb + c + 2, w0);
                    if (c + 1 < nch) dot16(w1, kg + (c + 1) * CH);
                }
            }
            float c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if (KS > 1) {
                if (ks > 0) part[ks * NW * 64 + j] = c;
                __syncthreads();
                if (ks == 0)
                    for (int t = 1; t < KS; ++t) c += part[t * NW * 64 + j];
            }
            if (comp) {
                Node<float>::step(c, gp, L.wEE, L.wEI, L.wIE, nz.x, nz.y, L.SE, L.SI, L.IE, L.II, L.rE, L.rI);
                Sb[(buf ^ 1) * ldS + j] = L.SE;
            }
            __syncthreads();
            buf ^= 1;
        }
        if (comp) L.bw.step(L.SE);
        if (ts % a.bold_tr == 0) {
            if (comp) sample(a, sim, j, vol, L);
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (comp && ts >= a.i_start && ts <= a.i_end) L.mIE += (double)L.IE;
            if (ts == a.i_end) {
                const bool out = comp && fic_decide<float>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        if (comp) { L.reset(); Sb[buf * ldS + j] = 0.001f; }
                        ts = 0; vol = 0; ++trial;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
    }
    finish(a, sim, j, comp, vol, trial, failed, L);
}

// ---------------------------------------------------------------------------
// sparse (N <= 512): the SC rows as lists of their nonzeros, for connectomes
// of partial density (the synthetic 20-40 % SCs).  LS = 1 or 2 lanes per node
// (LS = 2: each lane takes half of the row's k-range and the two partial dots
// are summed with a permlane32 swap), one workgroup of ceil(N * LS / 64)
// waves per simulation, one barrier per substep.  R = float or double (the
// element type of the weights and of S_E in LDS; ES = sizeof(R) bytes).
//
// The nonzeros of a lane's range, in increasing k, become P slots (P the
// same for every lane): slot p holds the weight w_p and the step
// d_p = ES * (k_p - k_{p-1}) in bytes (from the range's first node) as one
// byte, four steps per 32-bit register.  A gap of more than 255 / ES nodes
// takes zero-weight slots of the largest step; a lane with fewer nonzeros
// gets zero-weight slots of step 0 between its own, every item placed near
// slot k * P / (range width) (build_sparse_sc), so that at each slot all
// lanes gather S_E of nearby nodes, in distinct LDS banks.  The lane walks
// its slots with a running LDS byte address into the S_E buffer (one
// integer add per slot, the byte picked by an SDWA select), gathers S_E with
// one LDS read and multiplies it in, eight slots per batch.  The first RW
// weights are register-resident; in FP32 the remaining P - RW (at most LLMAX)
// sit in LDS as [slot / 4][LS * N] float4; every step is in registers.
// Nothing is streamed: at N = 400 and 40 % density (P = 200) the whole
// matrix is on the CU, where the dense FP32 matrix (640 KB) does not fit.
//
// STRICT (FP32): the node update is NodeStrict (hi / lo coefficients, S as
// S + Sl, FIC in FP64, extended-output sums in place in a.ext); otherwise
// Node<R> (FP64: the reference's expression order).  With two lanes per node
// the strict and the FP64 update are split by pool (the lower lane of the
// pair steps the excitatory pool, the upper the inhibitory, both from the
// previous S_E and S_I exchanged by a permlane32 swap; PoolStrict gives
// NodeStrict's bits, ei_step_f64 Node<double>'s), which halves the update's
// VALU work per lane.
//
// A zero weight adds +0 to the dense dot as well (S_E is finite and >= 0),
// so this kernel and the dense ones differ only in the grouping of the sum:
// four partial sums by slot index here, by k in the dense kernels.  Lanes
// past the last node walk node N - 1's slots (the same addresses: LDS
// broadcasts, no extra bank traffic) and keep nothing.
// ---------------------------------------------------------------------------
template <class R, int RW, int LLMAX, int MAXT, int LS, bool STRICT>
__global__ __launch_bounds__(MAXT) void k_sim_sparse(const R* __restrict__ Wsp, const unsigned* __restrict__ Dsp,
                                                    SimArgs a, int P) {
    static_assert(RW % 8 == 0 && LLMAX % 8 == 0, "slots come in groups of eight");
    static_assert(std::is_same<R, float>::value || (!STRICT && LLMAX == 0),
                  "FP64: the reference-order update, all weights in registers");
    constexpr unsigned ES = sizeof(R);
    // two lanes per node split the node update by pool (the lower lane of the
    // pair excitatory, the upper inhibitory), the partner's S from a
    // permlane32 swap: strict FP32 with PoolStrict, FP64 with ei_step_f64
    // (up to 48 register slots: its per-lane pool constants take 20 VGPRs,
    // which the 64-slot build does not have without spills)
    constexpr bool F64 = std::is_same<R, double>::value;
    constexpr bool SPLIT = LS == 2 && (STRICT || (F64 && RW <= 48));
    static_assert(LS == 1 || LS == 2, "one or two lanes per node");
    constexpr int DW = (RW + LLMAX) / 4;
    constexpr int GN = 64 / LS;   // nodes per wave
    const int sim = blockIdx.x;
    const int N = a.N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sl = lane / GN;                      // the lane's slice of the row
    const int j = wave * GN + lane % GN;
    const bool comp = j < N;
    const bool writer = comp && sl == 0;           // publishes S_E and the outputs
    const int jc = comp ? j : N - 1;
    const size_t li = (size_t)sl * N + jc;         // the lane's column in W / D / LDS weights
    const int ldS = (N + 3) & ~3;
    const int KH = (N + LS - 1) / LS;              // nodes per slice
    const int nl4 = P > RW ? (P - RW) / 4 : 0;   // LDS-resident groups of four slots
    const int nl8 = nl4 / 2;                      // ... of eight (P - RW is a multiple of 8)
    extern __shared__ unsigned char smem_raw[];
    R* Sb = reinterpret_cast<R*>(smem_raw);                                   // [2][ldS]
    float4* Wl = reinterpret_cast<float4*>(smem_raw + 2 * ES * (size_t)ldS);  // [nl4][LS * N] (FP32)
    const unsigned sbase = lds_byte_addr(Sb);
    for (int i = threadIdx.x; i < 2 * ldS; i += blockDim.x) Sb[i] = (i % ldS) < N ? R(0.001) : R(0);
    // Wsp / Dsp hold kSparseRW + kSparseLL slots (zero past P), so every
    // build reads them unguarded
    R w[RW];
    unsigned dl[DW];
    const R* wp = Wsp + li;
    const unsigned* dp = Dsp + li;
#pragma unroll
    for (int p = 0; p < RW; ++p) w[p] = wp[(size_t)p * LS * N];
#pragma unroll
    for (int i = 0; i < DW; ++i) dl[i] = dp[(size_t)i * LS * N];
    if constexpr (LLMAX > 0) {
        if (comp)
            for (int c = 0; c < nl4; ++c) {
                const R* s = wp + (size_t)(RW + 4 * c) * LS * N;
                Wl[(size_t)c * LS * N + li] = float4{s[0], s[LS * (size_t)N], s[2 * LS * (size_t)N], s[3 * LS * (size_t)N]};
            }
    }

    Lane<R> L;
    if (comp) load_params(a, sim, j, L);
    // strict update: coefficient hi / lo pairs, S lo parts, FIC in FP64
    NodeStrict ns;
    float SEl = NodeStrict::kS0l, SIl = NodeStrict::kS0l;
    double wIEd = 0.0, deltad = a.init_delta;
    // (the strict form sums its extended outputs in place in a.ext, once per
    // TR sample, which keeps six FP64 sums out of the VGPRs; split, each lane
    // owns its pool's three)
    double* extg = a.ext + (size_t)sim * 6 * N + j;
    const bool ext_owner = SPLIT ? comp : writer;
    const int e0 = SPLIT ? 3 * sl : 0, e1 = SPLIT ? e0 + 3 : 6;
    PoolStrict ps;
    R S_own = R(0.001), r_own = 0, I_own = 0;   // split: this lane's pool (+ Sl_own, strict)
    float Sl_own = NodeStrict::kS0l;
    // FP64 split: the pool's constants of ei_step_f64 (the inhibitory pool
    // has no coupling, kSI = 1 and m = 0)
    const double pGl = sl ? 0.0 : a.G[sim], pk0 = sl ? C64::wI_I0 : C64::wE_I0;
    const double pa = sl ? C64::a_I : C64::a_E, pb = sl ? C64::b_I : C64::b_E, pd = sl ? C64::d_I : C64::d_E;
    const double pg = sl ? C64::gamma_I : C64::gamma_E, pm = sl ? 0.0 : 1.0, ptau = sl ? C64::tau_I : C64::tau_E;
    const double pkSE = sl ? (double)L.wEI : (double)L.wEE;
    if constexpr (STRICT) {
        ns.set_G(a.G[sim]);
        if (comp) {
            const size_t o = (size_t)sim * N + j;
            wIEd = a.wIE_in[o];
            if constexpr (SPLIT) ps.init(sl, a.G[sim], a.wEE[o], a.wEI[o], wIEd);
            else ns.set_w(a.wEE[o], a.wEI[o], wIEd);
        }
        if (ext_owner)
            for (int i = e0; i < e1; ++i) extg[(size_t)i * N] = 0.0;
    }
    const R gp = gain<R>(a.G[sim]);
    int ts = 0, vol = 0, trial = 0, buf = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const float2* noise2 = reinterpret_cast<const float2*>(a.noise);
    // noise one substep ahead: substep s loads s + 1's pair, the tenth the
    // next millisecond's first (reloaded when a FIC restart changes it)
    long long pre_ms = -1;
    float2 nzn{0.0f, 0.0f};
    __syncthreads();
    while (ts <= a.time_steps) {
        const long long ms = ts % a.noise_period;
        const float2* nzp = noise2 + (size_t)ms * 10 * N + j;
        if (pre_ms != ms && comp) nzn = nzp[0];
#pragma unroll 1
        for (int s = 0; s < 10; ++s) {
            const float2 nz = nzn;
            if (s < 9) {
                if (comp) nzn = nzp[(size_t)(s + 1) * N];
            } else {
                pre_ms = (ts + 1) % a.noise_period;
                if (comp) nzn = noise2[(size_t)pre_ms * 10 * N + j];
            }
            unsigned addr = sbase + (unsigned)(buf * ldS + sl * KH) * ES;
            R acc[4] = {0, 0, 0, 0};
            // eight slots: the eight addresses (byte b of the packed steps d0,
            // d1 added in turn), the eight gathers, then the eight products,
            // so that eight LDS reads are in flight
            // (each batch its own scheduling region: hoisting later batches'
            // address chains costs registers the weights need)
            auto slot8 = [&](const R (&wv)[8], unsigned d0, unsigned d1) {
                __builtin_amdgcn_sched_barrier(0);
                unsigned ad[8];
                ad[0] = add_step<0>(addr, d0);
                ad[1] = add_step<1>(ad[0], d0);
                ad[2] = add_step<2>(ad[1], d0);
                ad[3] = add_step<3>(ad[2], d0);
                ad[4] = add_step<0>(ad[3], d1);
                ad[5] = add_step<1>(ad[4], d1);
                ad[6] = add_step<2>(ad[5], d1);
                ad[7] = add_step<3>(ad[6], d1);
                addr = ad[7];
                R sv[8];
#pragma unroll
                for (int q = 0; q < 8; ++q) sv[q] = lds_load<R>(ad[q]);
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    if constexpr (std::is_same<R, float>::value) acc[q & 3] = fmaf(wv[q], sv[q], acc[q & 3]);
                    else acc[q & 3] = fma(wv[q], sv[q], acc[q & 3]);
                }
            };
#pragma unroll
            for (int p = 0; p < RW; p += 8) {
                const R wv[8] = {w[p], w[p + 1], w[p + 2], w[p + 3], w[p + 4], w[p + 5], w[p + 6], w[p + 7]};
                slot8(wv, dl[p >> 2], dl[(p >> 2) + 1]);
            }
#pragma unroll
            for (int c = 0; c < LLMAX / 8; ++c)
                if (c < nl8) {
                    const float4 v0 = Wl[(size_t)(2 * c) * LS * N + li], v1 = Wl[(size_t)(2 * c + 1) * LS * N + li];
                    const R wv[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
                    slot8(wv, dl[RW / 4 + 2 * c], dl[RW / 4 + 2 * c + 1]);
                }
            R c = (acc[0] + acc[1]) + (acc[2] + acc[3]);
            if constexpr (LS == 2) c = lane_sum32<R>(c);   // the same sum in both lanes of a node
            // the dot is complete here, on every lane: left to itself the
            // compiler sinks the products into the node-lane branch below and
            // keeps every gathered value live until then
            asm volatile("" : : "v"(c));
            if constexpr (SPLIT) {
                R SEv, SIv;   // S_E from the lower lane of the pair, S_I from the upper
                pair32(S_own, SEv, SIv);
                if (comp) {
                    if constexpr (STRICT)
                        ps.step(c, sl ? nz.y : nz.x, SEv, SIv, S_own, Sl_own, r_own, I_own);
                    else
                        ei_step_f64(c, pGl, pk0, pkSE, sl ? 1.0 : L.wIE, pa, pb, pd, pg, pm, ptau,
                                    (double)(sl ? nz.y : nz.x), SEv, SIv, S_own, I_own, r_own);
                    if (sl == 0) Sb[(buf ^ 1) * ldS + j] = S_own;
                }
            } else if (comp) {
                if constexpr (STRICT)
                    ns.step(c, nz.x, nz.y, L.SE, SEl, L.SI, SIl, L.rE, L.rI, &L.IE, &L.II);
                else
                    Node<R>::step(c, gp, L.wEE, L.wEI, L.wIE, (R)nz.x, (R)nz.y, L.SE, L.SI, L.IE, L.II, L.rE, L.rI);
                if (sl == 0) Sb[(buf ^ 1) * ldS + j] = L.SE;
            }
            __syncthreads();
            buf ^= 1;
        }
        if (comp) L.bw.step(SPLIT ? S_own : L.SE);
        if (ts % a.bold_tr == 0) {
            if constexpr (SPLIT) {
                if (writer) {
                    R* bold = reinterpret_cast<R*>(a.bold);
                    bold[((size_t)sim * a.output_ts + vol) * N + j] = L.bw.bold(a.bw_k1, a.bw_k2, a.bw_k3);
                }
                if (comp && vol >= a.n_vols_remove) {
                    if constexpr (STRICT) {
                        extg[(size_t)e0 * N] += (double)S_own + (double)Sl_own;
                        extg[(size_t)(e0 + 1) * N] += (double)I_own;
                        extg[(size_t)(e0 + 2) * N] += (double)r_own;
                    } else {
                        L.ext[0] += (double)S_own;
                        L.ext[1] += (double)I_own;
                        L.ext[2] += (double)r_own;
                    }
                }
            } else if (writer) {
                if constexpr (STRICT) {
                    float* bold = reinterpret_cast<float*>(a.bold);
                    bold[((size_t)sim * a.output_ts + vol) * N + j] = L.bw.bold(a.bw_k1, a.bw_k2, a.bw_k3);
                    if (vol >= a.n_vols_remove) {
                        extg[0] += (double)L.SE + (double)SEl;
                        extg[(size_t)N] += (double)L.IE;
                        extg[2 * (size_t)N] += (double)L.rE;
                        extg[3 * (size_t)N] += (double)L.SI + (double)SIl;
                        extg[4 * (size_t)N] += (double)L.II;
                        extg[5 * (size_t)N] += (double)L.rI;
                    }
                } else {
                    sample(a, sim, j, vol, L);
                }
            }
            ++vol;
        }
        ++ts;
        if (adjusting) {
            if (comp && ts >= a.i_start && ts <= a.i_end) L.mIE += (double)(SPLIT ? I_own : L.IE);
            if (ts == a.i_end) {
                bool out = false;
                if (comp) {
                    if constexpr (SPLIT) {
                        // the excitatory lane decides; the inhibitory pool has no wIE
                        if (sl == 0) {
                            if constexpr (STRICT) {
                                out = fic_decide<double>(L.mIE, dur, trial, a.max_fic_trials, wIEd, deltad);
                                ps.set_wIE(wIEd);
                            } else {
                                out = fic_decide<R>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                            }
                        }
                    } else if constexpr (STRICT) {
                        out = fic_decide<double>(L.mIE, dur, trial, a.max_fic_trials, wIEd, deltad);
                        ns.set_wIE(wIEd);
                    } else {
                        out = fic_decide<R>(L.mIE, dur, trial, a.max_fic_trials, L.wIE, L.delta);
                    }
                }
                if (__syncthreads_or(out)) {
                    if (trial < a.max_fic_trials) {
                        if (comp) {
                            L.reset();
                            SEl = SIl = NodeStrict::kS0l;
                            S_own = R(0.001);
                            Sl_own = NodeStrict::kS0l;
                            if (sl == 0) Sb[buf * ldS + j] = R(0.001);
                        }
                        if (STRICT && ext_owner)
                            for (int i = e0; i < e1; ++i) extg[(size_t)i * N] = 0.0;
                        ts = 0; vol = 0; ++trial;
                        __syncthreads();
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
        }
    }
    if constexpr (STRICT) {
        if (writer) a.wIE_out[(size_t)sim * N + j] = wIEd;
        if (ext_owner) {
            const int n = vol - a.n_vols_remove;
            for (int i = e0; i < e1; ++i) extg[(size_t)i * N] = n > 0 ? extg[(size_t)i * N] / n : 0.0;
        }
        if (threadIdx.x == 0) {
            a.fic_failed[sim] = failed ? 1 : 0;
            a.fic_trials[sim] = trial;
        }
    } else if constexpr (SPLIT) {
        // finish() for a lane that holds one pool's three sums
        if (comp) {
            if (sl == 0) a.wIE_out[(size_t)sim * N + j] = (double)L.wIE;
            const int n = vol - a.n_vols_remove;
            for (int i = 0; i < 3; ++i) a.ext[((size_t)sim * 6 + e0 + i) * N + j] = n > 0 ? L.ext[i] / n : 0.0;
        }
        if (threadIdx.x == 0) {
            a.fic_failed[sim] = failed ? 1 : 0;
            a.fic_trials[sim] = trial;
        }
    } else {
        finish(a, sim, j, writer, vol, trial, failed, L);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class R>
hipError_t launch_sim_plain(const R* SC, const SimArgs& a, hipStream_t st) {
    const int threads = ((a.N + 63) / 64) * 64;
    const size_t smem = (size_t)a.N * sizeof(R);
    hipLaunchKernelGGL(k_sim_plain<R>, dim3(a.B), dim3(threads), smem, st, SC, a);
    return hipGetLastError();
}

// Register budget of a lane: KC SC values, KC read-ahead S_E values, plus the
// node state, noise double buffer and FP64 sums (~130 VGPRs in FP32, ~200 in
// FP64); the workgroup size bound follows (4, 2 or 1 waves per SIMD).
constexpr int maxt_for(int KC, int elem) {
    const int regs = 2 * KC * (elem / 4) + (elem == 4 ? 130 : 200);
    return regs <= 128 ? 1024 : regs <= 256 ? 512 : 256;
}

static const int kKC[] = {8, 16, 28, 32, 52, 64, 100, 128};

FastShape fast_shape(int N, int LS, int KS, size_t elem) {
    FastShape s{};
    s.LS = LS;
    s.KS = KS;
    s.NW = (N * LS + 63) / 64;
    const int per = (N + LS * KS - 1) / (LS * KS);
    s.KC = 0;
    for (int kc : kKC)
        if (kc >= per) { s.KC = kc; break; }
    s.kc = s.KC;
    // slice stride: smallest SS >= KC (multiple of 4) whose LS slices start in
    // 4-bank windows that do not overlap (32 banks of 4 bytes)
    int SS = s.KC;
    const int wb = (int)(elem / 4) * 4;  // banks covered by one 16-byte read
    for (;; SS += 4) {
        bool ok = true;
        for (int x = 0; x < LS && ok; ++x)
            for (int y = x + 1; y < LS && ok; ++y) {
                const int bx = (int)((x * SS * elem / 4) % 32), by = (int)((y * SS * elem / 4) % 32);
                const int d = (by - bx + 32) % 32;
                if (d < wb || d > 32 - wb) ok = false;
            }
        if (ok || SS > s.KC + 64) break;
    }
    s.SS = SS;
    s.threads = s.NW * KS * 64;
    s.smem = (2 * (size_t)KS * LS * SS + (size_t)KS * s.NW * (64 / LS)) * elem;
    s.ok = s.KC > 0 && s.threads <= maxt_for(s.KC, (int)elem) && s.smem <= 64 * 1024 &&
           (LS == 1 || LS == 2 || LS == 4);
    return s;
}

template <class R, int KC, int LS>
static void launch_split(const R* SC, const SimArgs& a, const FastShape& s, hipStream_t st) {
    constexpr int MAXT = maxt_for(KC, (int)sizeof(R));
    hipLaunchKernelGGL((k_sim_split<R, KC, LS, MAXT>), dim3(a.B), dim3(s.threads), s.smem, st, SC, a,
                       s.NW, s.KS, s.SS);
}

template <class R, int KC>
static void launch_split_kc(const R* SC, const SimArgs& a, const FastShape& s, hipStream_t st) {
    if (s.LS == 1) launch_split<R, KC, 1>(SC, a, s, st);
    else if (s.LS == 2) launch_split<R, KC, 2>(SC, a, s, st);
    else launch_split<R, KC, 4>(SC, a, s, st);
}

template <class R>
hipError_t launch_sim_fast(const R* SC, const SimArgs& a, int LS, int KS, hipStream_t st) {
    const FastShape s = fast_shape(a.N, LS, KS, sizeof(R));
    if (!s.ok) return hipErrorInvalidValue;
    switch (s.KC) {
        case 8: launch_split_kc<R, 8>(SC, a, s, st); break;
        case 16: launch_split_kc<R, 16>(SC, a, s, st); break;
        case 28: launch_split_kc<R, 28>(SC, a, s, st); break;
        case 32: launch_split_kc<R, 32>(SC, a, s, st); break;
        case 52: launch_split_kc<R, 52>(SC, a, s, st); break;
        case 64: launch_split_kc<R, 64>(SC, a, s, st); break;
        case 100: launch_split_kc<R, 100>(SC, a, s, st); break;
        default: launch_split_kc<R, 128>(SC, a, s, st); break;
    }
    return hipGetLastError();
}

template <int KC, int LS, int FORM>
static void launch_ei(const float* SC, const SimArgs& a, const FastShape& s, hipStream_t st) {
    constexpr int MAXT = maxt_for(KC, 4);
    hipLaunchKernelGGL((k_sim_ei<KC, LS, MAXT, FORM>), dim3(a.B), dim3(s.threads), s.smem, st, SC, a, s.NW, s.SS);
}

template <int KC, int LS>
static hipError_t launch_ei_form(const float* SC, const SimArgs& a, const FastShape& s, int form, hipStream_t st) {
    switch (form) {
        case 0: launch_ei<KC, LS, 0>(SC, a, s, st); return hipSuccess;
        case 2: launch_ei<KC, LS, 2>(SC, a, s, st); return hipSuccess;
        case 3: launch_ei<KC, LS, 3>(SC, a, s, st); return hipSuccess;
        case 3 | 64: launch_ei<KC, LS, 3 | 64>(SC, a, s, st); return hipSuccess;
        default: break;
    }
    // precision diagnostics, N = 100 shape only
    if constexpr (KC == 28 && LS == 4) {
        switch (form) {
            case 3 | 64 | 8: launch_ei<KC, LS, 3 | 64 | 8>(SC, a, s, st); return hipSuccess;
            case 3 | 8 | 16: launch_ei<KC, LS, 3 | 8 | 16>(SC, a, s, st); return hipSuccess;
            case 3 | 8: launch_ei<KC, LS, 3 | 8>(SC, a, s, st); return hipSuccess;
            default: break;
        }
    }
    return hipErrorInvalidValue;
}

template <int KC>
static hipError_t launch_ei_kc(const float* SC, const SimArgs& a, const FastShape& s, int form, hipStream_t st) {
    if constexpr (KC <= 32) {
        if (form == 1) { launch_ei<KC, 4, 1>(SC, a, s, st); return hipSuccess; }
    }
    if (s.LS == 2) return launch_ei_form<KC, 2>(SC, a, s, form, st);
    return launch_ei_form<KC, 4>(SC, a, s, form, st);
}

FastShape ei_shape(int N, int LS, bool bc) {
    FastShape s = fast_shape(N, LS, 1, sizeof(float));
    s.ok = s.ok && (LS == 2 || LS == 4);
    if (bc) {
        // lane q of slice x reads slots x*SS + q and x*SS + 16 + q: SS = 16
        // keeps one-value slices contiguous; two-value slices take SS = 48
        // (in bounds, and the two slices of a half-wave on disjoint banks)
        s.ok = s.ok && LS == 4 && s.KC <= 32;
        s.SS = s.KC <= 16 ? 16 : 48;
        s.smem = (2 * (size_t)LS * s.SS + (size_t)s.NW * (64 / LS)) * sizeof(float);
        s.ok = s.ok && s.smem <= 64 * 1024;
    }
    return s;
}

hipError_t launch_sim_ei(const float* SC, const SimArgs& a, int LS, int form, hipStream_t st) {
    const FastShape s = ei_shape(a.N, LS, form == 1);
    if (!s.ok) return hipErrorInvalidValue;
    hipError_t e;
    switch (s.KC) {
        case 8: e = launch_ei_kc<8>(SC, a, s, form, st); break;
        case 16: e = launch_ei_kc<16>(SC, a, s, form, st); break;
        case 28: e = launch_ei_kc<28>(SC, a, s, form, st); break;
        case 32: e = launch_ei_kc<32>(SC, a, s, form, st); break;
        case 52: e = launch_ei_kc<52>(SC, a, s, form, st); break;
        case 64: e = launch_ei_kc<64>(SC, a, s, form, st); break;
        case 100: e = launch_ei_kc<100>(SC, a, s, form, st); break;
        default: e = launch_ei_kc<128>(SC, a, s, form, st); break;
    }
    return e != hipSuccess ? e : hipGetLastError();
}

// FP64 E/I workgroup bound: the kernel keeps its 28 SC doubles per lane and
// state within 256 VGPRs (two waves per SIMD) without spills up to KC = 28
constexpr int ei64_maxt(int KC) { return KC <= 28 ? 512 : 256; }

template <int KC, int LS>
static void launch_ei64(const double* SC, const SimArgs& a, const FastShape& s, hipStream_t st) {
    constexpr int MAXT = ei64_maxt(KC);
    hipLaunchKernelGGL((k_sim_ei64<KC, LS, MAXT>), dim3(a.B), dim3(s.threads), s.smem, st, SC, a, s.NW, s.SS);
}

FastShape ei64_shape(int N, int LS) {
    FastShape s = fast_shape(N, LS, 1, sizeof(double));
    s.ok = s.KC > 0 && s.threads <= ei64_maxt(s.KC) && s.smem <= 64 * 1024 && (LS == 2 || LS == 4);
    return s;
}

hipError_t launch_sim_ei64(const double* SC, const SimArgs& a, int LS, hipStream_t st) {
    const FastShape s = ei64_shape(a.N, LS);
    if (!s.ok) return hipErrorInvalidValue;
    auto go = [&](auto kc) {
        constexpr int KC = decltype(kc)::value;
        if (s.LS == 2) launch_ei64<KC, 2>(SC, a, s, st);
        else launch_ei64<KC, 4>(SC, a, s, st);
    };
    switch (s.KC) {
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 16: go(std::integral_constant<int, 16>{}); break;
        case 28: go(std::integral_constant<int, 28>{}); break;
        case 32: go(std::integral_constant<int, 32>{}); break;
        case 52: go(std::integral_constant<int, 52>{}); break;
        case 64: go(std::integral_constant<int, 64>{}); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

FastShape stream_shape(int N, int KS, size_t elem) {
    FastShape s{};
    s.LS = 1;
    s.KS = KS;
    s.NW = (N + 63) / 64;
    const int per = (N + KS - 1) / KS;
    s.kc = ((per + 15) / 16) * 16;
    s.KC = 0;
    s.SS = s.kc;
    s.threads = s.NW * KS * 64;
    s.smem = (2 * (size_t)KS * s.kc + (size_t)KS * s.NW * 64) * elem;
    s.ok = s.threads <= 1024 && s.smem <= 64 * 1024;
    return s;
}

template <class R>
hipError_t launch_sim_stream(const R* SCt, const SimArgs& a, int KS, hipStream_t st) {
    const FastShape s = stream_shape(a.N, KS, sizeof(R));
    if (!s.ok) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_sim_split_g<R>, dim3(a.B), dim3(s.threads), s.smem, st, SCt, a, s.NW, KS, s.kc);
    return hipGetLastError();
}

// largest workgroup a register part of RR values per lane compiles for
// without spilling (512 VGPRs per SIMD lane: 256 per wave at two waves)
constexpr int hyb_maxt(int RR) { return RR == 0 ? 1024 : RR <= 128 ? 512 : 256; }

// Hybrid shape: RR register-resident values per lane (with two waves per
// SIMD 128, or 112 for form 2's third chunk buffer; 160 / 144 with one; rr >= 0
// overrides), then as many LDS values as `lds_bytes` holds, rounded so the
// streamed rest is whole chunks.
FastShape hyb_shape(int N, int KS, size_t lds_bytes, int rr, int form) {
    FastShape s = stream_shape(N, KS, sizeof(float));
    // RR: the register part the workgroup size leaves room for (one or two
    // waves per SIMD), capped by the lane's k-range
    s.KC = rr >= 0 ? rr : s.threads <= 256 ? (form == 2 ? 144 : 160) : s.threads <= 512 ? (form == 2 ? 112 : 128) : 0;
    while (s.KC > s.kc) s.KC -= 16;
    const size_t base = (2 * (size_t)KS * s.kc + (size_t)KS * s.NW * 64) * sizeof(float);
    int LL = 0;
    if (lds_bytes > base + 16) LL = (int)((lds_bytes - base) / ((size_t)s.threads * sizeof(float)));
    LL = std::min(LL, s.kc - s.KC) / 16 * 16;   // whole chunks
    s.SS = LL;
    s.smem = base + (size_t)LL * s.threads * sizeof(float);
    s.ok = s.threads <= 1024 && s.smem <= lds_bytes && (s.kc - s.KC - LL) % 16 == 0 && s.KC % 16 == 0 &&
           s.KC >= 0 && s.KC <= 160 && s.threads <= hyb_maxt(s.KC);
    // form 2's third chunk buffer: without spills up to 112 values at two
    // waves per SIMD and 144 at one, not in the 1024-thread (RR = 0) build
    if (form == 2) s.ok = s.ok && s.KC >= 16 && s.KC <= (s.threads <= 256 ? 144 : 112);
    return s;
}

template <int RR>
static void launch_hyb(const float* SCc, const SimArgs& a, const FastShape& s, int form, hipStream_t st) {
    constexpr int MAXT = hyb_maxt(RR);
    if (form == 2)
        hipLaunchKernelGGL((k_sim_hyb<RR, MAXT, 2>), dim3(a.B), dim3(s.threads), s.smem, st, SCc, a, s.NW, s.KS,
                           s.kc, s.SS);
    else if (form == 1)
        hipLaunchKernelGGL((k_sim_hyb<RR, MAXT, 1>), dim3(a.B), dim3(s.threads), s.smem, st, SCc, a, s.NW, s.KS,
                           s.kc, s.SS);
    else
        hipLaunchKernelGGL((k_sim_hyb<RR, MAXT, 0>), dim3(a.B), dim3(s.threads), s.smem, st, SCc, a, s.NW, s.KS,
                           s.kc, s.SS);
}

hipError_t launch_sim_hyb(const float* SCc, const SimArgs& a, int KS, size_t lds_bytes, int rr, int form,
                          hipStream_t st) {
    const FastShape s = hyb_shape(a.N, KS, lds_bytes, rr, form);
    if (!s.ok) return hipErrorInvalidValue;
    switch (s.KC) {
        case 0: launch_hyb<0>(SCc, a, s, form, st); break;
        case 16: launch_hyb<16>(SCc, a, s, form, st); break;
        case 32: launch_hyb<32>(SCc, a, s, form, st); break;
        case 48: launch_hyb<48>(SCc, a, s, form, st); break;
        case 64: launch_hyb<64>(SCc, a, s, form, st); break;
        case 80: launch_hyb<80>(SCc, a, s, form, st); break;
        case 96: launch_hyb<96>(SCc, a, s, form, st); break;
        case 112: launch_hyb<112>(SCc, a, s, form, st); break;
        case 128: launch_hyb<128>(SCc, a, s, form, st); break;
        case 144: launch_hyb<144>(SCc, a, s, form, st); break;
        case 160: launch_hyb<160>(SCc, a, s, form, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Sparse row lists of an N x N row-major FP32 SC (see k_sim_sparse), each
// row cut into LS slices of ceil(N / LS) nodes: returns P (a multiple of 8;
// 0 for a matrix without nonzeros) and, when P fits the kernel (P <=
// kSparseRW + kSparseLL), fills W[p * LS*N + sl*N + j] (weights) and
// D[(p / 4) * LS*N + sl*N + j] (byte p % 4: the step of slot p, in bytes,
// from the slice's first node) for all kSparseRW + kSparseLL slots, zero
// past P.
template <class T>
int build_sparse_sc(const T* SC, int N, int LS, std::vector<T>& W, std::vector<unsigned>& D) {
    constexpr size_t kSlots = kSparseRW + kSparseLL;
    constexpr unsigned ES = sizeof(T), kMaxStep = 255 / ES * ES;   // the largest step one byte holds
    const int KH = (N + LS - 1) / LS;
    const size_t M = (size_t)LS * N;   // lane columns: slice sl of row j at sl * N + j
    struct Item { T w; unsigned step, at; };   // weight, step to it, its byte offset in the slice
    std::vector<std::vector<Item>> rows(M);
    size_t P = 0;
    for (int sl = 0; sl < LS; ++sl)
        for (int j = 0; j < N; ++j) {
            auto& row = rows[(size_t)sl * N + j];
            unsigned at = 0;   // byte offset of the previous slot's S_E from the slice start
            for (int k = sl * KH; k < std::min(N, (sl + 1) * KH); ++k) {
                const T v = SC[(size_t)j * N + k];
                if (v == T(0)) continue;
                const unsigned to = ES * (unsigned)(k - sl * KH);
                while (to - at > kMaxStep) {
                    at += kMaxStep;
                    row.push_back({T(0), kMaxStep, at});
                }
                row.push_back({v, to - at, to});
                at = to;
            }
            P = std::max(P, row.size());
        }
    P = (P + 7) / 8 * 8;
    if (P > kSlots) return (int)P;   // too many for the kernel; nothing filled
    W.assign(kSlots * M, T(0));
    D.assign(kSlots / 4 * M, 0u);
    // Slot placement by column: item i (node k_i of the slice) goes to slot
    // k_i * P / KH, moved later where the previous item took that slot and
    // earlier where the remaining items would not fit (at most P - n + i);
    // slots between items hold zero weights with step 0 (re-reads of the
    // current address).  All lanes then stand near the same node at every
    // slot, within 32 consecutive S_E that lie in distinct LDS banks (or at
    // one address), where items packed from slot 0 drift apart with the rows'
    // densities and collide in banks.
    for (size_t j = 0; j < M; ++j) {
        const size_t n = rows[j].size();
        size_t prev = 0;
        for (size_t i = 0; i < n; ++i) {
            size_t p = (size_t)(rows[j][i].at / ES) * P / (size_t)KH;
            if (i > 0) p = std::max(p, prev + 1);
            p = std::min(p, P - n + i);
            W[p * M + j] = rows[j][i].w;
            D[p / 4 * M + j] |= rows[j][i].step << (8 * (p % 4));
            prev = p;
        }
    }
    return (int)P;
}

template int build_sparse_sc<float>(const float*, int, int, std::vector<float>&, std::vector<unsigned>&);
template int build_sparse_sc<double>(const double*, int, int, std::vector<double>&, std::vector<unsigned>&);

// register part: the slot count rounded up to 16 up to kSparseRW, beyond it
// kSparseRW in registers and the rest (at most kSparseLL) in LDS; FP64
// (elem 8): at most kSparseRW64 slots, all in registers
SparseShape sparse_shape(int N, int P, int LS, size_t lds_bytes, size_t elem) {
    SparseShape s{};
    s.P = P;
    s.LS = LS;
    s.threads = ((N + 64 / LS - 1) / (64 / LS)) * 64;
    s.RW = P <= kSparseRW ? std::max(16, (P + 15) / 16 * 16) : kSparseRW;
    s.LL = P > kSparseRW ? P - kSparseRW : 0;
    const size_t ldS = (size_t)((N + 3) & ~3);
    s.smem = 2 * ldS * elem + (size_t)s.LL * LS * N * sizeof(float);
    if (elem == 8 && (s.LL > 0 || s.RW > kSparseRW64)) {
        s.ok = false;
        return s;
    }
    // (kStaticLds: the workgroup reductions' own LDS beside the dynamic part;
    // two lanes per node only without LDS-resident weights)
    s.ok = N >= 1 && (LS == 1 || (LS == 2 && s.LL == 0)) && s.threads <= 512 && P > 0 && P % 8 == 0 &&
           s.LL <= kSparseLL && s.smem + kStaticLds <= lds_bytes;
    return s;
}

template <int RW, int LLMAX, int LS>
static void launch_sparse(const float* W, const unsigned* D, const SimArgs& a, const SparseShape& s, bool strict,
                          hipStream_t st) {
    if (strict)
        hipLaunchKernelGGL((k_sim_sparse<float, RW, LLMAX, 512, LS, true>), dim3(a.B), dim3(s.threads), s.smem, st, W,
                           D, a, s.P);
    else
        hipLaunchKernelGGL((k_sim_sparse<float, RW, LLMAX, 512, LS, false>), dim3(a.B), dim3(s.threads), s.smem, st,
                           W, D, a, s.P);
}

template <int RW, int LS>
static void launch_sparse64(const double* W, const unsigned* D, const SimArgs& a, const SparseShape& s,
                            hipStream_t st) {
    hipLaunchKernelGGL((k_sim_sparse<double, RW, 0, 512, LS, false>), dim3(a.B), dim3(s.threads), s.smem, st, W, D,
                       a, s.P);
}

hipError_t launch_sim_sparse64(const double* W, const unsigned* D, const SimArgs& a, int P, int LS, size_t lds_bytes,
                               hipStream_t st) {
    const SparseShape s = sparse_shape(a.N, P, LS, lds_bytes, sizeof(double));
    if (!s.ok) return hipErrorInvalidValue;
    static_assert(kSparseRW64 == 64, "instantiations below");
    const bool two = LS == 2;
    switch (s.RW) {
        case 16: two ? launch_sparse64<16, 2>(W, D, a, s, st) : launch_sparse64<16, 1>(W, D, a, s, st); break;
        case 32: two ? launch_sparse64<32, 2>(W, D, a, s, st) : launch_sparse64<32, 1>(W, D, a, s, st); break;
        case 48: two ? launch_sparse64<48, 2>(W, D, a, s, st) : launch_sparse64<48, 1>(W, D, a, s, st); break;
        case 64: two ? launch_sparse64<64, 2>(W, D, a, s, st) : launch_sparse64<64, 1>(W, D, a, s, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int LS>
static hipError_t launch_sparse_ls(const float* W, const unsigned* D, const SimArgs& a, const SparseShape& s,
                                   bool strict, hipStream_t st) {
    static_assert(kSparseRW == 112 && kSparseLL == 96, "instantiations below");
    switch (s.RW) {
        case 16: launch_sparse<16, 0, LS>(W, D, a, s, strict, st); break;
        case 32: launch_sparse<32, 0, LS>(W, D, a, s, strict, st); break;
        case 48: launch_sparse<48, 0, LS>(W, D, a, s, strict, st); break;
        case 64: launch_sparse<64, 0, LS>(W, D, a, s, strict, st); break;
        case 80: launch_sparse<80, 0, LS>(W, D, a, s, strict, st); break;
        case 96: launch_sparse<96, 0, LS>(W, D, a, s, strict, st); break;
        case 112:
            if constexpr (LS == 1) {
                if (s.LL > 0) launch_sparse<112, 96, 1>(W, D, a, s, strict, st);
                else launch_sparse<112, 0, 1>(W, D, a, s, strict, st);
            } else {
                launch_sparse<112, 0, LS>(W, D, a, s, strict, st);
            }
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_sim_sparse(const float* W, const unsigned* D, const SimArgs& a, int P, int LS, bool strict,
                             size_t lds_bytes, hipStream_t st) {
    const SparseShape s = sparse_shape(a.N, P, LS, lds_bytes, sizeof(float));
    if (!s.ok) return hipErrorInvalidValue;
    return LS == 2 ? launch_sparse_ls<2>(W, D, a, s, strict, st) : launch_sparse_ls<1>(W, D, a, s, strict, st);
}

// FP64 hybrid shape: as hyb_shape with 8-byte values in 8-value chunks; the
// register part fits without spills up to 48 values at two waves per SIMD
// and 64 at one
constexpr int hyb64_maxt(int RR) { return RR <= 48 ? 512 : 256; }

FastShape hyb64_shape(int N, int KS, size_t lds_bytes, int rr) {
    FastShape s = stream_shape(N, KS, sizeof(double));
    s.KC = rr >= 0 ? rr : s.threads <= 256 ? 64 : s.threads <= 512 ? 48 : 0;
    while (s.KC > s.kc) s.KC -= 8;
    const size_t base = (2 * (size_t)KS * s.kc + (size_t)KS * s.NW * 64) * sizeof(double);
    int LL = 0;
    if (lds_bytes > base + 16) LL = (int)((lds_bytes - base) / ((size_t)s.threads * sizeof(double)));
    LL = std::min(LL, s.kc - s.KC) / 8 * 8;
    s.SS = LL;
    s.smem = base + (size_t)LL * s.threads * sizeof(double);
    s.ok = s.threads <= 1024 && s.KC > 0 && s.smem <= lds_bytes && (s.kc - s.KC - LL) % 8 == 0 &&
           s.KC % 8 == 0 && s.KC <= 64 && s.threads <= hyb64_maxt(s.KC);
    return s;
}

template <int RR>
static void launch_hyb64(const double* SCc, const SimArgs& a, const FastShape& s, hipStream_t st) {
    constexpr int MAXT = hyb64_maxt(RR);
    hipLaunchKernelGGL((k_sim_hyb64<RR, MAXT>), dim3(a.B), dim3(s.threads), s.smem, st, SCc, a, s.NW, s.KS, s.kc,
                       s.SS);
}

hipError_t launch_sim_hyb64(const double* SCc, const SimArgs& a, int KS, size_t lds_bytes, int rr, hipStream_t st) {
    const FastShape s = hyb64_shape(a.N, KS, lds_bytes, rr);
    if (!s.ok) return hipErrorInvalidValue;
    switch (s.KC) {
        case 8: launch_hyb64<8>(SCc, a, s, st); break;
        case 16: launch_hyb64<16>(SCc, a, s, st); break;
        case 24: launch_hyb64<24>(SCc, a, s, st); break;
        case 32: launch_hyb64<32>(SCc, a, s, st); break;
        case 40: launch_hyb64<40>(SCc, a, s, st); break;
        case 48: launch_hyb64<48>(SCc, a, s, st); break;
        case 56: launch_hyb64<56>(SCc, a, s, st); break;
        case 64: launch_hyb64<64>(SCc, a, s, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template hipError_t launch_sim_stream<double>(const double*, const SimArgs&, int, hipStream_t);
template hipError_t launch_sim_stream<float>(const float*, const SimArgs&, int, hipStream_t);
template hipError_t launch_sim_plain<double>(const double*, const SimArgs&, hipStream_t);
template hipError_t launch_sim_plain<float>(const float*, const SimArgs&, hipStream_t);
template hipError_t launch_sim_fast<double>(const double*, const SimArgs&, int, int, hipStream_t);
template hipError_t launch_sim_fast<float>(const float*, const SimArgs&, int, int, hipStream_t);

}  // namespace bnm
