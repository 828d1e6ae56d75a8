// sim_mfma.hip -- population-batched integration: the coupling of 16
// simulations that share one SC is one MFMA contraction per substep,
//   C[N x 16] = SC[N x N] . S_E[N x 16]
// (bnm.cu:186-189 computes the same dot per node and simulation).
//
// Workgroup = 16 simulations; wave w owns the 16-node row tile w (N <= 128,
// so at most 8 waves).  Lane l of wave w holds the MFMA output rows
// 4*(l/16)..+3 of column l%16, i.e. four (node, simulation) pairs, and keeps
// their state in registers for the whole run.  SC row tiles are loaded once
// into VGPRs as A fragments; S_E of the 16 simulations is double-buffered in
// LDS as [sim][k] rows (padded so a 16-lane group's 16-byte reads hit distinct
// banks) and read as B fragments.  The K axis is permuted so that lane group
// g = l/16 supplies the contiguous k-range [g*KQ, (g+1)*KQ), which lets every
// B fragment come from one vector load.
//
// MODE 0  FP32 operands, v_mfma_f32_16x16x4_f32
// MODE 1  BF16 operands, v_mfma_f32_16x16x32_bf16, FP32 accumulate
// MODE 2  BF16x3: SC = hi + lo and S_E = hi + lo in BF16; hi*hi + lo*hi + hi*lo
//         on three MFMAs (the lo*lo term is below FP32 rounding)
//
// Each simulation keeps its own clock, so FIC restarts (bnm.cu:268-283) and
// per-simulation durations are independent; finished simulations idle.
#include "kernels.hpp"
#include "model.hpp"

namespace bnm {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// float -> bf16 bits with gfx950's v_cvt_pk_bf16_f32: round to nearest even, the same bits as
// adding 0x7fff + lsb to the FP32 pattern for finite inputs
__device__ __forceinline__ unsigned short f2bf(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

__device__ __forceinline__ bf16x8 as_bf16x8(u16x8 v) { return __builtin_bit_cast(bf16x8, v); }

constexpr int kSims = 16;

}  // namespace

template <int MODE, int KQ>
__global__ __launch_bounds__(512) void k_sim_mfma(const float* __restrict__ SC, SimArgs a,
                                                  int n_sims_total, int LDSS) {
    static_assert(KQ % 8 == 0, "KQ must be a multiple of 8");
    constexpr int NSTEP = MODE == 0 ? KQ : KQ / 8;  // MFMA steps per row tile
    const int N = a.N;
    const int lane = threadIdx.x & 63, rt = threadIdx.x >> 6;
    const int n = lane & 15, g = lane >> 4;
    const int sim = blockIdx.x * kSims + n;
    const bool sim_ok = sim < n_sims_total;
    const int row_a = rt * 16 + n;            // A-fragment row of this lane
    const int j0 = rt * 16 + 4 * g;           // first of this lane's 4 output nodes

    extern __shared__ unsigned char smem_raw[];
    // FP32 mode: float S[2][16][LDSS]; BF16 modes: u16 hi[2][16][LDSS], lo[2][16][LDSS]
    float* S32 = reinterpret_cast<float*>(smem_raw);
    unsigned short* Shi = reinterpret_cast<unsigned short*>(smem_raw);
    unsigned short* Slo = Shi + 2 * kSims * LDSS;
    int* flags = reinterpret_cast<int*>(smem_raw + (MODE == 0 ? 2 * kSims * LDSS * 4
                                                              : 4 * kSims * LDSS * 2));

    // ---- A fragments (SC row tile rt) -------------------------------------
    float a32[MODE == 0 ? KQ : 1];
    u16x8 ahi[MODE == 0 ? 1 : NSTEP], alo[MODE == 2 ? NSTEP : 1];
    if constexpr (MODE == 0) {
#pragma unroll
        for (int kk = 0; kk < KQ; ++kk) {
            const int k = g * KQ + kk;
            a32[kk] = (row_a < N && k < N) ? SC[(size_t)row_a * N + k] : 0.0f;
        }
    } else {
#pragma unroll
        for (int t = 0; t < NSTEP; ++t)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int k = g * KQ + 8 * t + e;
                const float v = (row_a < N && k < N) ? SC[(size_t)row_a * N + k] : 0.0f;
                const unsigned short h = f2bf(v);
                ahi[t][e] = h;
                if constexpr (MODE == 2) alo[MODE == 2 ? t : 0][e] = f2bf(v - bf2f(h));
            }
    }
    // ---- LDS init: S_E = 0.001 on real nodes, 0 on padding ----------------
    const int nbuf = 2 * kSims * LDSS;
    for (int i = threadIdx.x; i < nbuf; i += blockDim.x) {
        const int k = i % LDSS;
        const float v = k < N ? 0.001f : 0.0f;
        if constexpr (MODE == 0) {
            S32[i] = v;
        } else {
            const unsigned short h = f2bf(v);
            Shi[i] = h;
            Slo[i] = f2bf(v - bf2f(h));
        }
    }
    if (threadIdx.x < kSims) flags[threadIdx.x] = 0;

    // ---- per-pair state ---------------------------------------------------
    bool node_ok[4];
    float SE[4], SI[4], IE[4], II[4], rE[4], rI[4], wEE[4], wEI[4], wIE[4], delta[4];
    double mIE[4];
    float ext[4][6];
    BW<float> bw[4];
    const float gj = sim_ok ? (float)(a.G[sim] * C64::J_NMDA) : 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int j = j0 + r;
        node_ok[r] = sim_ok && j < N;
        const size_t o = (size_t)sim * N + j;
        SE[r] = 0.001f; SI[r] = 0.001f;
        IE[r] = II[r] = rE[r] = rI[r] = 0.0f;
        wEE[r] = node_ok[r] ? (float)a.wEE[o] : 0.0f;
        wEI[r] = node_ok[r] ? (float)a.wEI[o] : 0.0f;
        wIE[r] = node_ok[r] ? (float)a.wIE_in[o] : 0.0f;
        delta[r] = (float)a.init_delta;
        mIE[r] = 0.0;
        for (int e = 0; e < 6; ++e) ext[r][e] = 0.0f;
    }
    int ts = 0, vol = 0, trial = 0;
    bool adjusting = a.adjust_fic != 0, failed = false;
    const int T = a.time_steps;
    const unsigned dur = (unsigned)(a.i_end - a.i_start + 1);
    const int P = (int)a.noise_period;
    const float2* noise2 = reinterpret_cast<const float2*>(a.noise);
    int buf = 0;
    __syncthreads();

    while (__syncthreads_or(sim_ok && ts <= T)) {
        const bool active = sim_ok && ts <= T;
        const int ms = ts % P;
        const float2* nzp = noise2 + (size_t)ms * 10 * N + j0;
        // noise of the lane's four nodes, one substep ahead: the loads for
        // substep s + 1 are in flight behind substep s
        float2 nz[4], nzn[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) nzn[r] = node_ok[r] && active ? nzp[r] : float2{0, 0};
        for (int s = 0; s < 10; ++s) {
#pragma unroll
            for (int r = 0; r < 4; ++r) nz[r] = nzn[r];
            if (s < 9) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    nzn[r] = node_ok[r] && active ? nzp[(size_t)(s + 1) * N + r] : float2{0, 0};
            }
            // ---- coupling on MFMA ------------------------------------------
            f32x4 acc = {0, 0, 0, 0};
            if constexpr (MODE == 0) {
                const f32x4* B4 = reinterpret_cast<const f32x4*>(S32 + (buf * kSims + n) * LDSS + g * KQ);
                f32x4 bv[KQ / 4];
#pragma unroll
                for (int q = 0; q < KQ / 4; ++q) bv[q] = B4[q];
#pragma unroll
                for (int q = 0; q < KQ / 4; ++q)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a32[4 * q + e], bv[q][e], acc, 0, 0, 0);
            } else {
                const u16x8* Bh = reinterpret_cast<const u16x8*>(Shi + (buf * kSims + n) * LDSS + g * KQ);
                const u16x8* Bl = reinterpret_cast<const u16x8*>(Slo + (buf * kSims + n) * LDSS + g * KQ);
                u16x8 bh[NSTEP], bl[MODE == 2 ? NSTEP : 1];
#pragma unroll
                for (int t = 0; t < NSTEP; ++t) {
                    bh[t] = Bh[t];
                    if constexpr (MODE == 2) bl[t] = Bl[t];
                }
#pragma unroll
                for (int t = 0; t < NSTEP; ++t) {
                    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ahi[t]), as_bf16x8(bh[t]), acc, 0, 0, 0);
                    if constexpr (MODE == 2) {
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(alo[t]), as_bf16x8(bh[t]), acc, 0, 0, 0);
                        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf16x8(ahi[t]), as_bf16x8(bl[t]), acc, 0, 0, 0);
                    }
                }
            }
            // ---- node updates ----------------------------------------------
            if (active) {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (node_ok[r])
                        Node<float>::step(acc[r], gj, wEE[r], wEI[r], wIE[r], nz[r].x, nz[r].y, SE[r],
                                          SI[r], IE[r], II[r], rE[r], rI[r]);
                const int base = ((buf ^ 1) * kSims + n) * LDSS + j0;
                if (j0 < N) {
                    if constexpr (MODE == 0) {
                        f32x4 v;
                        for (int r = 0; r < 4; ++r) v[r] = node_ok[r] ? SE[r] : 0.0f;
                        *reinterpret_cast<f32x4*>(S32 + base) = v;
                    } else {
                        u16x4 h, l;
                        for (int r = 0; r < 4; ++r) {
                            const float v = node_ok[r] ? SE[r] : 0.0f;
                            h[r] = f2bf(v);
                            l[r] = f2bf(v - bf2f(h[r]));
                        }
                        *reinterpret_cast<u16x4*>(Shi + base) = h;
                        if constexpr (MODE == 2) *reinterpret_cast<u16x4*>(Slo + base) = l;
                    }
                }
            }
            // (a finished simulation's column is independent of the others)
            __syncthreads();
            buf ^= 1;
        }
        if (active) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (node_ok[r]) bw[r].step(SE[r]);
            if (ts % a.bold_tr == 0) {
                float* bold = reinterpret_cast<float*>(a.bold);
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (node_ok[r]) {
                        bold[((size_t)sim * a.output_ts + vol) * N + j0 + r] = bw[r].bold(a.bw_k1, a.bw_k2, a.bw_k3);
                        if (vol >= a.n_vols_remove) {
                            ext[r][0] += SE[r]; ext[r][1] += IE[r]; ext[r][2] += rE[r];
                            ext[r][3] += SI[r]; ext[r][4] += II[r]; ext[r][5] += rI[r];
                        }
                    }
                ++vol;
            }
            ++ts;
        }
        // ---- numerical FIC, per simulation ------------------------------------
        const bool sample = active && adjusting && ts >= a.i_start && ts <= a.i_end;
        if (sample)
#pragma unroll
            for (int r = 0; r < 4; ++r) mIE[r] += (double)IE[r];
        const bool eval = active && adjusting && ts == a.i_end;
        if (__syncthreads_or(eval)) {
            bool out = false;
            if (eval)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (node_ok[r]) out |= fic_decide<float>(mIE[r], dur, trial, a.max_fic_trials, wIE[r], delta[r]);
            if (out) atomicOr(&flags[n], 1);
            __syncthreads();
            const bool needs = eval && flags[n] != 0;
            __syncthreads();
            if (threadIdx.x < kSims) flags[threadIdx.x] = 0;
            if (eval) {
                if (needs) {
                    if (trial < a.max_fic_trials) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            SE[r] = 0.001f; SI[r] = 0.001f; mIE[r] = 0.0;
                            bw[r].reset();
                            for (int e = 0; e < 6; ++e) ext[r][e] = 0.0f;
                        }
                        if (j0 < N) {
                            const int base = (buf * kSims + n) * LDSS + j0;
                            for (int r = 0; r < 4; ++r) {
                                const float v = node_ok[r] ? 0.001f : 0.0f;
                                if constexpr (MODE == 0) {
                                    S32[base + r] = v;
                                } else {
                                    const unsigned short h = f2bf(v);
                                    Shi[base + r] = h;
                                    if constexpr (MODE == 2) Slo[base + r] = f2bf(v - bf2f(h));
                                }
                            }
                        }
                        ts = 0; vol = 0; ++trial;
                    } else {
                        failed = true; adjusting = false;
                    }
                } else {
                    adjusting = false;
                }
            }
            __syncthreads();
        }
    }
    // ---- outputs ----------------------------------------------------------------
    const int n_ext = vol - a.n_vols_remove;
#pragma unroll
    for (int r = 0; r < 4; ++r)
        if (node_ok[r]) {
            const size_t o = (size_t)sim * N + j0 + r;
            a.wIE_out[o] = (double)wIE[r];
            for (int e = 0; e < 6; ++e)
                a.ext[((size_t)sim * 6 + e) * N + j0 + r] = n_ext > 0 ? (double)ext[r][e] / n_ext : 0.0;
        }
    if (rt == 0 && g == 0 && sim_ok) {
        a.fic_failed[sim] = failed ? 1 : 0;
        a.fic_trials[sim] = trial;
    }
}

MfmaShape mfma_shape(int N, int mode) {
    MfmaShape s{};
    s.mode = mode;
    const int KP = ((N + 31) / 32) * 32;
    s.KQ = KP / 4;
    s.waves = (N + 15) / 16;
    s.threads = s.waves * 64;
    // row stride of the LDS S_E rows: KP + pad, chosen so that (stride in
    // dwords) mod 32 == 4 and a 16-lane group's 16-byte reads are conflict-free
    const int el = mode == 0 ? 4 : 2;
    int ld = KP;
    while (((ld * el / 4) % 32) != 4) ld += 16 / el;
    s.LDSS = ld;
    s.smem = (size_t)(mode == 0 ? 2 : 4) * kSims * ld * el + kSims * sizeof(int);
    s.ok = N <= 128 && (mode >= 0 && mode <= 2) && (s.KQ == 8 || s.KQ == 16 || s.KQ == 24 || s.KQ == 32);
    return s;
}

template <int MODE>
static void launch_mode(const float* SC, const SimArgs& a, const MfmaShape& s, hipStream_t st) {
    const dim3 grid((a.B + kSims - 1) / kSims);
    switch (s.KQ) {
        case 8: hipLaunchKernelGGL((k_sim_mfma<MODE, 8>), grid, dim3(s.threads), s.smem, st, SC, a, a.B, s.LDSS); break;
        case 16: hipLaunchKernelGGL((k_sim_mfma<MODE, 16>), grid, dim3(s.threads), s.smem, st, SC, a, a.B, s.LDSS); break;
        case 24: hipLaunchKernelGGL((k_sim_mfma<MODE, 24>), grid, dim3(s.threads), s.smem, st, SC, a, a.B, s.LDSS); break;
        default: hipLaunchKernelGGL((k_sim_mfma<MODE, 32>), grid, dim3(s.threads), s.smem, st, SC, a, a.B, s.LDSS); break;
    }
}

hipError_t launch_sim_mfma(const float* SC, const SimArgs& a, int mode, hipStream_t st) {
    const MfmaShape s = mfma_shape(a.N, mode);
    if (!s.ok) return hipErrorInvalidValue;
    if (mode == 0) launch_mode<0>(SC, a, s, st);
    else if (mode == 1) launch_mode<1>(SC, a, s, st);
    else launch_mode<2>(SC, a, s, st);
    return hipGetLastError();
}

}  // namespace bnm
