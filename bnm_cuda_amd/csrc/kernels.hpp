// kernels.hpp -- host-side launch functions of the HIP kernels (internal to
// libbnm_hip.so; the public surface is include/bnm_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <vector>
#include "model.hpp"

namespace bnm {

// ---- integration (sim_kernels.hip) ----------------------------------------
struct FastShape {
    int LS;        // lanes per node within a wave (1, 2, 4)
    int KS;        // wave groups splitting the coupling dot
    int NW;        // waves per wave group
    int kc;        // SC values per lane
    int KC;        // register array size (template)
    int SS;        // LDS slice stride (elements)
    int threads;   // workgroup size
    size_t smem;   // LDS bytes
    bool ok;
};
FastShape fast_shape(int N, int LS, int KS, size_t elem);

template <class R>
hipError_t launch_sim_plain(const R* SC, const SimArgs& a, hipStream_t st);
template <class R>
hipError_t launch_sim_fast(const R* SC, const SimArgs& a, int LS, int KS, hipStream_t st);
// large N: SC streamed from a transposed copy SCt[k][j] in global memory
FastShape stream_shape(int N, int KS, size_t elem);
// zero k-rows the chunked FP32 SC carries after its N rows (>= the k-range
// padding of any hybrid shape: KS * 16 with KS <= 16)
constexpr size_t kSCtPad = 256;
// FP32 streaming kernel with register- and LDS-resident SC (N > 128); the
// shape's KC field is the register part, SS the LDS part (values per lane)
// (rr >= 0: register part forced, a multiple of 16 up to 160)
FastShape hyb_shape(int N, int KS, size_t lds_bytes, int rr = -1, int form = 0);
// SCc: chunked layout SCc[((k/16)*N + j)*16 + k%16] = SC[j][k], zero for k >= N
// form: 0 S_E by 16-byte LDS broadcast reads, 1 by DPP row broadcast, 2 DPP
// broadcast with three streamed chunks in flight
hipError_t launch_sim_hyb(const float* SCc, const SimArgs& a, int KS, size_t lds_bytes, int rr, int form,
                          hipStream_t st);
// Sparse kernel (N <= 512; FP32 and FP64): each SC row (or half-row, two
// lanes per node) as P slots of (weight, byte step to the next nonzero's
// S_E), the first kSparseRW weights per lane in registers and in FP32 up to
// kSparseLL more in LDS; no SC traffic per substep
constexpr int kSparseRW = 112, kSparseLL = 96;
constexpr int kSparseRW64 = 64;   // FP64: register-resident slots only
constexpr size_t kStaticLds = 1024;   // headroom for a kernel's static LDS
struct SparseShape {
    int P;         // slots per lane (multiple of 8)
    int LS;        // lanes per node (1, or 2: each takes half the row's k-range)
    int RW;        // register-resident slots (template)
    int LL;        // LDS-resident slots
    int threads;
    size_t smem;
    bool ok;
};
// (T = float: 4-byte S_E, steps of up to 63 nodes; T = double: 8-byte S_E,
// steps of up to 31 nodes)
template <class T>
int build_sparse_sc(const T* SC, int N, int LS, std::vector<T>& W, std::vector<unsigned>& D);
SparseShape sparse_shape(int N, int P, int LS, size_t lds_bytes, size_t elem);
// strict: the strict FP32 node update (NodeStrict, FIC in FP64), else Node<float>
hipError_t launch_sim_sparse(const float* W, const unsigned* D, const SimArgs& a, int P, int LS, bool strict,
                             size_t lds_bytes, hipStream_t st);
// FP64 sparse kernel: Node<double> in the reference's expression order, FP64
// weights and S_E, all slots in registers (P <= kSparseRW64)
hipError_t launch_sim_sparse64(const double* W, const unsigned* D, const SimArgs& a, int P, int LS, size_t lds_bytes,
                               hipStream_t st);
// FP32 split kernel with the node update divided between lanes by population
FastShape ei_shape(int N, int LS, bool bc = false);
// FP64 hybrid: register / LDS-resident SC, the rest streamed from the chunked
// copy SCc64[((k/8)*N + j)*8 + k%8] (zero for k >= N, kSCtPad padding rows)
FastShape hyb64_shape(int N, int KS, size_t lds_bytes, int rr = -1);
hipError_t launch_sim_hyb64(const double* SCc, const SimArgs& a, int KS, size_t lds_bytes, int rr, hipStream_t st);
// FP64 E/I split (reference expression order per pool; bit-identical to the
// FP64 split kernel with the same LS)
FastShape ei64_shape(int N, int LS);
hipError_t launch_sim_ei64(const double* SC, const SimArgs& a, int LS, hipStream_t st);
// form: 0 as written, 1 DPP row broadcast (LS = 4, N <= 128), 2 / 3 short-chain
// update with four / one partial sums in the coupling dot
hipError_t launch_sim_ei(const float* SC, const SimArgs& a, int LS, int form, hipStream_t st);
template <class R>
hipError_t launch_sim_stream(const R* SCt, const SimArgs& a, int KS, hipStream_t st);

// MFMA-batched integration (sim_mfma.hip): 16 simulations per workgroup,
// mode 0 FP32, 1 BF16, 2 BF16x3 coupling operands.
struct MfmaShape {
    int mode;
    int KQ;        // K values per lane group
    int waves;     // one 16-node row tile per wave
    int threads;
    int LDSS;      // LDS row stride (elements)
    size_t smem;
    bool ok;
};
MfmaShape mfma_shape(int N, int mode);
hipError_t launch_sim_mfma(const float* SC, const SimArgs& a, int mode, hipStream_t st);

// analytical FIC on the device (fic_kernel.hip), fic.cpp:130-207
hipError_t launch_fic_analytical(const double* SC, int N, int B, const double* G,
                                 const double* wEE, const double* wEI, double* wIE_out,
                                 int* unstable, hipStream_t st);

// ---- FC / FCD (fc_kernels.hip) ----------------------------------------------
// Per (sim, segment) node moments of BOLD over [seg_start, seg_start+seg_len):
// mean and 1/sqrt(sum (x-mean)^2), FP64 compensated sums.
template <class T>
hipError_t launch_seg_moments(const T* bold, int B, int N, int output_ts, int n_seg,
                              const int* seg_start, const int* seg_len, double* mean,
                              double* inv_ssd, hipStream_t st);
// Gram of z-scored columns.  Mode FC: items (sim, seg), X[t][n] =
// bold[(start+t)*N + n]; outputs the pair list (pair table pidx[N*N], -1 =
// skipped) into fc (seg 0) and wfc (seg >= 1).  Mode FCD: items sim, X[e][w]
// = wfc[(sim*nw + w)*np + e]; outputs the wi>wj triangle into fcd.
template <class Tin, class Tc, class Tout>
hipError_t launch_gram_fc(const Tin* bold, int B, int N, int output_ts, int n_seg,
                          const int* seg_start, const int* seg_len, const double* mean,
                          const double* inv_ssd, const int* pidx, int np, Tout* fc, Tout* wfc,
                          hipStream_t st);
template <class Tin, class Tc, class Tout>
hipError_t launch_gram_fcd(const Tin* wfc, int B, int nw, int np, const double* mean,
                           const double* inv_ssd, Tout* fcd, hipStream_t st, int col0 = 0, int K = -1);
// BF16-operand variants (FP32 accumulation, FP32 outputs); split = hi/lo
// three-product form
template <class Tin>
hipError_t launch_gram_fc_bf16(const Tin* bold, int B, int N, int output_ts, int n_seg,
                               const int* seg_start, const int* seg_len, const double* mean,
                               const double* inv_ssd, const int* pidx, int np, float* fc, float* wfc,
                               bool split, hipStream_t st);
hipError_t launch_gram_fcd_bf16(const float* wfc, int B, int nw, int np, const double* mean,
                                const double* inv_ssd, float* fcd, bool split, hipStream_t st, int col0 = 0,
                                int K = -1);
// FCD from z-scored rows: z-scores the FP32 window-FC rows into zhi (FP32
// [B*nw][npp], or BF16 hi) and zlo (BF16 lo, mode 2), npp = np rounded up to
// 64, then Grams them (mode 0 FP32, 1 BF16, 2 BF16 hi/lo three products)
// (tile 128: FP32 Gram over 128x128 output tiles; otherwise 64x64, the default)
hipError_t launch_fcd_z(const float* wfc, int B, int nw, int np, const double* mean, const double* inv_ssd,
                        void* zhi, unsigned short* zlo, int mode, float* fcd, hipStream_t st, int tile = 64);
// Row moments of the window-FC matrix (mean, 1/ssd over the np edges).
template <class T>
hipError_t launch_row_moments(const T* x, long long rows, int len, double* mean, double* inv_ssd,
                              hipStream_t st, int stride = -1, int offset = 0);

// ---- fit (gof_kernels.hip) ---------------------------------------------------
struct GofArgs {
    int B, N, np;
    long long nwp;               // FCD values per sim
    int m;                       // unique empirical FCD breakpoints
    const double* emp_fc_c;      // [np] empirical FC minus its mean
    double emp_fc_ssd;           // sqrt(sum emp_fc_c^2)
    double emp_fc_mean;
    const double* ks_u;          // [m] sorted unique empirical FCD values
    const double* ks_fe;         // [m] empirical CDF at ks_u
    int* bins;                   // [B][2m+1] scratch, zero before launch
    int* nan_flag;               // [B] scratch, zero before launch
    const double* ext;           // [B][6][N] (r_E = row 2)
    const int* fic_failed;       // [B]
    int sim_only, no_fcd, use_fc_diff, calc_penalty, fic_reject;
    double max_cost, fic_scale;
    double* corr;                // [B]
    double* fic_pen;             // [B]
};
template <class T>
hipError_t launch_ks_bin(const T* fcd, const GofArgs& g, hipStream_t st);
template <class T>
hipError_t launch_gof(const T* fc, const GofArgs& g, hipStream_t st);
hipError_t launch_fill_i32(int* p, long long n, int v, hipStream_t st);

}  // namespace bnm
