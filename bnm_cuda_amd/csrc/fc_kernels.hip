// fc_kernels.hip -- FC, windowed FC and FCD as Gram products of z-scored
// columns on MFMA (bold_stats / window_bold_stats / fc / window_fc_stats /
// fcd of bnm.cu:317-551; fc_tril / fcd_tril of check_fit.cpp:19-136).
//
// Pearson r(x_i, x_j) = sum_t z_ti z_tj with z_ti = (x_ti - mean_i) / ssd_i and
// ssd_i = sqrt(sum_t (x_ti - mean_i)^2).  The moments are FP64 sums with
// Neumaier compensation; the cross products run on MFMA in the Gram precision
// (FP64: v_mfma_f64_16x16x4f64, FP32: v_mfma_f32_16x16x4f32).
#include "kernels.hpp"

namespace bnm {

namespace {

// Neumaier-compensated running sum
struct KSum {
    double s = 0, c = 0;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

template <class T> __device__ __forceinline__ double to_d(T v) { return (double)v; }

// block-wide sum of a double (blockDim.x multiple of 64, <= 1024)
__device__ double block_sum(double v, double* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    const int nwv = blockDim.x >> 6;
    for (int i = 0; i < nwv; ++i) t += red[i];
    return t;
}

}  // namespace

// ---- node moments per (sim, segment) -----------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_seg_moments(const T* __restrict__ bold, int N,
                                                     int output_ts, int n_seg,
                                                     const int* __restrict__ seg_start,
                                                     const int* __restrict__ seg_len,
                                                     double* __restrict__ mean,
                                                     double* __restrict__ inv_ssd) {
    const int seg = blockIdx.x, sim = blockIdx.y;
    const int t0 = seg_start[seg], len = seg_len[seg];
    const T* x = bold + ((size_t)sim * output_ts + t0) * N;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        KSum s;
        for (int t = 0; t < len; ++t) s.add(to_d(x[(size_t)t * N + j]));
        const double m = s.value() / len;
        KSum q;
        for (int t = 0; t < len; ++t) {
            const double d = to_d(x[(size_t)t * N + j]) - m;
            q.add(d * d);
        }
        const size_t o = ((size_t)sim * n_seg + seg) * N + j;
        mean[o] = m;
        inv_ssd[o] = 1.0 / sqrt(q.value());
    }
}

// ---- row moments (window-FC rows over the edge axis) -------------------------
template <class T>
__global__ __launch_bounds__(256) void k_row_moments(const T* __restrict__ x, int len, int stride,
                                                     int offset, double* __restrict__ mean,
                                                     double* __restrict__ inv_ssd) {
    __shared__ double red[16];
    const size_t row = blockIdx.x;
    const T* r = x + row * stride + offset;
    KSum s;
    for (int i = threadIdx.x; i < len; i += blockDim.x) s.add(to_d(r[i]));
    const double tot = block_sum(s.s, red) + block_sum(s.c, red);
    const double m = tot / len;
    KSum q;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const double d = to_d(r[i]) - m;
        q.add(d * d);
    }
    const double ss = block_sum(q.s, red) + block_sum(q.c, red);
    if (threadIdx.x == 0) {
        mean[row] = m;
        inv_ssd[row] = 1.0 / sqrt(ss);
    }
}

// ---- MFMA Gram ------------------------------------------------------------------
namespace {

template <class Tc> struct MF;
// Output layout of the 16x16 result: element r of lane l sits at
// (row(l, r), l % 16).  v_mfma_f32_16x16x4f32 stacks each lane's 4 values
// down consecutive rows; v_mfma_f64_16x16x4f64 interleaves them by 4 rows.
template <> struct MF<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
    static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct MF<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ int row(int lane, int r) { return 4 * r + (lane >> 4); }
    static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

constexpr int TM = 64;    // output tile edge (4 waves x 32x32)
constexpr int TK = 16;    // K chunk staged in LDS
constexpr int LDT = TM + 16;

struct GramItem {
    const void* base;     // element (k, m) at base[k*ldk + m*ldm]
    int K, M, ldk, ldm;
    const double* mu;     // [M]
    const double* is;     // [M]
    void* out;            // pair-indexed output of this item
};

__device__ __forceinline__ void tile_pair(int p, int& ti, int& tj) {
    int t = (int)((sqrt(8.0 * p + 1.0) - 1.0) * 0.5);
    while ((t + 1) * (t + 2) / 2 <= p) ++t;
    while (t * (t + 1) / 2 > p) --t;
    ti = t;
    tj = p - t * (t + 1) / 2;
}

// One 64x64 lower-triangle output tile (ti >= tj) of the Gram of item `it`.
// pidx: pair index table [M*M] (-1 = not in the list) or nullptr for the
// closed form i*(i-1)/2 + j.
template <class Tin, class Tc, class Tout>
__device__ void gram_tile(const GramItem& it, int ti, int tj, const int* __restrict__ pidx,
                          Tc (*As)[LDT], Tc (*Bs)[LDT]) {
    using V4 = typename MF<Tc>::v4;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wi = w >> 1, wj = w & 1;
    const Tin* X = reinterpret_cast<const Tin*>(it.base);
    V4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = V4{0, 0, 0, 0};
    const bool mfast = it.ldm == 1;
    for (int k0 = 0; k0 < it.K; k0 += TK) {
#pragma unroll
        for (int r = 0; r < (TK * TM) / 256; ++r) {
            const int e = tid + 256 * r;
            const int m = mfast ? (e % TM) : (e / TK);
            const int k = mfast ? (e / TM) : (e % TK);
            const int kg = k0 + k;
            const int mi = ti * TM + m, mj = tj * TM + m;
            Tc za = 0, zb = 0;
            if (kg < it.K) {
                if (mi < it.M)
                    za = (Tc)((to_d(X[(size_t)kg * it.ldk + (size_t)mi * it.ldm]) - it.mu[mi]) * it.is[mi]);
                if (mj < it.M)
                    zb = (Tc)((to_d(X[(size_t)kg * it.ldk + (size_t)mj * it.ldm]) - it.mu[mj]) * it.is[mj]);
            }
            As[k][m] = za;
            Bs[k][m] = zb;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK / 4; ++kk) {
            const int kr = kk * 4 + (lane >> 4);
            Tc a[2], b[2];
            for (int q = 0; q < 2; ++q) {
                a[q] = As[kr][wi * 32 + q * 16 + (lane & 15)];
                b[q] = Bs[kr][wj * 32 + q * 16 + (lane & 15)];
            }
            for (int qa = 0; qa < 2; ++qa)
                for (int qb = 0; qb < 2; ++qb) acc[qa][qb] = MF<Tc>::mma(a[qa], b[qb], acc[qa][qb]);
        }
        __syncthreads();
    }
    Tout* out = reinterpret_cast<Tout*>(it.out);
    for (int qa = 0; qa < 2; ++qa)
        for (int qb = 0; qb < 2; ++qb)
            for (int r = 0; r < 4; ++r) {
                const int row = ti * TM + wi * 32 + qa * 16 + MF<Tc>::row(lane, r);
                const int col = tj * TM + wj * 32 + qb * 16 + (lane & 15);
                if (row < it.M && col < row) {
                    const long long p = pidx ? pidx[(size_t)row * it.M + col]
                                             : (long long)row * (row - 1) / 2 + col;
                    if (p >= 0) out[p] = (Tout)acc[qa][qb][r];
                }
            }
}

// BF16 operands, FP32 accumulation (v_mfma_f32_16x16x32_bf16): z-scored
// values are formed in FP64, rounded to bfloat16 (nearest even) in the tile
// loader and staged k-contiguous per row, so each lane fetches its 8
// consecutive k (lane group g = lane / 16 holds k = 8g .. 8g+7) with one
// 16-byte LDS read.  Output layout as v_mfma_f32_16x16x4f32.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
constexpr int TKB = 32;          // K chunk
constexpr int LDB = TKB + 8;     // u16 per staged row (80 B: rows stay 16-byte aligned)

// gfx950's v_cvt_pk_bf16_f32: round to nearest even, the same bits as
// adding 0x7fff + lsb to the FP32 pattern for finite inputs
__device__ __forceinline__ unsigned short bf16_bits(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }

// SPLIT: each z-score is carried as hi = bf16(z) and lo = bf16(z - hi), and
// the product as hi*hi + hi*lo + lo*hi (three MFMAs; the dropped lo*lo term
// is 2^-16 relative), which brings the Gram to FP32-like accuracy.  The
// As/Bs arrays then hold [hi rows | lo rows] (2*TM rows).
template <class Tin, bool SPLIT>
__device__ void gram_tile_bf16(const GramItem& it, int ti, int tj, const int* __restrict__ pidx,
                               unsigned short (*As)[LDB], unsigned short (*Bs)[LDB]) {
    typedef float v4 __attribute__((ext_vector_type(4)));
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wi = w >> 1, wj = w & 1;
    const Tin* X = reinterpret_cast<const Tin*>(it.base);
    v4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{0, 0, 0, 0};
    const bool mfast = it.ldm == 1;
    for (int k0 = 0; k0 < it.K; k0 += TKB) {
#pragma unroll
        for (int r = 0; r < (TKB * TM) / 256; ++r) {
            const int e = tid + 256 * r;
            const int m = mfast ? (e % TM) : (e / TKB);
            const int k = mfast ? (e / TM) : (e % TKB);
            const int kg = k0 + k;
            const int mi = ti * TM + m, mj = tj * TM + m;
            float za = 0, zb = 0;
            if (kg < it.K) {
                if (mi < it.M)
                    za = (float)((to_d(X[(size_t)kg * it.ldk + (size_t)mi * it.ldm]) - it.mu[mi]) * it.is[mi]);
                if (mj < it.M)
                    zb = (float)((to_d(X[(size_t)kg * it.ldk + (size_t)mj * it.ldm]) - it.mu[mj]) * it.is[mj]);
            }
            const unsigned short ha = bf16_bits(za), hb = bf16_bits(zb);
            As[m][k] = ha;
            Bs[m][k] = hb;
            if constexpr (SPLIT) {
                As[TM + m][k] = bf16_bits(za - __uint_as_float((unsigned)ha << 16));
                Bs[TM + m][k] = bf16_bits(zb - __uint_as_float((unsigned)hb << 16));
            }
        }
        __syncthreads();
        u16x8 a[2], b[2], al[2], bl[2];
        for (int q = 0; q < 2; ++q) {
            const int ra = wi * 32 + q * 16 + (lane & 15), rb = wj * 32 + q * 16 + (lane & 15);
            a[q] = *reinterpret_cast<const u16x8*>(&As[ra][8 * (lane >> 4)]);
            b[q] = *reinterpret_cast<const u16x8*>(&Bs[rb][8 * (lane >> 4)]);
            if constexpr (SPLIT) {
                al[q] = *reinterpret_cast<const u16x8*>(&As[TM + ra][8 * (lane >> 4)]);
                bl[q] = *reinterpret_cast<const u16x8*>(&Bs[TM + rb][8 * (lane >> 4)]);
            }
        }
        for (int qa = 0; qa < 2; ++qa)
            for (int qb = 0; qb < 2; ++qb) {
                acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, a[qa]), __builtin_bit_cast(bf16x8, b[qb]), acc[qa][qb], 0, 0, 0);
                if constexpr (SPLIT) {
                    acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, a[qa]), __builtin_bit_cast(bf16x8, bl[qb]), acc[qa][qb], 0, 0, 0);
                    acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, al[qa]), __builtin_bit_cast(bf16x8, b[qb]), acc[qa][qb], 0, 0, 0);
                }
            }
        __syncthreads();
    }
    float* out = reinterpret_cast<float*>(it.out);
    for (int qa = 0; qa < 2; ++qa)
        for (int qb = 0; qb < 2; ++qb)
            for (int r = 0; r < 4; ++r) {
                const int row = ti * TM + wi * 32 + qa * 16 + 4 * (lane >> 4) + r;
                const int col = tj * TM + wj * 32 + qb * 16 + (lane & 15);
                if (row < it.M && col < row) {
                    const long long p = pidx ? pidx[(size_t)row * it.M + col]
                                             : (long long)row * (row - 1) / 2 + col;
                    if (p >= 0) out[p] = acc[qa][qb][r];
                }
            }
}

}  // namespace

template <class Tin, bool SPLIT>
__global__ __launch_bounds__(256) void k_gram_fc_bf16(const Tin* __restrict__ bold, int N, int output_ts,
                                                      int n_seg, const int* __restrict__ seg_start,
                                                      const int* __restrict__ seg_len,
                                                      const double* __restrict__ mean,
                                                      const double* __restrict__ inv_ssd,
                                                      const int* __restrict__ pidx, int np,
                                                      float* __restrict__ fc, float* __restrict__ wfc) {
    __shared__ __align__(16) unsigned short As[SPLIT ? 2 * TM : TM][LDB];
    __shared__ __align__(16) unsigned short Bs[SPLIT ? 2 * TM : TM][LDB];
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int seg = blockIdx.y, sim = blockIdx.z;
    GramItem it;
    it.base = bold + ((size_t)sim * output_ts + seg_start[seg]) * N;
    it.K = seg_len[seg];
    it.M = N;
    it.ldk = N;
    it.ldm = 1;
    const size_t mo = ((size_t)sim * n_seg + seg) * N;
    it.mu = mean + mo;
    it.is = inv_ssd + mo;
    it.out = seg == 0 ? (void*)(fc + (size_t)sim * np)
                      : (void*)(wfc + ((size_t)sim * (n_seg - 1) + (seg - 1)) * np);
    gram_tile_bf16<Tin, SPLIT>(it, ti, tj, pidx, As, Bs);
}

template <bool SPLIT>
__global__ __launch_bounds__(256) void k_gram_fcd_bf16(const float* __restrict__ wfc, int nw, int np, int col0,
                                                       int K, const double* __restrict__ mean,
                                                       const double* __restrict__ inv_ssd,
                                                       float* __restrict__ fcd) {
    __shared__ __align__(16) unsigned short As[SPLIT ? 2 * TM : TM][LDB];
    __shared__ __align__(16) unsigned short Bs[SPLIT ? 2 * TM : TM][LDB];
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int sim = blockIdx.z;
    GramItem it;
    it.base = wfc + (size_t)sim * nw * np + col0;
    it.K = K;
    it.M = nw;
    it.ldk = 1;
    it.ldm = np;
    it.mu = mean + (size_t)sim * nw;
    it.is = inv_ssd + (size_t)sim * nw;
    it.out = fcd + (size_t)sim * ((long long)nw * (nw - 1) / 2);
    gram_tile_bf16<float, SPLIT>(it, ti, tj, nullptr, As, Bs);
}

template <class Tin, class Tc, class Tout>
__global__ __launch_bounds__(256) void k_gram_fc(const Tin* __restrict__ bold, int N, int output_ts,
                                                 int n_seg, const int* __restrict__ seg_start,
                                                 const int* __restrict__ seg_len,
                                                 const double* __restrict__ mean,
                                                 const double* __restrict__ inv_ssd,
                                                 const int* __restrict__ pidx, int np,
                                                 Tout* __restrict__ fc, Tout* __restrict__ wfc) {
    __shared__ Tc As[TK][LDT];
    __shared__ Tc Bs[TK][LDT];
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int seg = blockIdx.y, sim = blockIdx.z;
    GramItem it;
    it.base = bold + ((size_t)sim * output_ts + seg_start[seg]) * N;
    it.K = seg_len[seg];
    it.M = N;
    it.ldk = N;
    it.ldm = 1;
    const size_t mo = ((size_t)sim * n_seg + seg) * N;
    it.mu = mean + mo;
    it.is = inv_ssd + mo;
    it.out = seg == 0 ? (void*)(fc + (size_t)sim * np)
                      : (void*)(wfc + ((size_t)sim * (n_seg - 1) + (seg - 1)) * np);
    gram_tile<Tin, Tc, Tout>(it, ti, tj, pidx, As, Bs);
}

template <class Tin, class Tc, class Tout>
__global__ __launch_bounds__(256) void k_gram_fcd(const Tin* __restrict__ wfc, int nw, int np, int col0,
                                                  int K, const double* __restrict__ mean,
                                                  const double* __restrict__ inv_ssd,
                                                  Tout* __restrict__ fcd) {
    __shared__ Tc As[TK][LDT];
    __shared__ Tc Bs[TK][LDT];
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int sim = blockIdx.z;
    GramItem it;
    it.base = wfc + (size_t)sim * nw * np + col0;
    it.K = K;
    it.M = nw;
    it.ldk = 1;
    it.ldm = np;
    it.mu = mean + (size_t)sim * nw;
    it.is = inv_ssd + (size_t)sim * nw;
    it.out = fcd + (size_t)sim * ((long long)nw * (nw - 1) / 2);
    gram_tile<Tin, Tc, Tout>(it, ti, tj, nullptr, As, Bs);
}

// ---- FCD Gram over pre-z-scored rows -----------------------------------------
// The window-FC matrix (per simulation nw rows of np edges) is z-scored once
// per row into Z [rows][npp] (npp = np rounded up to 64, zero padded): FP32,
// or BF16 hi (+ lo = bf16(z - hi) for the split form).  The Gram tiles then
// move 16-byte vectors from global memory to LDS with no per-element
// arithmetic, and the MFMA issue rate bounds the loop instead of the loader.
namespace {
constexpr int ZT = 64;         // output tile edge
constexpr int ZK32 = 32;       // K chunk, FP32 operands
constexpr int ZKB = 64;        // K chunk, BF16 operands
constexpr int ZLD32 = ZK32 + 4;   // floats per staged row (bank offset)
constexpr int ZLDB = ZKB + 8;     // u16 per staged row
}  // namespace

template <int MODE>   // 0: FP32, 1: BF16, 2: BF16 hi + lo
__global__ __launch_bounds__(256) void k_zscore_rows(const float* __restrict__ x, int np, int npp,
                                                     const double* __restrict__ mean,
                                                     const double* __restrict__ inv_ssd, void* __restrict__ zhi,
                                                     unsigned short* __restrict__ zlo) {
    const size_t row = blockIdx.x;
    const double m = mean[row], is = inv_ssd[row];
    const float* r = x + row * np;
    for (int i = threadIdx.x; i < npp; i += blockDim.x) {
        const float z = i < np ? (float)(((double)r[i] - m) * is) : 0.0f;
        if constexpr (MODE == 0) {
            reinterpret_cast<float*>(zhi)[row * npp + i] = z;
        } else {
            const unsigned short h = bf16_bits(z);
            reinterpret_cast<unsigned short*>(zhi)[row * npp + i] = h;
            if constexpr (MODE == 2) zlo[row * npp + i] = bf16_bits(z - __uint_as_float((unsigned)h << 16));
        }
    }
}

// FP32 operands: tile rows ti*64.., tj*64.. of one simulation's Z
__global__ __launch_bounds__(256) void k_gram_z32(const float* __restrict__ Z, int nw, int npp,
                                                  float* __restrict__ fcd) {
    __shared__ __align__(16) float As[ZT][ZLD32];
    __shared__ __align__(16) float Bs[ZT][ZLD32];
    typedef float v4 __attribute__((ext_vector_type(4)));
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int sim = blockIdx.z;
    const float* Zs = Z + (size_t)sim * nw * npp;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wi = w >> 1, wj = w & 1;
    v4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{0, 0, 0, 0};
    // loader: 64 rows x 32 k = 512 float4 per operand, two per thread
    const int lr = tid >> 3, lc = (tid & 7) * 4;   // rows lr and lr + 32, k lc..lc+3
    auto ld = [&](int t, int r, int k0) -> float4 {
        const int row = t * ZT + r;
        return row < nw ? *reinterpret_cast<const float4*>(Zs + (size_t)row * npp + k0 + lc)
                        : float4{0, 0, 0, 0};
    };
    float4 a0 = ld(ti, lr, 0), a1 = ld(ti, lr + 32, 0), b0 = ld(tj, lr, 0), b1 = ld(tj, lr + 32, 0);
    for (int k0 = 0; k0 < npp; k0 += ZK32) {
        *reinterpret_cast<float4*>(&As[lr][lc]) = a0;
        *reinterpret_cast<float4*>(&As[lr + 32][lc]) = a1;
        *reinterpret_cast<float4*>(&Bs[lr][lc]) = b0;
        *reinterpret_cast<float4*>(&Bs[lr + 32][lc]) = b1;
        __syncthreads();
        if (k0 + ZK32 < npp) {   // next chunk in flight behind the MFMAs
            a0 = ld(ti, lr, k0 + ZK32); a1 = ld(ti, lr + 32, k0 + ZK32);
            b0 = ld(tj, lr, k0 + ZK32); b1 = ld(tj, lr + 32, k0 + ZK32);
        }
#pragma unroll
        for (int s = 0; s < ZK32 / 4; ++s) {
            float a[2], b[2];
            for (int q = 0; q < 2; ++q) {
                a[q] = As[wi * 32 + q * 16 + (lane & 15)][4 * s + (lane >> 4)];
                b[q] = Bs[wj * 32 + q * 16 + (lane & 15)][4 * s + (lane >> 4)];
            }
            for (int qa = 0; qa < 2; ++qa)
                for (int qb = 0; qb < 2; ++qb)
                    acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[qa], b[qb], acc[qa][qb], 0, 0, 0);
        }
        __syncthreads();
    }
    float* out = fcd + (size_t)sim * ((long long)nw * (nw - 1) / 2);
    for (int qa = 0; qa < 2; ++qa)
        for (int qb = 0; qb < 2; ++qb)
            for (int r = 0; r < 4; ++r) {
                const int row = ti * ZT + wi * 32 + qa * 16 + 4 * (lane >> 4) + r;
                const int col = tj * ZT + wj * 32 + qb * 16 + (lane & 15);
                if (row < nw && col < row) out[(long long)row * (row - 1) / 2 + col] = acc[qa][qb][r];
            }
}

// BF16 operands (SPLIT: hi and lo arrays, three products)
template <bool SPLIT>
__global__ __launch_bounds__(256) void k_gram_zbf16(const unsigned short* __restrict__ Zh,
                                                    const unsigned short* __restrict__ Zl, int nw, int npp,
                                                    float* __restrict__ fcd) {
    constexpr int NB = SPLIT ? 2 : 1;
    __shared__ __align__(16) unsigned short As[NB][ZT][ZLDB];
    __shared__ __align__(16) unsigned short Bs[NB][ZT][ZLDB];
    typedef float v4 __attribute__((ext_vector_type(4)));
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int sim = blockIdx.z;
    const size_t base = (size_t)sim * nw * npp;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wi = w >> 1, wj = w & 1;
    v4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = v4{0, 0, 0, 0};
    // loader: 64 rows x 64 k of u16 = 512 x 16 B per operand and array, two per thread
    const int lr = tid >> 3, lc = (tid & 7) * 8;
    auto ld = [&](const unsigned short* Zp, int t, int r, int k0) -> u16x8 {
        const int row = t * ZT + r;
        return row < nw ? *reinterpret_cast<const u16x8*>(Zp + base + (size_t)row * npp + k0 + lc)
                        : u16x8{0, 0, 0, 0, 0, 0, 0, 0};
    };
    u16x8 ra[NB][2], rb[NB][2];
    const unsigned short* Zp[2] = {Zh, Zl};
    for (int h = 0; h < NB; ++h)
        for (int q = 0; q < 2; ++q) { ra[h][q] = ld(Zp[h], ti, lr + 32 * q, 0); rb[h][q] = ld(Zp[h], tj, lr + 32 * q, 0); }
    for (int k0 = 0; k0 < npp; k0 += ZKB) {
        for (int h = 0; h < NB; ++h)
            for (int q = 0; q < 2; ++q) {
                *reinterpret_cast<u16x8*>(&As[h][lr + 32 * q][lc]) = ra[h][q];
                *reinterpret_cast<u16x8*>(&Bs[h][lr + 32 * q][lc]) = rb[h][q];
            }
        __syncthreads();
        if (k0 + ZKB < npp)
            for (int h = 0; h < NB; ++h)
                for (int q = 0; q < 2; ++q) {
                    ra[h][q] = ld(Zp[h], ti, lr + 32 * q, k0 + ZKB);
                    rb[h][q] = ld(Zp[h], tj, lr + 32 * q, k0 + ZKB);
                }
#pragma unroll
        for (int s = 0; s < ZKB / 32; ++s) {
            u16x8 a[NB][2], b[NB][2];
            for (int h = 0; h < NB; ++h)
                for (int q = 0; q < 2; ++q) {
                    a[h][q] = *reinterpret_cast<const u16x8*>(&As[h][wi * 32 + q * 16 + (lane & 15)][32 * s + 8 * (lane >> 4)]);
                    b[h][q] = *reinterpret_cast<const u16x8*>(&Bs[h][wj * 32 + q * 16 + (lane & 15)][32 * s + 8 * (lane >> 4)]);
                }
            for (int qa = 0; qa < 2; ++qa)
                for (int qb = 0; qb < 2; ++qb) {
                    acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, a[0][qa]), __builtin_bit_cast(bf16x8, b[0][qb]), acc[qa][qb], 0, 0, 0);
                    if constexpr (SPLIT) {
                        acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, a[0][qa]), __builtin_bit_cast(bf16x8, b[1][qb]), acc[qa][qb], 0, 0, 0);
                        acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            __builtin_bit_cast(bf16x8, a[1][qa]), __builtin_bit_cast(bf16x8, b[0][qb]), acc[qa][qb], 0, 0, 0);
                    }
                }
        }
        __syncthreads();
    }
    float* out = fcd + (size_t)sim * ((long long)nw * (nw - 1) / 2);
    for (int qa = 0; qa < 2; ++qa)
        for (int qb = 0; qb < 2; ++qb)
            for (int r = 0; r < 4; ++r) {
                const int row = ti * ZT + wi * 32 + qa * 16 + 4 * (lane >> 4) + r;
                const int col = tj * ZT + wj * 32 + qb * 16 + (lane & 15);
                if (row < nw && col < row) out[(long long)row * (row - 1) / 2 + col] = acc[qa][qb][r];
            }
}

// FP32 operands, 128x128 output tiles: each wave computes a 64x64 block as a
// 4x4 grid of 16x16 MFMA tiles, so a K step reads 8 fragments from LDS for
// 16 MFMAs (the 64x64 tile reads 4 for 4), and every z-scored row is read
// from memory by half as many tiles.  In a diagonal tile the wave whose block
// lies wholly above the diagonal (rows 0-63, columns 64-127) only loads.
__global__ __launch_bounds__(256) void k_gram_z32_t128(const float* __restrict__ Z, int nw, int npp,
                                                       float* __restrict__ fcd) {
    constexpr int TT = 128;
    __shared__ __align__(16) float As[TT][ZLD32];
    __shared__ __align__(16) float Bs[TT][ZLD32];
    typedef float v4 __attribute__((ext_vector_type(4)));
    int ti, tj;
    tile_pair(blockIdx.x, ti, tj);
    const int sim = blockIdx.z;
    const float* Zs = Z + (size_t)sim * nw * npp;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wi = w >> 1, wj = w & 1;
    const bool idle = ti == tj && wi < wj;   // block above the diagonal
    v4 acc[4][4];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) acc[a][b] = v4{0, 0, 0, 0};
    // loader: 128 rows x 32 k = 1024 float4 per operand, four per thread
    const int lr = tid >> 3, lc = (tid & 7) * 4;   // rows lr + 32q, k lc..lc+3
    auto ld = [&](int t, int r, int k0) -> float4 {
        const int row = t * TT + r;
        return row < nw ? *reinterpret_cast<const float4*>(Zs + (size_t)row * npp + k0 + lc)
                        : float4{0, 0, 0, 0};
    };
    float4 ra[4], rb[4];
    for (int q = 0; q < 4; ++q) { ra[q] = ld(ti, lr + 32 * q, 0); rb[q] = ld(tj, lr + 32 * q, 0); }
    for (int k0 = 0; k0 < npp; k0 += ZK32) {
        for (int q = 0; q < 4; ++q) {
            *reinterpret_cast<float4*>(&As[lr + 32 * q][lc]) = ra[q];
            *reinterpret_cast<float4*>(&Bs[lr + 32 * q][lc]) = rb[q];
        }
        __syncthreads();
        if (k0 + ZK32 < npp)   // next chunk in flight behind the MFMAs
            for (int q = 0; q < 4; ++q) {
                ra[q] = ld(ti, lr + 32 * q, k0 + ZK32);
                rb[q] = ld(tj, lr + 32 * q, k0 + ZK32);
            }
        if (!idle) {
#pragma unroll
            for (int s = 0; s < ZK32 / 4; ++s) {
                float a[4], b[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    a[q] = As[wi * 64 + q * 16 + (lane & 15)][4 * s + (lane >> 4)];
                    b[q] = Bs[wj * 64 + q * 16 + (lane & 15)][4 * s + (lane >> 4)];
                }
#pragma unroll
                for (int qa = 0; qa < 4; ++qa)
#pragma unroll
                    for (int qb = 0; qb < 4; ++qb)
                        acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[qa], b[qb], acc[qa][qb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    if (idle) return;
    float* out = fcd + (size_t)sim * ((long long)nw * (nw - 1) / 2);
    for (int qa = 0; qa < 4; ++qa)
        for (int qb = 0; qb < 4; ++qb)
            for (int r = 0; r < 4; ++r) {
                const int row = ti * TT + wi * 64 + qa * 16 + 4 * (lane >> 4) + r;
                const int col = tj * TT + wj * 64 + qb * 16 + (lane & 15);
                if (row < nw && col < row) out[(long long)row * (row - 1) / 2 + col] = acc[qa][qb][r];
            }
}

hipError_t launch_fcd_z(const float* wfc, int B, int nw, int np, const double* mean, const double* inv_ssd,
                        void* zhi, unsigned short* zlo, int mode, float* fcd, hipStream_t st, int tile) {
    const int npp = (np + 63) / 64 * 64;
    const unsigned rows = (unsigned)((long long)B * nw);
    if (mode == 0)
        hipLaunchKernelGGL(k_zscore_rows<0>, dim3(rows), dim3(256), 0, st, wfc, np, npp, mean, inv_ssd, zhi, zlo);
    else if (mode == 1)
        hipLaunchKernelGGL(k_zscore_rows<1>, dim3(rows), dim3(256), 0, st, wfc, np, npp, mean, inv_ssd, zhi, zlo);
    else
        hipLaunchKernelGGL(k_zscore_rows<2>, dim3(rows), dim3(256), 0, st, wfc, np, npp, mean, inv_ssd, zhi, zlo);
    const int nt = (nw + ZT - 1) / ZT;
    const dim3 grid(nt * (nt + 1) / 2, 1, B);
    const int nt2 = (nw + 127) / 128;
    if (mode == 0 && tile == 128)
        hipLaunchKernelGGL(k_gram_z32_t128, dim3(nt2 * (nt2 + 1) / 2, 1, B), dim3(256), 0, st, (const float*)zhi,
                           nw, npp, fcd);
    else if (mode == 0)
        hipLaunchKernelGGL(k_gram_z32, grid, dim3(256), 0, st, (const float*)zhi, nw, npp, fcd);
    else if (mode == 1)
        hipLaunchKernelGGL(k_gram_zbf16<false>, grid, dim3(256), 0, st, (const unsigned short*)zhi, zlo, nw, npp,
                           fcd);
    else
        hipLaunchKernelGGL(k_gram_zbf16<true>, grid, dim3(256), 0, st, (const unsigned short*)zhi, zlo, nw, npp,
                           fcd);
    return hipGetLastError();
}

// ---- launchers ---------------------------------------------------------------
template <class T>
hipError_t launch_seg_moments(const T* bold, int B, int N, int output_ts, int n_seg,
                              const int* seg_start, const int* seg_len, double* mean,
                              double* inv_ssd, hipStream_t st) {
    const int threads = N >= 256 ? 256 : ((N + 63) / 64) * 64;
    hipLaunchKernelGGL(k_seg_moments<T>, dim3(n_seg, B), dim3(threads), 0, st, bold, N, output_ts,
                       n_seg, seg_start, seg_len, mean, inv_ssd);
    return hipGetLastError();
}

template <class T>
hipError_t launch_row_moments(const T* x, long long rows, int len, double* mean, double* inv_ssd,
                              hipStream_t st, int stride, int offset) {
    if (stride < 0) stride = len;
    hipLaunchKernelGGL(k_row_moments<T>, dim3((unsigned)rows), dim3(256), 0, st, x, len, stride, offset, mean,
                       inv_ssd);
    return hipGetLastError();
}

template <class Tin, class Tc, class Tout>
hipError_t launch_gram_fc(const Tin* bold, int B, int N, int output_ts, int n_seg,
                          const int* seg_start, const int* seg_len, const double* mean,
                          const double* inv_ssd, const int* pidx, int np, Tout* fc, Tout* wfc,
                          hipStream_t st) {
    const int nt = (N + TM - 1) / TM;
    hipLaunchKernelGGL((k_gram_fc<Tin, Tc, Tout>), dim3(nt * (nt + 1) / 2, n_seg, B), dim3(256), 0,
                       st, bold, N, output_ts, n_seg, seg_start, seg_len, mean, inv_ssd, pidx, np,
                       fc, wfc);
    return hipGetLastError();
}

template <class Tin, class Tc, class Tout>
hipError_t launch_gram_fcd(const Tin* wfc, int B, int nw, int np, const double* mean,
                           const double* inv_ssd, Tout* fcd, hipStream_t st, int col0, int K) {
    if (K < 0) K = np;
    const int nt = (nw + TM - 1) / TM;
    hipLaunchKernelGGL((k_gram_fcd<Tin, Tc, Tout>), dim3(nt * (nt + 1) / 2, 1, B), dim3(256), 0, st,
                       wfc, nw, np, col0, K, mean, inv_ssd, fcd);
    return hipGetLastError();
}

template <class Tin>
hipError_t launch_gram_fc_bf16(const Tin* bold, int B, int N, int output_ts, int n_seg,
                               const int* seg_start, const int* seg_len, const double* mean,
                               const double* inv_ssd, const int* pidx, int np, float* fc, float* wfc,
                               bool split, hipStream_t st) {
    const int nt = (N + TM - 1) / TM;
    const dim3 grid(nt * (nt + 1) / 2, n_seg, B);
    if (split)
        hipLaunchKernelGGL((k_gram_fc_bf16<Tin, true>), grid, dim3(256), 0, st, bold, N, output_ts, n_seg,
                           seg_start, seg_len, mean, inv_ssd, pidx, np, fc, wfc);
    else
        hipLaunchKernelGGL((k_gram_fc_bf16<Tin, false>), grid, dim3(256), 0, st, bold, N, output_ts, n_seg,
                           seg_start, seg_len, mean, inv_ssd, pidx, np, fc, wfc);
    return hipGetLastError();
}
template hipError_t launch_gram_fc_bf16<float>(const float*, int, int, int, int, const int*, const int*,
                                               const double*, const double*, const int*, int, float*,
                                               float*, bool, hipStream_t);
template hipError_t launch_gram_fc_bf16<double>(const double*, int, int, int, int, const int*, const int*,
                                                const double*, const double*, const int*, int, float*,
                                                float*, bool, hipStream_t);

hipError_t launch_gram_fcd_bf16(const float* wfc, int B, int nw, int np, const double* mean,
                                const double* inv_ssd, float* fcd, bool split, hipStream_t st, int col0, int K) {
    if (K < 0) K = np;
    const int nt = (nw + TM - 1) / TM;
    const dim3 grid(nt * (nt + 1) / 2, 1, B);
    if (split)
        hipLaunchKernelGGL(k_gram_fcd_bf16<true>, grid, dim3(256), 0, st, wfc, nw, np, col0, K, mean, inv_ssd, fcd);
    else
        hipLaunchKernelGGL(k_gram_fcd_bf16<false>, grid, dim3(256), 0, st, wfc, nw, np, col0, K, mean, inv_ssd, fcd);
    return hipGetLastError();
}

#define BNM_INST_MOM(T)                                                                           \
    template hipError_t launch_seg_moments<T>(const T*, int, int, int, int, const int*, const int*, \
                                              double*, double*, hipStream_t);                     \
    template hipError_t launch_row_moments<T>(const T*, long long, int, double*, double*, hipStream_t, int, int);
BNM_INST_MOM(float)
BNM_INST_MOM(double)

#define BNM_INST_FC(Tin, Tc)                                                                      \
    template hipError_t launch_gram_fc<Tin, Tc, Tc>(const Tin*, int, int, int, int, const int*,   \
                                                    const int*, const double*, const double*,     \
                                                    const int*, int, Tc*, Tc*, hipStream_t);
BNM_INST_FC(double, double)
BNM_INST_FC(double, float)
BNM_INST_FC(float, double)
BNM_INST_FC(float, float)

#define BNM_INST_FCD(Tc)                                                                          \
    template hipError_t launch_gram_fcd<Tc, Tc, Tc>(const Tc*, int, int, int, const double*,      \
                                                    const double*, Tc*, hipStream_t, int, int);
BNM_INST_FCD(double)
BNM_INST_FCD(float)

}  // namespace bnm
