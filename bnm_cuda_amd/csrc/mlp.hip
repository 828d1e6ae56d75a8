// mlp.hip -- a small multilayer perceptron trained on the GPU: dense layers
// with GELU, mean-squared-error loss, Adam.  Used as the learned FIC
// initialiser (per-node wIE from G, wEE, wEI, SC strength, map value) and as
// the fit surrogate (per-simulation FC r and FCD KS from the parameters).  No
// counterpart in the reference.
//
// All matrix products go through one 64x64-tiled GEMM kernel on MFMA:
// BF16 operands (v_mfma_f32_16x16x32_bf16) or FP32 operands
// (v_mfma_f32_16x16x4_f32), FP32 accumulation, FP32 parameters.  Epilogues:
//   forward  Z = X W^T + b;  A = gelu(Z)          (hidden layers)
//   backward dZ_prev = (dZ W) * gelu'(Z_prev)
//   weights  dW = dZ^T X  (split over the batch, partial sums reduced in a
//            fixed order), db = column sums of dZ
// gelu(z) = z/2 (1 + erf(z/sqrt 2)).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/bnm_hip.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// gfx950's v_cvt_pk_bf16_f32: round to nearest even, the same bits as
// adding 0x7fff + lsb to the FP32 pattern for finite inputs
__device__ __forceinline__ unsigned short f2bf(float x) { return __builtin_bit_cast(unsigned short, (__bf16)x); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.3989422804014327f * expf(-0.5f * z * z);
}

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

enum Epi { EPI_STORE = 0, EPI_BIAS_GELU = 1, EPI_BIAS = 2, EPI_GELU_GRAD = 3 };

// C[M x N] (+)= op(A)[M x K] . op(B)[K x N]; element (m, k) of op(A) at
// A[m*sam + k*sak], element (k, n) of op(B) at B[k*sbk + n*sbn].  The K range
// of this workgroup is [blockIdx.z*klen, min(K, (blockIdx.z+1)*klen)); its
// output goes to C + blockIdx.z*cz.
template <bool BF16, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const float* __restrict__ A, long sam, long sak,
                                              const float* __restrict__ B, long sbk, long sbn,
                                              float* __restrict__ C, long ldc, long cz, int M, int N,
                                              int K, int klen, const float* __restrict__ bias,
                                              float* __restrict__ Zout, const float* __restrict__ Zin) {
    __shared__ float As32[BF16 ? 1 : BM * LDK], Bs32[BF16 ? 1 : BN * LDK];
    __shared__ unsigned short As16[BF16 ? BM * LDK : 1], Bs16[BF16 ? BN * LDK : 1];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kb = blockIdx.z * klen, ke = min(K, kb + klen);
    f32x4 acc[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0, 0, 0, 0};
    for (int k0 = kb; k0 < ke; k0 += BK) {
        // stage a BM x BK tile of op(A) and a BN x BK tile of op(B)^T, row = m or n, col = k
        for (int e = tid; e < BM * BK; e += 256) {
            const int r = e / BK, kk = e % BK, k = k0 + kk;
            const float va = (m0 + r < M && k < ke) ? A[(long)(m0 + r) * sam + (long)k * sak] : 0.0f;
            const float vb = (n0 + r < N && k < ke) ? B[(long)k * sbk + (long)(n0 + r) * sbn] : 0.0f;
            if constexpr (BF16) {
                As16[r * LDK + kk] = f2bf(va);
                Bs16[r * LDK + kk] = f2bf(vb);
            } else {
                As32[r * LDK + kk] = va;
                Bs32[r * LDK + kk] = vb;
            }
        }
        __syncthreads();
        if constexpr (BF16) {
            u16x8 a[2], b[2];
            for (int q = 0; q < 2; ++q) {
                a[q] = *reinterpret_cast<const u16x8*>(&As16[(wm * 32 + q * 16 + (lane & 15)) * LDK + 8 * (lane >> 4)]);
                b[q] = *reinterpret_cast<const u16x8*>(&Bs16[(wn * 32 + q * 16 + (lane & 15)) * LDK + 8 * (lane >> 4)]);
            }
            for (int qa = 0; qa < 2; ++qa)
                for (int qb = 0; qb < 2; ++qb)
                    acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(bf16x8, a[qa]), __builtin_bit_cast(bf16x8, b[qb]), acc[qa][qb], 0, 0, 0);
        } else {
            for (int s = 0; s < BK / 4; ++s) {
                float a[2], b[2];
                for (int q = 0; q < 2; ++q) {
                    a[q] = As32[(wm * 32 + q * 16 + (lane & 15)) * LDK + 4 * s + (lane >> 4)];
                    b[q] = Bs32[(wn * 32 + q * 16 + (lane & 15)) * LDK + 4 * s + (lane >> 4)];
                }
                for (int qa = 0; qa < 2; ++qa)
                    for (int qb = 0; qb < 2; ++qb)
                        acc[qa][qb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[qa], b[qb], acc[qa][qb], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    float* Cz = C + (long)blockIdx.z * cz;
    for (int qa = 0; qa < 2; ++qa)
        for (int qb = 0; qb < 2; ++qb)
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm * 32 + qa * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn * 32 + qb * 16 + (lane & 15);
                if (m >= M || n >= N) continue;
                float v = acc[qa][qb][r];
                const long o = (long)m * ldc + n;
                if constexpr (EPI == EPI_BIAS_GELU) {
                    v += bias[n];
                    Zout[o] = v;
                    v = gelu(v);
                } else if constexpr (EPI == EPI_BIAS) {
                    v += bias[n];
                } else if constexpr (EPI == EPI_GELU_GRAD) {
                    v *= gelu_grad(Zin[o]);
                }
                Cz[o] = v;
            }
}

// out[i] = sum_z part[z*n + i], fixed order
__global__ void k_sum_parts(const float* __restrict__ part, int nz, long n, float* __restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        float s = 0;
        for (int z = 0; z < nz; ++z) s += part[(long)z * n + i];
        out[i] = s;
    }
}

// db[n] = sum_m dZ[m][n] (one workgroup per column block, fixed order)
__global__ void k_colsum(const float* __restrict__ dZ, int M, int N, float* __restrict__ db) {
    __shared__ float red[256];
    const int n = blockIdx.x;
    float s = 0;
    for (int m = threadIdx.x; m < M; m += blockDim.x) s += dZ[(long)m * N + n];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[n] = red[0];
}

// dY = 2 (Y - T) / (M*N) elementwise (gradient of the mean squared error);
// loss partial sums per workgroup
// loss_part[blockIdx.x] is overwritten, or accumulated into when `accum`
__global__ void k_mse(const float* __restrict__ Y, const float* __restrict__ T, long n, float scale,
                      float* __restrict__ dY, float* __restrict__ loss_part, int accum) {
    __shared__ float red[256];
    float s = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float d = Y[i] - T[i];
        dY[i] = 2.0f * d * scale;
        s += d * d;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * scale + (accum ? loss_part[blockIdx.x] : 0.0f);
}

__global__ void k_fill_f32(float* __restrict__ p, int n, float v) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1 - b1) * gi;
        const float b = b2 * m2[i] + (1 - b2) * gi * gi;
        m1[i] = a;
        m2[i] = b;
        p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
    }
}

// gather rows idx[0..n) of X [*, d] into out [n, d]
__global__ void k_gather(const float* __restrict__ X, const int* __restrict__ idx, int n, int d,
                         float* __restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < (long)n * d; i += (long)gridDim.x * blockDim.x)
        out[i] = X[(long)idx[i / d] * d + i % d];
}

thread_local std::string g_mlp_err;

}  // namespace

struct bnm_mlp {
    int in = 0, hid = 0, depth = 0, out = 0, bf16 = 1;
    std::vector<int> dims;          // in, hid x depth, out
    std::vector<long> woff, boff;   // offsets into the flat parameter vector
    long nparam = 0;
    float *p = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
    int cap = 0;                    // activation rows allocated
    std::vector<float*> Z, A, dZ;   // per layer [cap x dims[l+1]] (A[0] = input)
    float *part = nullptr, *loss_part = nullptr, *target = nullptr;
    long part_cap = 0;
    float *X = nullptr, *Y = nullptr, *bx = nullptr, *by = nullptr;
    int *idx = nullptr;
    int n_data = 0;
    long step = 0;
    hipStream_t st = nullptr;
};

namespace {

int mlp_fail(const std::string& m) {
    g_mlp_err = m;
    return BNM_E_HIP;
}

#define MCHK(expr)                                                                   \
    do {                                                                             \
        hipError_t _e = (expr);                                                      \
        if (_e != hipSuccess) return mlp_fail(std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

int ensure_rows(bnm_mlp* m, int rows) {
    if (rows <= m->cap) return BNM_OK;
    for (auto* v : {&m->Z, &m->A, &m->dZ})
        for (float*& q : *v)
            if (q) { (void)hipFree(q); q = nullptr; }
    const int L = m->depth + 1;
    m->Z.assign(L, nullptr);
    m->A.assign(L + 1, nullptr);
    m->dZ.assign(L + 1, nullptr);
    for (int l = 0; l <= L; ++l) {
        MCHK(hipMalloc(&m->A[l], (size_t)rows * m->dims[l] * sizeof(float)));
        MCHK(hipMalloc(&m->dZ[l], (size_t)rows * m->dims[l] * sizeof(float)));
    }
    for (int l = 0; l < L; ++l) MCHK(hipMalloc(&m->Z[l], (size_t)rows * m->dims[l + 1] * sizeof(float)));
    if (m->target) (void)hipFree(m->target);
    MCHK(hipMalloc(&m->target, (size_t)rows * m->out * sizeof(float)));
    m->cap = rows;
    return BNM_OK;
}

template <bool BF16, int EPI>
void gemm(bnm_mlp* m, const float* A, long sam, long sak, const float* B, long sbk, long sbn, float* C,
          long ldc, long cz, int M, int N, int K, int ksplit, const float* bias, float* Zout,
          const float* Zin) {
    int klen = (K + ksplit - 1) / ksplit;
    klen = ((klen + BK - 1) / BK) * BK;
    const int kz = (K + klen - 1) / klen;
    const dim3 grid((M + BM - 1) / BM, (N + BN - 1) / BN, kz);
    hipLaunchKernelGGL((k_gemm<BF16, EPI>), grid, dim3(256), 0, m->st, A, sam, sak, B, sbk, sbn, C, ldc,
                       cz, M, N, K, klen, bias, Zout, Zin);
}

// forward over `rows` rows already in A[0]; output in A[L]
int forward(bnm_mlp* m, int rows) {
    const int L = m->depth + 1;
    for (int l = 0; l < L; ++l) {
        const int K = m->dims[l], N = m->dims[l + 1];
        const float* W = m->p + m->woff[l];
        const float* b = m->p + m->boff[l];
        const bool last = l == L - 1;
        if (m->bf16) {
            if (last) gemm<true, EPI_BIAS>(m, m->A[l], K, 1, W, 1, K, m->A[l + 1], N, 0, rows, N, K, 1, b, nullptr, nullptr);
            else gemm<true, EPI_BIAS_GELU>(m, m->A[l], K, 1, W, 1, K, m->A[l + 1], N, 0, rows, N, K, 1, b, m->Z[l], nullptr);
        } else {
            if (last) gemm<false, EPI_BIAS>(m, m->A[l], K, 1, W, 1, K, m->A[l + 1], N, 0, rows, N, K, 1, b, nullptr, nullptr);
            else gemm<false, EPI_BIAS_GELU>(m, m->A[l], K, 1, W, 1, K, m->A[l + 1], N, 0, rows, N, K, 1, b, m->Z[l], nullptr);
        }
    }
    MCHK(hipGetLastError());
    return BNM_OK;
}

constexpr int kLossBlocks = 64;

// gradient into m->g for `rows` rows with targets in m->target; the loss goes
// to m->loss_part (kLossBlocks partial sums, added to what is there when
// `accum`), and to *loss when given (synchronises)
int backward(bnm_mlp* m, int rows, float* loss, int accum = 0) {
    const int L = m->depth + 1;
    const long nout = (long)rows * m->out;
    const int lb = kLossBlocks;
    hipLaunchKernelGGL(k_mse, dim3(lb), dim3(256), 0, m->st, m->A[L], m->target, nout, 1.0f / (float)nout,
                       m->dZ[L], m->loss_part, accum);
    for (int l = L - 1; l >= 0; --l) {
        const int K = m->dims[l], N = m->dims[l + 1];
        const float* W = m->p + m->woff[l];
        // dW[N x K] = dZ_{l+1}^T [N x rows] . A_l [rows x K], split over rows
        const int ks = std::max(1, std::min(64, rows / 256));
        const long wsz = (long)N * K;
        if (m->part_cap < wsz * ks) {
            if (m->part) (void)hipFree(m->part);
            MCHK(hipMalloc(&m->part, (size_t)wsz * ks * sizeof(float)));
            m->part_cap = wsz * ks;
        }
        if (m->bf16)
            gemm<true, EPI_STORE>(m, m->dZ[l + 1], 1, N, m->A[l], K, 1, m->part, K, wsz, N, K, rows, ks, nullptr, nullptr, nullptr);
        else
            gemm<false, EPI_STORE>(m, m->dZ[l + 1], 1, N, m->A[l], K, 1, m->part, K, wsz, N, K, rows, ks, nullptr, nullptr, nullptr);
        int klen = (rows + ks - 1) / ks;
        klen = ((klen + BK - 1) / BK) * BK;
        const int kz = (rows + klen - 1) / klen;
        hipLaunchKernelGGL(k_sum_parts, dim3(256), dim3(256), 0, m->st, m->part, kz, wsz, m->g + m->woff[l]);
        hipLaunchKernelGGL(k_colsum, dim3(N), dim3(256), 0, m->st, m->dZ[l + 1], rows, N, m->g + m->boff[l]);
        if (l > 0) {
            // dZ_l = (dZ_{l+1} . W) * gelu'(Z_{l-1})
            if (m->bf16)
                gemm<true, EPI_GELU_GRAD>(m, m->dZ[l + 1], N, 1, W, K, 1, m->dZ[l], K, 0, rows, K, N, 1, nullptr, nullptr, m->Z[l - 1]);
            else
                gemm<false, EPI_GELU_GRAD>(m, m->dZ[l + 1], N, 1, W, K, 1, m->dZ[l], K, 0, rows, K, N, 1, nullptr, nullptr, m->Z[l - 1]);
        }
    }
    MCHK(hipGetLastError());
    if (loss) {
        std::vector<float> lp(lb);
        MCHK(hipMemcpyAsync(lp.data(), m->loss_part, lb * sizeof(float), hipMemcpyDeviceToHost, m->st));
        MCHK(hipStreamSynchronize(m->st));
        double s = 0;
        for (float v : lp) s += v;
        *loss = (float)s;
    }
    return BNM_OK;
}

}  // namespace

extern "C" {

const char* bnm_mlp_last_error(void) { return g_mlp_err.c_str(); }

bnm_mlp* bnm_mlp_create(int in_dim, int hidden, int depth, int out_dim, int bf16, unsigned seed,
                        int* status) {
    auto set = [&](int s) { if (status) *status = s; };
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_mlp_err = "no HIP device";
        set(BNM_E_NODEV);
        return nullptr;
    }
    if (in_dim < 1 || hidden < 1 || depth < 1 || out_dim < 1) {
        g_mlp_err = "bad MLP shape";
        set(BNM_E_ARG);
        return nullptr;
    }
    auto* m = new bnm_mlp;
    m->in = in_dim; m->hid = hidden; m->depth = depth; m->out = out_dim; m->bf16 = bf16 != 0;
    m->dims.push_back(in_dim);
    for (int i = 0; i < depth; ++i) m->dims.push_back(hidden);
    m->dims.push_back(out_dim);
    long off = 0;
    for (int l = 0; l <= depth; ++l) {
        m->woff.push_back(off);
        off += (long)m->dims[l] * m->dims[l + 1];
        m->boff.push_back(off);
        off += m->dims[l + 1];
    }
    m->nparam = off;
    // He-style uniform init in [-sqrt(6/fan_in), sqrt(6/fan_in)] (weights), zero biases
    std::vector<float> h(off, 0.0f);
    std::mt19937 gen(seed);
    for (int l = 0; l <= depth; ++l) {
        const float a = std::sqrt(6.0f / m->dims[l]) * (l == depth ? 0.5f : 1.0f);
        std::uniform_real_distribution<float> u(-a, a);
        for (long i = 0; i < (long)m->dims[l] * m->dims[l + 1]; ++i) h[m->woff[l] + i] = u(gen);
    }
    bool ok = hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&m->p, off * sizeof(float)) == hipSuccess &&
              hipMalloc(&m->g, off * sizeof(float)) == hipSuccess &&
              hipMalloc(&m->m1, off * sizeof(float)) == hipSuccess &&
              hipMalloc(&m->m2, off * sizeof(float)) == hipSuccess &&
              hipMalloc(&m->loss_part, 64 * sizeof(float)) == hipSuccess &&
              hipMemcpy(m->p, h.data(), off * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
              hipMemset(m->m1, 0, off * sizeof(float)) == hipSuccess &&
              hipMemset(m->m2, 0, off * sizeof(float)) == hipSuccess;
    if (!ok) {
        g_mlp_err = "MLP allocation failed";
        set(BNM_E_HIP);
        bnm_mlp_destroy(m);
        return nullptr;
    }
    set(BNM_OK);
    return m;
}

void bnm_mlp_destroy(bnm_mlp* m) {
    if (!m) return;
    for (float* q : {m->p, m->g, m->m1, m->m2, m->part, m->loss_part, m->target, m->X, m->Y, m->bx, m->by})
        if (q) (void)hipFree(q);
    if (m->idx) (void)hipFree(m->idx);
    for (auto* v : {&m->Z, &m->A, &m->dZ})
        for (float* q : *v)
            if (q) (void)hipFree(q);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

long bnm_mlp_num_params(bnm_mlp* m) { return m ? m->nparam : -1; }

int bnm_mlp_get_params(bnm_mlp* m, float* out) {
    if (!m || !out) return BNM_E_ARG;
    MCHK(hipStreamSynchronize(m->st));
    MCHK(hipMemcpy(out, m->p, m->nparam * sizeof(float), hipMemcpyDeviceToHost));
    return BNM_OK;
}

int bnm_mlp_set_params(bnm_mlp* m, const float* in) {
    if (!m || !in) return BNM_E_ARG;
    MCHK(hipStreamSynchronize(m->st));
    MCHK(hipMemcpy(m->p, in, m->nparam * sizeof(float), hipMemcpyHostToDevice));
    return BNM_OK;
}

// Y[n x out] = f(X[n x in])
int bnm_mlp_predict(bnm_mlp* m, const float* X, int n, float* Y) {
    if (!m || !X || !Y || n < 1) return BNM_E_ARG;
    const int chunk = 65536;
    for (int s = 0; s < n; s += chunk) {
        const int r = std::min(chunk, n - s);
        int e = ensure_rows(m, r);
        if (e) return e;
        MCHK(hipMemcpyAsync(m->A[0], X + (long)s * m->in, (size_t)r * m->in * sizeof(float), hipMemcpyHostToDevice, m->st));
        if ((e = forward(m, r))) return e;
        MCHK(hipMemcpyAsync(Y + (long)s * m->out, m->A[m->depth + 1], (size_t)r * m->out * sizeof(float),
                            hipMemcpyDeviceToHost, m->st));
        MCHK(hipStreamSynchronize(m->st));
    }
    return BNM_OK;
}

// Loss (mean squared error over n x out) and its gradient w.r.t. the flat
// parameter vector at the current parameters.
int bnm_mlp_loss_grad(bnm_mlp* m, const float* X, const float* T, int n, float* loss, float* grad) {
    if (!m || !X || !T || n < 1) return BNM_E_ARG;
    int e = ensure_rows(m, n);
    if (e) return e;
    MCHK(hipMemcpyAsync(m->A[0], X, (size_t)n * m->in * sizeof(float), hipMemcpyHostToDevice, m->st));
    MCHK(hipMemcpyAsync(m->target, T, (size_t)n * m->out * sizeof(float), hipMemcpyHostToDevice, m->st));
    if ((e = forward(m, n))) return e;
    if ((e = backward(m, n, loss))) return e;
    if (grad) MCHK(hipMemcpy(grad, m->g, m->nparam * sizeof(float), hipMemcpyDeviceToHost));
    MCHK(hipStreamSynchronize(m->st));
    return BNM_OK;
}

// Uploads a training set (X [n x in], T [n x out]) to the device.
int bnm_mlp_set_data(bnm_mlp* m, const float* X, const float* T, int n) {
    if (!m || !X || !T || n < 1) return BNM_E_ARG;
    MCHK(hipStreamSynchronize(m->st));
    for (float** q : {&m->X, &m->Y})
        if (*q) { (void)hipFree(*q); *q = nullptr; }
    MCHK(hipMalloc(&m->X, (size_t)n * m->in * sizeof(float)));
    MCHK(hipMalloc(&m->Y, (size_t)n * m->out * sizeof(float)));
    MCHK(hipMemcpy(m->X, X, (size_t)n * m->in * sizeof(float), hipMemcpyHostToDevice));
    MCHK(hipMemcpy(m->Y, T, (size_t)n * m->out * sizeof(float), hipMemcpyHostToDevice));
    m->n_data = n;
    return BNM_OK;
}

// Adam on shuffled mini-batches of the uploaded set; loss[e] = mean batch loss
// of epoch e.
int bnm_mlp_train(bnm_mlp* m, int epochs, int batch, float lr, unsigned seed, float* loss) {
    if (!m || m->n_data < 1 || batch < 1 || epochs < 1) return BNM_E_ARG;
    batch = std::min(batch, m->n_data);
    int e = ensure_rows(m, batch);
    if (e) return e;
    if (m->idx) { (void)hipFree(m->idx); m->idx = nullptr; }
    MCHK(hipMalloc(&m->idx, (size_t)m->n_data * sizeof(int)));
    std::mt19937 gen(seed);
    std::vector<int> perm(m->n_data);
    for (int i = 0; i < m->n_data; ++i) perm[i] = i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int nb = m->n_data / batch;
    float lp[kLossBlocks];
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), gen);
        // the previous epoch's kernels read m->idx: the synchronous copy waits for them
        MCHK(hipStreamSynchronize(m->st));
        MCHK(hipMemcpy(m->idx, perm.data(), (size_t)m->n_data * sizeof(int), hipMemcpyHostToDevice));
        hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, m->st, m->loss_part, kLossBlocks, 0.0f);
        for (int b = 0; b < nb; ++b) {
            const int* ib = m->idx + (long)b * batch;
            hipLaunchKernelGGL(k_gather, dim3(256), dim3(256), 0, m->st, m->X, ib, batch, m->in, m->A[0]);
            hipLaunchKernelGGL(k_gather, dim3(256), dim3(256), 0, m->st, m->Y, ib, batch, m->out, m->target);
            if ((e = forward(m, batch))) return e;
            if ((e = backward(m, batch, nullptr, 1))) return e;
            ++m->step;
            const float c1 = 1 - std::pow(b1, (float)m->step), c2 = 1 - std::pow(b2, (float)m->step);
            hipLaunchKernelGGL(k_adam, dim3(256), dim3(256), 0, m->st, m->p, m->g, m->m1, m->m2, m->nparam, lr, b1,
                               b2, eps, c1, c2);
        }
        MCHK(hipGetLastError());
        if (loss) {
            MCHK(hipMemcpyAsync(lp, m->loss_part, sizeof(lp), hipMemcpyDeviceToHost, m->st));
            MCHK(hipStreamSynchronize(m->st));
            double ls = 0;
            for (float v : lp) ls += v;
            loss[ep] = (float)(ls / nb);
        }
    }
    MCHK(hipStreamSynchronize(m->st));
    return BNM_OK;
}

}  // extern "C"
