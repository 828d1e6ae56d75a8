// gof_kernels.hip -- goodness of fit and FIC penalty on the device
// (GOF block of bnm.cu:726-769; check_fit_bold of check_fit.cpp:138-217;
// fic_penalty of bnm.cpp:390-405).
//
// cost = 1 - r(emp FC, sim FC) [+ |mean emp FC - mean sim FC|] + KS(emp FCD,
// sim FCD), or max_cost when a simulated FC/FCD value is NaN.
//
// Two-sample KS without sorting the simulated values: the empirical sample is
// sorted and de-duplicated once (breakpoints u_0 < ... < u_{m-1} with
// empirical CDF Fe_i).  Each simulated value v falls in bin 2i+1 when v == u_i
// or bin 2i when u_{i-1} < v < u_i (bin 2m: v > u_{m-1}).  With C_b the
// cumulative count through bin b, the simulated CDF at u_i is C_{2i+1}/n and
// its largest value inside (u_i, u_{i+1}) is C_{2i+2}/n, while the empirical
// CDF there is Fe_i; so D = max_b |C_b/n - Fe_{(b-1)/2}| (Fe_{-1} = 0), which
// equals sup_x |F_sim(x) - F_emp(x)| exactly.
#include "kernels.hpp"

namespace bnm {

namespace {

__device__ double block_reduce(double v, double* red, bool is_max) {
    for (int o = 32; o > 0; o >>= 1) {
        const double u = __shfl_down(v, o, 64);
        v = is_max ? fmax(v, u) : v + u;
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    double t = is_max ? 0.0 : 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t = is_max ? fmax(t, red[i]) : t + red[i];
    return t;
}

}  // namespace

__global__ void k_fill_i32(int* p, long long n, int v) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

template <class T>
__global__ __launch_bounds__(256) void k_ks_bin(const T* __restrict__ fcd, GofArgs g) {
    const int sim = blockIdx.y;
    const T* x = fcd + (size_t)sim * g.nwp;
    int* bins = g.bins + (size_t)sim * (2 * g.m + 1);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < g.nwp;
         i += (long long)gridDim.x * blockDim.x) {
        const double v = (double)x[i];
        if (isnan(v)) {
            atomicOr(&g.nan_flag[sim], 1);
            continue;
        }
        int lo = 0, hi = g.m;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (g.ks_u[mid] < v) lo = mid + 1; else hi = mid;
        }
        const int bin = (lo < g.m && g.ks_u[lo] == v) ? 2 * lo + 1 : 2 * lo;
        atomicAdd(&bins[bin], 1);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_gof(const T* __restrict__ fc, GofArgs g) {
    __shared__ double red[16];
    __shared__ long long scan[256];
    const int sim = blockIdx.x, tid = threadIdx.x;
    // FIC penalty from the mean r_E of each node (bnm.cu:752-767)
    if (g.calc_penalty) {
        double p = 0;
        const double* rE = g.ext + ((size_t)sim * 6 + 2) * g.N;
        for (int j = tid; j < g.N; j += blockDim.x) {
            const double d = fabs(rE[j] - 3);
            if (d > 1) p += 1 - exp(-0.05 * (d - 1));
        }
        p = block_reduce(p, red, false);
        p *= g.fic_scale;
        p /= g.N;
        if (g.fic_reject && g.fic_failed[sim]) p += 1;
        if (tid == 0) g.fic_pen[sim] = p;
    }
    if (g.sim_only) {
        if (tid == 0) g.corr[sim] = g.max_cost;
        return;
    }
    // FC: NaN check, mean, Pearson against the (pre-centred) empirical FC
    const T* f = fc + (size_t)sim * g.np;
    double s = 0, nan = 0;
    for (int i = tid; i < g.np; i += blockDim.x) {
        const double v = (double)f[i];
        if (isnan(v)) nan = 1;
        s += v;
    }
    const double any_nan = block_reduce(nan, red, true);
    const double mean_s = block_reduce(s, red, false) / g.np;
    double cross = 0, ss = 0;
    for (int i = tid; i < g.np; i += blockDim.x) {
        const double d = (double)f[i] - mean_s;
        cross += g.emp_fc_c[i] * d;
        ss += d * d;
    }
    cross = block_reduce(cross, red, false);
    ss = block_reduce(ss, red, false);
    double cost = 1 - cross / (g.emp_fc_ssd * sqrt(ss));
    if (g.use_fc_diff) cost += fabs(g.emp_fc_mean - mean_s);
    bool bad = any_nan != 0;
    if (!g.no_fcd) {
        bad = bad || g.nan_flag[sim] != 0;
        // KS scan over the 2m+1 bins: contiguous chunk per thread
        const long long nb = 2LL * g.m + 1;
        const long long chunk = (nb + blockDim.x - 1) / blockDim.x;
        const long long b0 = tid * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
        const int* bins = g.bins + (size_t)sim * nb;
        long long local = 0;
        for (long long b = b0; b < b1; ++b) local += bins[b];
        scan[tid] = local;
        __syncthreads();
        for (int o = 1; o < (int)blockDim.x; o <<= 1) {   // inclusive Hillis-Steele scan
            const long long v = tid >= o ? scan[tid - o] : 0;
            __syncthreads();
            scan[tid] += v;
            __syncthreads();
        }
        long long run = scan[tid] - local;
        const double inv_n = 1.0 / (double)g.nwp;
        double d = 0;
        for (long long b = b0; b < b1; ++b) {
            run += bins[b];
            const double fe = b == 0 ? 0.0 : g.ks_fe[(b - 1) >> 1];
            d = fmax(d, fabs((double)run * inv_n - fe));
        }
        d = block_reduce(d, red, true);
        cost += d;
    }
    if (tid == 0) g.corr[sim] = bad ? g.max_cost : cost;
}

hipError_t launch_fill_i32(int* p, long long n, int v, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    long long blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

template <class T>
hipError_t launch_ks_bin(const T* fcd, const GofArgs& g, hipStream_t st) {
    long long blocks = (g.nwp + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_ks_bin<T>, dim3((unsigned)blocks, g.B), dim3(256), 0, st, fcd, g);
    return hipGetLastError();
}

template <class T>
hipError_t launch_gof(const T* fc, const GofArgs& g, hipStream_t st) {
    hipLaunchKernelGGL(k_gof<T>, dim3(g.B), dim3(256), 0, st, fc, g);
    return hipGetLastError();
}

template hipError_t launch_ks_bin<float>(const float*, const GofArgs&, hipStream_t);
template hipError_t launch_ks_bin<double>(const double*, const GofArgs&, hipStream_t);
template hipError_t launch_gof<float>(const float*, const GofArgs&, hipStream_t);
template hipError_t launch_gof<double>(const double*, const GofArgs&, hipStream_t);

}  // namespace bnm
