// api.hip -- the C ABI of include/bnm_hip.h: config, index tables, noise, and
// the persistent context that runs the whole path on one GPU.
//
// Host-side pieces restate the reference driver: env defaults (constants.h:93-116,
// helpers.cpp:4-20), output_ts / pair / window tables (run_CMAES.cpp:476-594),
// the noise array (run_CMAES.cpp:640-657).  All compute runs on the device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/bnm_hip.h"
#include "kernels.hpp"

using namespace bnm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(BNM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));      \
    } while (0)

// get_env_or_default of helpers.cpp:4-20: the int overload parses with atoi,
// the double overload with atof and returns float.
int env_int(const char* k, int dflt) {
    const char* v = std::getenv(k);
    return v ? std::atoi(v) : dflt;
}
float env_float(const char* k, double dflt) {
    const char* v = std::getenv(k);
    return v ? (float)std::atof(v) : (float)dflt;
}

template <class T>
int dalloc(T** p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess) return fail(BNM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return BNM_OK;
}

}  // namespace

// One launchable integration kernel shape.
struct Variant {
    int kind;   // 0 split (SC in registers), 1 stream (SC from L2), 2 mfma, 3 plain,
                // 4 split with the E/I pools on separate lanes (FP32; mode = form:
                //   0 as written, 1 S_E by DPP row broadcast, 2 / 3 short-chain
                //   update with four / one partial sums),
                // 6 FP64 split with the E/I pools on separate lanes,
                // 7 FP64 stream with register/LDS-resident SC,
                // 5 stream with register/LDS-resident SC (FP32; mode 1: S_E by
                //   DPP row broadcast, 2: and three streamed chunks in flight)
    int ls, ks, mode;
};

// Scratch of the no-allocation entry point: one static device array per GPU
// (each device the library runs on holds its own copy until the process
// ends).  Host code never uses the array's name as an address: the address is
// taken once per device with hipGetSymbolAddress (scratch_base below).
static __device__ __align__(256) unsigned char scratch[BNM_NOALLOC_SCRATCH_BYTES];

struct bnm_ctx {
    bnm_config cfg{};
    int N = 0, B = 0, ots = 0, np = 0, nw = 0, nseg = 0, m = 0;
    long long nwp = 0;
    bool f64 = true;      // integration precision
    bool fc64 = true;     // Gram precision
    bool fcbf16 = false;  // BF16 Gram operands (FP32 accumulation and storage)
    bool fcsplit = false; // ... as hi/lo pairs (three products)
    bool fcd_direct = false;
    int fcd_tile = 64;
    bool have_bold = false;  // BOLD holds a run's or a caller's series
    hipStream_t st = nullptr;
    double* SC64 = nullptr;
    float* SC32 = nullptr;
    double* SCt64 = nullptr;  // transposed copies for the streamed large-N kernel
    float* SCt32 = nullptr;
    float* SCc32 = nullptr;   // chunked [k/16][j][k%16] copy for the hybrid kernel
    double* SCc64 = nullptr;  // chunked [k/8][j][k%8] copy for the FP64 hybrid kernel
    // sparse row lists for the sparse FP32 kernel (N <= 512), one lane per
    // node ([0]) and two (each half the row's k-range, N <= 256, [1]):
    // weights [kSparseRW + kSparseLL][LS * N], packed byte steps [that / 4]
    // [LS * N]; sp_P the slots per lane of the current SC (0: no SC yet, or
    // too dense to fit)
    float* SPw[2] = {nullptr, nullptr};
    double* SPw64[2] = {nullptr, nullptr};   // FP64 contexts: FP64 weights (8-byte steps)
    unsigned* SPd[2] = {nullptr, nullptr};
    int sp_P[2] = {0, 0};
    size_t lds_max = 64 * 1024;  // dynamic LDS a workgroup may use
    float* noise = nullptr;
    long long noise_cap = 0;
    // wIEfic holds the FIC starting wIE (the caller's or the analytical
    // one) and is never written by the integration, so every run and graph
    // replay starts from the same values; wIEadj receives the adjusted wIE
    double *G = nullptr, *wEE = nullptr, *wEI = nullptr, *wIE = nullptr, *wIEfic = nullptr,
           *wIEadj = nullptr, *wIEscratch = nullptr;
    void* bold = nullptr;
    double* ext = nullptr;
    int *failed = nullptr, *trials = nullptr, *unstable = nullptr;
    int *seg_start = nullptr, *seg_len = nullptr, *pidx = nullptr;
    double *mean = nullptr, *inv = nullptr, *wmean = nullptr, *winv = nullptr;
    void *fc = nullptr, *wfc = nullptr, *fcd = nullptr;
    // z-scored window-FC rows for the FP32 / BF16 FCD Gram ([B*nw][np rounded
    // up to 64]; zlo: BF16 low parts for the hi/lo split)
    void* wz = nullptr;
    unsigned short* wzlo = nullptr;
    // hemisphere FCDs (save_hemis): first / last n_hemi edges of each window FC
    void *fcd_h[2] = {nullptr, nullptr};
    double *wmean_h[2] = {nullptr, nullptr}, *winv_h[2] = {nullptr, nullptr};
    int n_hemi = 0;
    double *emp_c = nullptr, *ks_u = nullptr, *ks_fe = nullptr;
    double emp_mean = 0, emp_ssd = 0;
    int* bins = nullptr;
    int* nanf = nullptr;
    double *corr = nullptr, *pen = nullptr;
    hipEvent_t ev[6] = {};
    double timings[6] = {0, 0, 0, 0, 0, 0};
    std::string kname = "none";
    std::string tune_log;
    bool tuned = false;
    double ms_per_sim_ms = -1;   // integration time per simulated ms of the chosen variant (tuner)
    int autotune = 1, tune_ms = 300;
    Variant variant{0, 4, 1, 0};
    hipGraphExec_t graph_exec = nullptr;
    bool have_sc = false, have_noise = false, have_emp = false, have_params = false;
    int fast_ls = 0, fast_ks = 0, stream_ks = 0, force_stream = 0, no_ei = 0, ei_ls = 0, ei_bc = 0, ei_form = -1, no_hyb = 0, hyb_ks = 0, hyb_bc = -1, hyb_rr = -1;
    int no_sparse = 0, sparse_only = 0, sparse_mode = 64;
    int sc_gen = 0;   // bumped when a new SC changes the sparse kernel's shape
    // no-allocation mode (bnm_run_simulations_noalloc): every device buffer is
    // carved out of `arena` (the library's static scratch array), the null
    // stream is used and no events, tuner windows or graphs are created
    unsigned char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
};

// Device buffer of `bytes` for the context: from the arena (256-byte aligned
// bump allocation, BNM_E_SIZE when it does not fit) or hipMalloc.
static int ctx_alloc_bytes(bnm_ctx* x, void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 1;
    if (x->arena) {
        const size_t off = (x->arena_used + 255) & ~(size_t)255;
        if (off > x->arena_cap || bytes > x->arena_cap - off)
            return fail(BNM_E_SIZE, "configuration needs more than the " + std::to_string(x->arena_cap >> 20) +
                                        " MiB no-allocation scratch (" + std::to_string((off + bytes) >> 20) +
                                        "+ MiB so far)");
        *p = x->arena + off;
        x->arena_used = off + bytes;
        return BNM_OK;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) return fail(BNM_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return BNM_OK;
}

template <class T>
static int cdalloc(bnm_ctx* x, T** p, size_t n) {
    void* v = nullptr;
    const int r = ctx_alloc_bytes(x, &v, (n ? n : 1) * sizeof(T));
    *p = static_cast<T*>(v);
    return r;
}

extern "C" {

const char* bnm_last_error(void) { return g_err.c_str(); }

int bnm_default_config(bnm_config* c) {
    if (!c) return fail(BNM_E_ARG, "null config");
    std::memset(c, 0, sizeof(*c));
    c->nodes = 0;
    c->n_sims = 0;
    c->time_steps = 0;
    c->bold_tr = 1000;
    c->n_vols_remove = -1;
    c->window_size = 30;
    c->window_step = 5;
    c->exc_interhemispheric = env_int("BNM_EXC_INTERHEM", 1) != 0;
    c->do_fic = 1;
    c->adjust_fic = env_int("BNM_NUMERICAL_FIC", 1) != 0;
    c->max_fic_trials = env_int("BNM_MAX_FIC_TRIALS_CMAES", 10);
    c->i_sampling_start = env_int("BNM_I_SAMPLING_START", 1000);
    c->i_sampling_end = env_int("BNM_I_SAMPLING_END", 10000);
    c->init_delta = (double)env_float("BNM_INIT_DELTA", 0.02);
    c->noise_period_ms = 0;   // finalize: the whole run, time_steps + 1 ms
    c->calculate_fic_penalty = 1;
    c->sim_only = 0;
    c->no_fcd = 0;
    c->use_fc_diff = env_int("BNM_USE_FC_DIFF", 1) != 0;
    c->max_cost = (double)env_int("BNM_MAX_COST", 2);
    c->fic_penalty_scale = (double)env_float("BNM_CMAES_FIC_PENALTY_SCALE", 2.0);
    c->fic_reject_failed = env_int("BNM_CMAES_FIC_REJECT_FAILED", 0) != 0;
    c->analytical_fic = 1;
    // the reference's default build computes in double (u_real = double
    // unless USE_FLOATS, constants.h:8-22)
    c->precision = BNM_PREC_F64;
    c->fc_precision = BNM_FC_F64;
    c->kernel = BNM_KERNEL_AUTO;
    c->sims_per_block = 0;
    c->autotune = env_int("BNM_AUTOTUNE", 1) != 0;
    c->use_graph = 0;
    c->bw_friston2003 = env_int("BNM_BW_FRISTON2003", 0) != 0;
    return BNM_OK;
}

int bnm_output_ts(int time_steps, int bold_tr) {
    if (bold_tr <= 0 || time_steps < 0) return fail(BNM_E_ARG, "bad time_steps / bold_tr");
    // run_CMAES.cpp:476-477 evaluates int(T / (TR_s / 0.001)) + 1 in floating
    // point; the kernels sample every ts with ts % TR == 0, i.e. T/TR + 1 times.
    const int n = time_steps / bold_tr + 1;
    const double TR = (double)bold_tr / 1000;
    const int ref = (int)(time_steps / (TR / 0.001)) + 1;
    if (ref != n) return fail(BNM_E_UNSUP, "output_ts of the reference formula differs from T/TR+1");
    return n;
}

int bnm_n_pairs(int nodes, int exc) {
    if (nodes < 2) return fail(BNM_E_ARG, "nodes < 2");
    if (exc && (nodes % 2)) return fail(BNM_E_ARG, "exc_interhemispheric needs an even node count");
    int n = nodes * (nodes - 1) / 2;
    if (exc) n -= (nodes / 2) * (nodes / 2);
    return n;
}

int bnm_finalize_config(bnm_config* c) {
    if (!c) return fail(BNM_E_ARG, "null config");
    if (c->nodes < 2 || c->nodes > 1024) return fail(BNM_E_ARG, "nodes must be in [2, 1024]");
    if (c->n_sims < 1) return fail(BNM_E_ARG, "n_sims < 1");
    if (c->bold_tr <= 0 || c->time_steps < 0) return fail(BNM_E_ARG, "bad time_steps / bold_tr");
    if (c->n_vols_remove < 0) c->n_vols_remove = env_int("BNM_BOLD_REMOVE", 30) * 1000 / c->bold_tr;
    if (c->noise_period_ms == 0) c->noise_period_ms = (long long)c->time_steps + 1;
    if (c->noise_period_ms < 1) return fail(BNM_E_ARG, "noise_period_ms < 1");
    if (c->window_step < 1 || c->window_size < 1) return fail(BNM_E_ARG, "bad FCD window");
    if (c->i_sampling_end < c->i_sampling_start) return fail(BNM_E_ARG, "bad FIC sampling window");
    const int ots = bnm_output_ts(c->time_steps, c->bold_tr);
    if (ots < 0) return ots;
    if (!c->sim_only) {
        if (ots - c->n_vols_remove < 2) return fail(BNM_E_ARG, "fewer than 2 volumes left for FC");
        const int np = bnm_n_pairs(c->nodes, c->exc_interhemispheric);
        if (np < 0) return np;
        if (!c->no_fcd && bnm_n_windows(c) < 2) return fail(BNM_E_ARG, "fewer than 2 FCD windows");
    }
    return BNM_OK;
}

int bnm_window_bounds(const bnm_config* c, int* starts, int* ends) {
    const int ots = c->time_steps / c->bold_tr + 1;
    const int corr_len = ots - c->n_vols_remove;
    const int first = c->window_size / 2 + c->n_vols_remove;
    const int last = corr_len - 1 - c->window_size / 2 + c->n_vols_remove;
    int n = 0;
    for (int ctr = first; ctr <= last; ctr += c->window_step) {
        int s = ctr - c->window_size / 2, e = ctr + c->window_size / 2;
        if (s < 0) s = 0;
        if (e >= ots) e = ots - 1;
        if (starts) { starts[n] = s; ends[n] = e; }
        ++n;
    }
    return n;
}

int bnm_n_windows(const bnm_config* c) {
    if (!c || c->bold_tr <= 0 || c->window_step < 1) return fail(BNM_E_ARG, "bad config");
    return bnm_window_bounds(c, nullptr, nullptr);
}

int bnm_regional_params(int nodes, int n_maps, const double* maps, double base,
                        const double* scales, double* out) {
    if (nodes < 1 || n_maps < 0 || !out || (n_maps > 0 && (!maps || !scales)))
        return fail(BNM_E_ARG, "bad regional parameter arguments");
    double mn = 1000;
    for (int j = 0; j < nodes; ++j) {
        double scaler = 1;
        for (int m = 0; m < n_maps; ++m) scaler += scales[m] * maps[(size_t)m * nodes + j];
        out[j] = base * scaler;
        if (out[j] < mn) mn = out[j];
    }
    if (mn < 0.001)
        for (int j = 0; j < nodes; ++j) out[j] -= mn - 0.001;
    return BNM_OK;
}

int bnm_make_noise(unsigned seed, long long count, float* out) {
    if (!out || count < 0) return fail(BNM_E_ARG, "bad noise buffer");
    std::mt19937 gen(seed);
    std::normal_distribution<float> nd(0, 1);
    for (long long i = 0; i < count; ++i) out[i] = nd(gen);
    return BNM_OK;
}

int bnm_sparse_layout(int nodes, const double* SC, int lanes_per_node, int fp64, int* slots, double* W,
                      unsigned* D) {
    if (nodes < 1 || !SC || !slots || (lanes_per_node != 1 && lanes_per_node != 2))
        return fail(BNM_E_ARG, "bnm_sparse_layout: bad arguments");
    const int N = nodes, LS = lanes_per_node;
    *slots = 0;
    if (N > 512 / LS) return BNM_OK;
    const size_t M = (size_t)LS * N;
    std::vector<unsigned> d;
    int P = 0;
    if (fp64) {
        std::vector<double> w;
        P = build_sparse_sc(SC, N, LS, w, d);
        if (P <= 0 || P > kSparseRW64) return BNM_OK;
        if (W) std::copy(w.begin(), w.begin() + (size_t)P * M, W);
    } else {
        std::vector<float> f((size_t)N * N), w;
        for (size_t i = 0; i < f.size(); ++i) f[i] = (float)SC[i];
        P = build_sparse_sc(f.data(), N, LS, w, d);
        if (P <= 0 || P > kSparseRW + kSparseLL) return BNM_OK;
        if (W) for (size_t i = 0; i < (size_t)P * M; ++i) W[i] = (double)w[i];
    }
    if (D) std::copy(d.begin(), d.begin() + (size_t)P / 4 * M, D);
    *slots = P;
    return BNM_OK;
}

int bnm_make_noise_chunked(unsigned seed, long long count, float* out, int threads) {
    if (!out || count < 0) return fail(BNM_E_ARG, "bad noise buffer");
    constexpr long long kChunk = 1LL << 20;
    const long long nchunks = (count + kChunk - 1) / kChunk;
    int nt = threads > 0 ? threads : (int)std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > nchunks) nt = (int)(nchunks > 0 ? nchunks : 1);
    auto work = [&](int t) {
        for (long long c = t; c < nchunks; c += nt) {
            std::seed_seq sq{seed, (unsigned)(c & 0xffffffffu), (unsigned)(c >> 32)};
            std::mt19937 gen(sq);
            std::normal_distribution<float> nd(0, 1);
            const long long b = c * kChunk, e = std::min(count, b + kChunk);
            for (long long i = b; i < e; ++i) out[i] = nd(gen);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto& th : pool) th.join();
    return BNM_OK;
}

void bnm_ctx_destroy(bnm_ctx* x) {
    if (!x) return;
    void* bufs[] = {x->SC64, x->SC32, x->SCt64, x->SCt32, x->SCc32, x->SCc64, x->SPw[0], x->SPw[1], x->SPw64[0], x->SPw64[1], x->SPd[0], x->SPd[1], x->noise, x->G, x->wEE, x->wEI, x->wIE, x->wIEfic,
                    x->wIEadj, x->wIEscratch, x->bold, x->ext, x->failed, x->trials, x->unstable,
                    x->seg_start, x->seg_len, x->pidx, x->mean, x->inv, x->wmean, x->winv, x->fc,
                    x->wfc, x->wz, x->wzlo, x->fcd, x->emp_c, x->ks_u, x->ks_fe, x->bins, x->nanf, x->corr, x->pen,
                    x->fcd_h[0], x->fcd_h[1], x->wmean_h[0], x->wmean_h[1], x->winv_h[0], x->winv_h[1]};
    if (!x->arena)
        for (void* p : bufs)
            if (p) (void)hipFree(p);
    for (auto& e : x->ev)
        if (e) (void)hipEventDestroy(e);
    if (x->graph_exec) (void)hipGraphExecDestroy(x->graph_exec);
    if (x->st) (void)hipStreamDestroy(x->st);
    delete x;
}

static int ctx_alloc(bnm_ctx* x) {
    const bnm_config& c = x->cfg;
    const size_t N = x->N, B = x->B;
    const size_t rsz = x->f64 ? 8 : 4, fsz = x->fc64 ? 8 : 4;
    int r;
    if ((r = cdalloc(x, &x->SC64, N * N))) return r;
    if ((r = cdalloc(x, &x->SC32, N * N))) return r;
    if ((r = cdalloc(x, &x->SCt64, N * N))) return r;
    // kSCtPad zero rows after the N rows: the hybrid kernel's padded k-ranges
    // read them without bounds checks
    if ((r = cdalloc(x, &x->SCt32, (N + kSCtPad) * N))) return r;
    HIPCHK(hipMemset(x->SCt32, 0, (N + kSCtPad) * N * sizeof(float)));
    if ((r = cdalloc(x, &x->SCc32, (N + kSCtPad) * N))) return r;
    HIPCHK(hipMemset(x->SCc32, 0, (N + kSCtPad) * N * sizeof(float)));
    if (x->f64) {
        if ((r = cdalloc(x, &x->SCc64, (N + kSCtPad) * N))) return r;
        HIPCHK(hipMemset(x->SCc64, 0, (N + kSCtPad) * N * sizeof(double)));
    }
    for (int i = 0; i < 2; ++i) {
        const size_t ls = i + 1;
        if (N > 512 / ls) continue;
        if (x->f64) {
            if ((r = cdalloc(x, &x->SPw64[i], (size_t)(kSparseRW + kSparseLL) * ls * N))) return r;
        } else {
            if ((r = cdalloc(x, &x->SPw[i], (size_t)(kSparseRW + kSparseLL) * ls * N))) return r;
        }
        if ((r = cdalloc(x, &x->SPd[i], (size_t)(kSparseRW + kSparseLL) / 4 * ls * N))) return r;
    }
    if ((r = cdalloc(x, &x->G, B))) return r;
    if ((r = cdalloc(x, &x->wEE, B * N))) return r;
    if ((r = cdalloc(x, &x->wEI, B * N))) return r;
    if ((r = cdalloc(x, &x->wIE, B * N))) return r;
    if ((r = cdalloc(x, &x->wIEfic, B * N))) return r;
    if ((r = cdalloc(x, &x->wIEadj, B * N))) return r;
    if ((r = cdalloc(x, &x->wIEscratch, B * N))) return r;
    if ((r = ctx_alloc_bytes(x, &x->bold, B * x->ots * N * rsz))) return r;
    if ((r = cdalloc(x, &x->ext, B * 6 * N))) return r;
    if ((r = cdalloc(x, &x->failed, B))) return r;
    if ((r = cdalloc(x, &x->trials, B))) return r;
    if ((r = cdalloc(x, &x->unstable, B))) return r;
    if ((r = cdalloc(x, &x->corr, B))) return r;
    if ((r = cdalloc(x, &x->pen, B))) return r;
    if (!c.sim_only) {
        std::vector<int> st(x->nseg), ln(x->nseg);
        st[0] = c.n_vols_remove;
        ln[0] = x->ots - c.n_vols_remove;
        std::vector<int> ws(x->nw > 0 ? x->nw : 1), we(x->nw > 0 ? x->nw : 1);
        bnm_window_bounds(&c, ws.data(), we.data());
        for (int w = 0; w + 1 < x->nseg; ++w) {
            st[w + 1] = ws[w];
            ln[w + 1] = we[w] - ws[w] + 1;
        }
        if ((r = cdalloc(x, &x->seg_start, x->nseg))) return r;
        if ((r = cdalloc(x, &x->seg_len, x->nseg))) return r;
        HIPCHK(hipMemcpy(x->seg_start, st.data(), x->nseg * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(x->seg_len, ln.data(), x->nseg * sizeof(int), hipMemcpyHostToDevice));
        // pair table (run_CMAES.cpp:494-524): pidx[i*N+j] = position of pair (i,j), i > j
        std::vector<int> pidx(N * N, -1);
        const int rh = x->N / 2;
        int p = 0;
        for (int i = 0; i < x->N; ++i)
            for (int j = 0; j < i; ++j) {
                if (c.exc_interhemispheric && ((i < rh) ^ (j < rh))) continue;
                pidx[(size_t)i * N + j] = p++;
            }
        if ((r = cdalloc(x, &x->pidx, N * N))) return r;
        HIPCHK(hipMemcpy(x->pidx, pidx.data(), N * N * sizeof(int), hipMemcpyHostToDevice));
        if ((r = cdalloc(x, &x->mean, B * x->nseg * N))) return r;
        if ((r = cdalloc(x, &x->inv, B * x->nseg * N))) return r;
        if ((r = ctx_alloc_bytes(x, &x->fc, B * x->np * fsz))) return r;
        if ((r = cdalloc(x, &x->emp_c, x->np))) return r;
        if (!c.no_fcd) {
            if ((r = ctx_alloc_bytes(x, &x->wfc, B * (size_t)x->nw * x->np * fsz))) return r;
            if ((r = cdalloc(x, &x->wmean, B * x->nw))) return r;
            if ((r = cdalloc(x, &x->winv, B * x->nw))) return r;
            if (!x->fc64 && !x->fcd_direct) {
                const size_t zn = B * (size_t)x->nw * ((x->np + 63) / 64 * 64);
                if ((r = ctx_alloc_bytes(x, &x->wz, zn * (x->fcbf16 ? 2 : 4)))) return r;
                if (x->fcsplit && (r = cdalloc(x, &x->wzlo, zn))) return r;
            }
            if ((r = ctx_alloc_bytes(x, &x->fcd, B * (size_t)x->nwp * fsz))) return r;
            if (c.save_hemis) {
                x->n_hemi = (x->N / 2) * (x->N / 2 - 1) / 2;
                for (int h = 0; h < 2; ++h) {
                    if ((r = ctx_alloc_bytes(x, &x->fcd_h[h], B * (size_t)x->nwp * fsz))) return r;
                    if ((r = cdalloc(x, &x->wmean_h[h], B * x->nw))) return r;
                    if ((r = cdalloc(x, &x->winv_h[h], B * x->nw))) return r;
                }
            }
            if ((r = cdalloc(x, &x->nanf, B))) return r;
        }
    }
    if (!x->arena)
        for (auto& e : x->ev) HIPCHK(hipEventCreate(&e));
    return BNM_OK;
}

// BNM_OK when the current HIP device is a gfx950, else BNM_E_NODEV
static int check_device() {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(BNM_E_NODEV, "no HIP device");
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess ||
        std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(BNM_E_NODEV, "the current HIP device is not gfx950");
    return BNM_OK;
}

// Analytical FIC for B parameter sets on the device (the reference's
// analytical_fic, fic.cpp:130-207, called per particle by calculate_params,
// run_CMAES.cpp:144-174).
int bnm_analytical_fic(int nodes, int n_sims, const double* SC, const double* G, const double* wEE,
                       const double* wEI, double* wIE_out, int* unstable) {
    if (nodes < 1 || n_sims < 1 || !SC || !G || !wEE || !wEI || !wIE_out)
        return fail(BNM_E_ARG, "bnm_analytical_fic: bad arguments");
    int r = check_device();
    if (r) return r;
    const size_t N = nodes, B = n_sims;
    double *dSC = nullptr, *dG = nullptr, *dEE = nullptr, *dEI = nullptr, *dIE = nullptr;
    int* dU = nullptr;
    auto run = [&]() -> int {
        HIPCHK(hipMalloc(&dSC, N * N * sizeof(double)));
        HIPCHK(hipMalloc(&dG, B * sizeof(double)));
        HIPCHK(hipMalloc(&dEE, B * N * sizeof(double)));
        HIPCHK(hipMalloc(&dEI, B * N * sizeof(double)));
        HIPCHK(hipMalloc(&dIE, B * N * sizeof(double)));
        HIPCHK(hipMalloc(&dU, B * sizeof(int)));
        HIPCHK(hipMemcpy(dSC, SC, N * N * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dG, G, B * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dEE, wEE, B * N * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(dEI, wEI, B * N * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(launch_fic_analytical(dSC, nodes, n_sims, dG, dEE, dEI, dIE, dU, nullptr));
        HIPCHK(hipMemcpy(wIE_out, dIE, B * N * sizeof(double), hipMemcpyDeviceToHost));
        if (unstable) HIPCHK(hipMemcpy(unstable, dU, B * sizeof(int), hipMemcpyDeviceToHost));
        return BNM_OK;
    };
    r = run();
    for (void* p : {(void*)dSC, (void*)dG, (void*)dEE, (void*)dEI, (void*)dIE, (void*)dU})
        if (p) (void)hipFree(p);
    return r;
}

// Context on the current device; with `arena` every device buffer comes from
// that memory (no-allocation mode, see bnm_run_simulations_noalloc).
static bnm_ctx* ctx_create(const bnm_config* cfg, int* status, unsigned char* arena, size_t arena_cap) {
    auto set = [&](int s) { if (status) *status = s; };
    if (!cfg) { set(fail(BNM_E_ARG, "null config")); return nullptr; }
    if (int r = check_device()) { set(r); return nullptr; }
    bnm_config c = *cfg;
    int r = bnm_finalize_config(&c);
    if (r) { set(r); return nullptr; }
    if (c.precision == BNM_PREC_F64 && (c.kernel == BNM_KERNEL_MFMA)) {
        set(fail(BNM_E_UNSUP, "the MFMA integration kernel runs in FP32/BF16 only"));
        return nullptr;
    }
    if (c.precision < BNM_PREC_F64 || c.precision > BNM_PREC_BF16X3) {
        set(fail(BNM_E_ARG, "unknown precision"));
        return nullptr;
    }
    if (c.precision >= BNM_PREC_BF16 && c.kernel != BNM_KERNEL_MFMA && c.kernel != BNM_KERNEL_AUTO) {
        set(fail(BNM_E_UNSUP, "BF16 coupling needs the MFMA kernel"));
        return nullptr;
    }
    auto* x = new bnm_ctx;
    x->cfg = c;
    x->N = c.nodes;
    x->B = c.n_sims;
    x->ots = c.time_steps / c.bold_tr + 1;
    x->f64 = c.precision == BNM_PREC_F64;
    if (c.fc_precision < BNM_FC_F64 || c.fc_precision > BNM_FC_BF16X3) {
        set(fail(BNM_E_ARG, "unknown fc_precision"));
        delete x;
        return nullptr;
    }
    x->fc64 = c.fc_precision == BNM_FC_F64;
    x->fcbf16 = c.fc_precision == BNM_FC_BF16 || c.fc_precision == BNM_FC_BF16X3;
    x->fcsplit = c.fc_precision == BNM_FC_BF16X3;
    // BNM_FCD_DIRECT=1: z-score inside the FCD Gram's tile loader instead of
    // a separate pass (no z buffer; slower, kept for comparison)
    x->fcd_direct = env_int("BNM_FCD_DIRECT", 0) != 0;
    // BNM_FCD_TILE=128: the FP32 FCD Gram over 128x128 output tiles (bit-identical;
    // slower at fcd400: 28.8 vs 24.8 ms, a third as many workgroups)
    x->fcd_tile = env_int("BNM_FCD_TILE", 64);
    if (!c.sim_only) {
        x->np = bnm_n_pairs(c.nodes, c.exc_interhemispheric);
        x->nw = c.no_fcd ? 0 : bnm_n_windows(&c);
        x->nwp = (long long)x->nw * (x->nw - 1) / 2;
        x->nseg = 1 + x->nw;
    }
    x->fast_ls = env_int("BNM_FAST_LS", 0);
    x->fast_ks = env_int("BNM_FAST_KS", 0);
    x->stream_ks = env_int("BNM_STREAM_KS", 0);
    x->force_stream = env_int("BNM_FORCE_STREAM", 0);
    x->no_ei = env_int("BNM_NO_EI", 0);
    x->ei_ls = env_int("BNM_EI_LS", 0);
    // BNM_EI_BC=1 adds the DPP-broadcast E/I variant to tuning (or pins it
    // with BNM_EI_LS=4); off by default: bit-identical to the LDS form and
    // slower at N = 100 (0.447 vs 0.382 us per substep)
    x->ei_bc = env_int("BNM_EI_BC", 0);
    // BNM_EI_FORM pins the E/I update form (0 as written, 2 / 3 short chain,
    // 67 = 3 | 64 strict short chain, the default; 75 / 11: strict / plain
    // with the FP64 BW step, 27: FP64 update, diagnostics of the N = 100
    // shape)
    x->ei_form = env_int("BNM_EI_FORM", -1);
    // BNM_HYB_BC: 0 / 1 / 2 restricts the hybrid kernel to LDS / DPP-broadcast
    // S_E / DPP broadcast with three streamed chunks in flight
    x->hyb_bc = env_int("BNM_HYB_BC", -1);
    // BNM_HYB_RR forces the hybrid kernel's register part (values per lane)
    x->hyb_rr = env_int("BNM_HYB_RR", -1);
    x->no_hyb = env_int("BNM_NO_HYB", 0);
    // BNM_SPARSE pins the sparse FP32 kernel where the SC fits it (1: one
    // lane per node, 2: two, 3: either, tuned); BNM_NO_SPARSE=1 leaves it out
    // of the tuner's candidates
    x->sparse_only = env_int("BNM_SPARSE", 0);
    x->no_sparse = env_int("BNM_NO_SPARSE", 0);
    // BNM_SPARSE_STRICT=0: the sparse kernel with the plain FP32 node update
    // instead of the strict one (the default)
    x->sparse_mode = env_int("BNM_SPARSE_STRICT", 1) ? 64 : 0;
    x->hyb_ks = env_int("BNM_HYB_KS", 0);
    {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            x->lds_max = (size_t)v;
    }
    x->autotune = c.autotune;
    x->tune_ms = env_int("BNM_TUNE_MS", 300);
    if (arena) {
        // the first candidate kernel, no timing windows, no graph, null stream
        x->arena = arena;
        x->arena_cap = arena_cap;
        x->autotune = 0;
        x->cfg.use_graph = 0;
    } else if (hipStreamCreateWithFlags(&x->st, hipStreamNonBlocking) != hipSuccess) {
        set(fail(BNM_E_HIP, "hipStreamCreate failed"));
        delete x;
        return nullptr;
    }
    r = ctx_alloc(x);
    if (r) {
        bnm_ctx_destroy(x);
        set(r);
        return nullptr;
    }
    set(BNM_OK);
    return x;
}

bnm_ctx* bnm_ctx_create(const bnm_config* cfg, int* status) { return ctx_create(cfg, status, nullptr, 0); }

int bnm_ctx_set_sc(bnm_ctx* x, const double* SC) {
    if (!x || !SC) return fail(BNM_E_ARG, "null argument");
    const size_t N = x->N, n = N * N;
    std::vector<float> f(n), ft(n);
    std::vector<double> dt(n);
    for (size_t j = 0; j < N; ++j)
        for (size_t k = 0; k < N; ++k) {
            f[j * N + k] = (float)SC[j * N + k];
            ft[k * N + j] = (float)SC[j * N + k];
            dt[k * N + j] = SC[j * N + k];
        }
    HIPCHK(hipMemcpy(x->SC64, SC, n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->SC32, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->SCt64, dt.data(), n * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->SCt32, ft.data(), n * sizeof(float), hipMemcpyHostToDevice));
    // chunked copy: SCc[((k/16)*N + j)*16 + k%16] = SC[j][k]; k >= N stays zero
    std::vector<float> fc(((N + 15) / 16) * 16 * N, 0.0f);
    for (size_t j = 0; j < N; ++j)
        for (size_t k = 0; k < N; ++k) fc[((k / 16) * N + j) * 16 + k % 16] = (float)SC[j * N + k];
    HIPCHK(hipMemcpy(x->SCc32, fc.data(), fc.size() * sizeof(float), hipMemcpyHostToDevice));
    if (x->SCc64) {
        // SCc64[((k/8)*N + j)*8 + k%8] = SC[j][k]
        std::vector<double> dc(((N + 7) / 8) * 8 * N, 0.0);
        for (size_t j = 0; j < N; ++j)
            for (size_t k = 0; k < N; ++k) dc[((k / 8) * N + j) * 8 + k % 8] = SC[j * N + k];
        HIPCHK(hipMemcpy(x->SCc64, dc.data(), dc.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    int sp[2] = {0, 0};
    for (int i = 0; i < 2; ++i) {
        if (!x->SPd[i]) continue;
        std::vector<unsigned> d;
        int P = 0;
        if (x->f64) {
            std::vector<double> w;
            P = build_sparse_sc(SC, (int)N, i + 1, w, d);
            if (P > 0 && P <= kSparseRW64)
                HIPCHK(hipMemcpy(x->SPw64[i], w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
            else
                P = 0;   // too dense for the FP64 kernel (or empty)
        } else {
            std::vector<float> w;
            P = build_sparse_sc(f.data(), (int)N, i + 1, w, d);
            if (P > 0 && P <= kSparseRW + kSparseLL)
                HIPCHK(hipMemcpy(x->SPw[i], w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice));
            else
                P = 0;   // too dense for the kernel (or empty)
        }
        if (P > 0) HIPCHK(hipMemcpy(x->SPd[i], d.data(), d.size() * sizeof(unsigned), hipMemcpyHostToDevice));
        sp[i] = P;
    }
    {
        // the slot counts are launch arguments and set the kernel's shape: a
        // new count re-runs the kernel choice and the graph capture
        if ((sp[0] != x->sp_P[0] || sp[1] != x->sp_P[1]) && x->tuned) {
            x->tuned = false;
            if (x->graph_exec) {
                (void)hipGraphExecDestroy(x->graph_exec);
                x->graph_exec = nullptr;
            }
            ++x->sc_gen;
        }
        x->sp_P[0] = sp[0];
        x->sp_P[1] = sp[1];
    }
    x->have_sc = true;
    return BNM_OK;
}

int bnm_ctx_set_noise(bnm_ctx* x, const float* noise, long long len) {
    if (!x || !noise) return fail(BNM_E_ARG, "null argument");
    const long long need = x->cfg.noise_period_ms * 10LL * x->N * 2;
    if (len < need) return fail(BNM_E_ARG, "noise array shorter than noise_period_ms*10*N*2");
    if (x->noise_cap < need) {
        if (x->noise && !x->arena) (void)hipFree(x->noise);
        x->noise = nullptr;
        if (int r = cdalloc(x, &x->noise, (size_t)need)) return r;
        x->noise_cap = need;
    }
    HIPCHK(hipMemcpy(x->noise, noise, need * sizeof(float), hipMemcpyHostToDevice));
    x->have_noise = true;
    return BNM_OK;
}

int bnm_ctx_set_empirical(bnm_ctx* x, const double* emp_fc, int n_fc, const double* emp_fcd,
                          int n_fcd) {
    if (!x) return fail(BNM_E_ARG, "null context");
    if (x->cfg.sim_only) { x->have_emp = true; return BNM_OK; }
    if (!emp_fc || n_fc != x->np) return fail(BNM_E_SIZE, "empirical FC size does not match the pair count");
    // centred empirical FC and its norm (the constant half of the Pearson)
    long double s = 0;
    for (int i = 0; i < n_fc; ++i) s += emp_fc[i];
    const double mean = (double)(s / n_fc);
    std::vector<double> c(n_fc);
    long double ss = 0;
    for (int i = 0; i < n_fc; ++i) {
        c[i] = emp_fc[i] - mean;
        ss += (long double)c[i] * c[i];
    }
    x->emp_mean = mean;
    x->emp_ssd = std::sqrt((double)ss);
    HIPCHK(hipMemcpy(x->emp_c, c.data(), n_fc * sizeof(double), hipMemcpyHostToDevice));
    if (!x->cfg.no_fcd) {
        if (!emp_fcd || n_fcd < 1) return fail(BNM_E_SIZE, "empty empirical FCD");
        std::vector<double> u(emp_fcd, emp_fcd + n_fcd);
        std::sort(u.begin(), u.end());
        std::vector<double> uu, fe;
        for (int i = 0; i < n_fcd; ++i) {
            if (i + 1 < n_fcd && u[i + 1] == u[i]) continue;  // last of a run of ties
            uu.push_back(u[i]);
            fe.push_back((double)(i + 1) / n_fcd);
        }
        if (!x->arena) {
            if (x->ks_u) (void)hipFree(x->ks_u);
            if (x->ks_fe) (void)hipFree(x->ks_fe);
            if (x->bins) (void)hipFree(x->bins);
        }
        x->ks_u = x->ks_fe = nullptr;
        x->bins = nullptr;
        x->m = (int)uu.size();
        int r;
        if ((r = cdalloc(x, &x->ks_u, uu.size()))) return r;
        if ((r = cdalloc(x, &x->ks_fe, fe.size()))) return r;
        if ((r = cdalloc(x, &x->bins, (size_t)x->B * (2 * x->m + 1)))) return r;
        HIPCHK(hipMemcpy(x->ks_u, uu.data(), uu.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(x->ks_fe, fe.data(), fe.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    x->have_emp = true;
    return BNM_OK;
}

int bnm_ctx_set_params(bnm_ctx* x, const double* G, const double* wEE, const double* wEI,
                       const double* wIE, const double* wIE_fic) {
    if (!x || !G || !wEE || !wEI) return fail(BNM_E_ARG, "null argument");
    const size_t B = x->B, N = x->N;
    HIPCHK(hipMemcpy(x->G, G, B * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->wEE, wEE, B * N * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(x->wEI, wEI, B * N * sizeof(double), hipMemcpyHostToDevice));
    if (wIE) HIPCHK(hipMemcpy(x->wIE, wIE, B * N * sizeof(double), hipMemcpyHostToDevice));
    else HIPCHK(hipMemset(x->wIE, 0, B * N * sizeof(double)));
    if (x->cfg.do_fic && !x->cfg.analytical_fic) {
        if (!wIE_fic) return fail(BNM_E_ARG, "w_IE_fic is required when analytical_fic == 0");
        HIPCHK(hipMemcpy(x->wIEfic, wIE_fic, B * N * sizeof(double), hipMemcpyHostToDevice));
    }
    x->have_params = true;
    return BNM_OK;
}

// ---- integration variants ------------------------------------------------------

static bool variant_ok(const bnm_ctx* x, const Variant& v) {
    const size_t el = x->f64 ? 8 : 4;
    switch (v.kind) {
        case 0: return fast_shape(x->N, v.ls, v.ks, el).ok;
        case 1: return stream_shape(x->N, v.ks, el).ok;
        case 2: return !x->f64 && mfma_shape(x->N, v.mode).ok;
        case 4: return !x->f64 && ei_shape(x->N, v.ls, v.mode == 1).ok;
        case 5: return !x->f64 && hyb_shape(x->N, v.ks, x->lds_max, x->hyb_rr, v.mode).ok;
        case 6: return x->f64 && ei64_shape(x->N, v.ls).ok;
        case 7: return x->f64 && hyb64_shape(x->N, v.ks, x->lds_max, x->hyb_rr).ok;
        case 8:
            return (v.ls == 1 || v.ls == 2) && x->sp_P[v.ls - 1] > 0 && !(x->f64 && (v.mode & 64)) &&
                   sparse_shape(x->N, x->sp_P[v.ls - 1], v.ls, x->lds_max, el).ok;
        default: return true;
    }
}

static std::string variant_name(const bnm_ctx* x, const Variant& v) {
    const size_t el = x->f64 ? 8 : 4;
    const char* p = x->f64 ? "f64" : "f32";
    char name[160];
    if (v.kind == 0) {
        const FastShape s = fast_shape(x->N, v.ls, v.ks, el);
        std::snprintf(name, sizeof name, "split_%s(LS=%d,KS=%d,NW=%d,KC=%d,SS=%d,threads=%d)", p,
                      s.LS, s.KS, s.NW, s.KC, s.SS, s.threads);
    } else if (v.kind == 1) {
        const FastShape g = stream_shape(x->N, v.ks, el);
        std::snprintf(name, sizeof name, "stream_%s(KS=%d,NW=%d,kc=%d,threads=%d)", p, g.KS, g.NW,
                      g.kc, g.threads);
    } else if (v.kind == 2) {
        static const char* mn[] = {"f32", "bf16", "bf16x3"};
        const MfmaShape s = mfma_shape(x->N, v.mode);
        std::snprintf(name, sizeof name, "mfma_%s(sims/block=16,waves=%d,KQ=%d,threads=%d)",
                      mn[v.mode], s.waves, s.KQ, s.threads);
    } else if (v.kind == 5) {
        const FastShape g = hyb_shape(x->N, v.ks, x->lds_max, x->hyb_rr, v.mode);
        std::snprintf(name, sizeof name, "hyb_f32(KS=%d,NW=%d,kc=%d,regs=%d,lds=%d,streamed=%d,threads=%d%s)", g.KS,
                      g.NW, g.kc, g.KC, g.SS, g.kc - g.KC - g.SS, g.threads,
                      v.mode == 2 ? ",dpp,pf3" : v.mode == 1 ? ",dpp" : "");
    } else if (v.kind == 8) {
        const SparseShape g = sparse_shape(x->N, x->sp_P[v.ls - 1], v.ls, x->lds_max, el);
        std::snprintf(name, sizeof name, "sparse_%s(LS=%d,P=%d,regs=%d,lds=%d,threads=%d%s)", p, g.LS, g.P, g.RW,
                      g.LL, g.threads, v.mode & 64 ? ",strict" : "");
    } else if (v.kind == 7) {
        const FastShape g = hyb64_shape(x->N, v.ks, x->lds_max, x->hyb_rr);
        std::snprintf(name, sizeof name, "hyb_f64(KS=%d,NW=%d,kc=%d,regs=%d,lds=%d,streamed=%d,threads=%d)", g.KS,
                      g.NW, g.kc, g.KC, g.SS, g.kc - g.KC - g.SS, g.threads);
    } else if (v.kind == 6) {
        const FastShape s = ei64_shape(x->N, v.ls);
        std::snprintf(name, sizeof name, "split_ei_f64(LS=%d,NW=%d,KC=%d,SS=%d,threads=%d)", s.LS, s.NW, s.KC,
                      s.SS, s.threads);
    } else if (v.kind == 4) {
        const FastShape s = ei_shape(x->N, v.ls, v.mode == 1);
        static const char* fm[] = {"", ",dpp", ",short4", ",short1"};
        std::snprintf(name, sizeof name, "split_ei_f32(LS=%d,NW=%d,KC=%d,SS=%d,threads=%d%s%s%s%s)", s.LS, s.NW,
                      s.KC, s.SS, s.threads, fm[v.mode & 3], v.mode & 64 ? ",strict" : "",
                      v.mode & 8 ? ",bw64" : "", v.mode & 16 ? ",upd64" : "");
    } else {
        std::snprintf(name, sizeof name, "plain_%s(threads=%d)", p, ((x->N + 63) / 64) * 64);
    }
    return name;
}

static hipError_t launch_variant(bnm_ctx* x, const Variant& v, const SimArgs& a, hipStream_t st) {
    switch (v.kind) {
        case 0:
            return x->f64 ? launch_sim_fast<double>(x->SC64, a, v.ls, v.ks, st)
                          : launch_sim_fast<float>(x->SC32, a, v.ls, v.ks, st);
        case 1:
            return x->f64 ? launch_sim_stream<double>(x->SCt64, a, v.ks, st)
                          : launch_sim_stream<float>(x->SCt32, a, v.ks, st);
        case 2: return launch_sim_mfma(x->SC32, a, v.mode, st);
        case 4: return launch_sim_ei(x->SC32, a, v.ls, v.mode, st);
        case 5: return launch_sim_hyb(x->SCc32, a, v.ks, x->lds_max, x->hyb_rr, v.mode, st);
        case 6: return launch_sim_ei64(x->SC64, a, v.ls, st);
        case 7: return launch_sim_hyb64(x->SCc64, a, v.ks, x->lds_max, x->hyb_rr, st);
        case 8:
            if (x->f64)
                return launch_sim_sparse64(x->SPw64[v.ls - 1], x->SPd[v.ls - 1], a, x->sp_P[v.ls - 1], v.ls, x->lds_max,
                                           st);
            return launch_sim_sparse(x->SPw[v.ls - 1], x->SPd[v.ls - 1], a, x->sp_P[v.ls - 1], v.ls, (v.mode & 64) != 0,
                                     x->lds_max, st);
        default:
            return x->f64 ? launch_sim_plain<double>(x->SC64, a, st)
                          : launch_sim_plain<float>(x->SC32, a, st);
    }
}

// Candidates that implement the configured precision with the configured kernel.
static std::vector<Variant> candidates(const bnm_ctx* x) {
    const bnm_config& c = x->cfg;
    std::vector<Variant> out;
    const int mode = c.precision == BNM_PREC_BF16 ? 1 : c.precision == BNM_PREC_BF16X3 ? 2 : 0;
    auto add = [&](Variant v) { if (variant_ok(x, v)) out.push_back(v); };
    if (c.kernel == BNM_KERNEL_PLAIN) {
        add({3, 0, 0, 0});
        return out;
    }
    if (c.precision >= BNM_PREC_BF16) {
        add({2, 0, 0, mode});
        return out;
    }
    if (c.kernel == BNM_KERNEL_MFMA) {
        add({2, 0, 0, 0});
        return out;
    }
    const int spm = x->f64 ? 0 : x->sparse_mode;   // FP64: the reference-order update
    if (x->sparse_only) {
        if (x->sparse_only & 1) add({8, 1, 1, spm});
        if (x->sparse_only & 2) add({8, 2, 1, spm});
        if (!out.empty()) return out;
    }
    if (x->hyb_ks > 0) {
        if (x->f64) add({7, 1, x->hyb_ks, 0});
        else add({5, 1, x->hyb_ks, x->hyb_bc > 0 ? x->hyb_bc : 0});
        if (!out.empty()) return out;
    }
    if (x->ei_ls > 0) {
        if (x->f64) add({6, x->ei_ls, 1, 0});
        else add({4, x->ei_ls, 1, x->ei_form >= 0 ? x->ei_form : x->ei_bc == 1 ? 1 : (3 | 64)});
        if (!out.empty()) return out;
    }
    const bool pinned = x->fast_ls > 0 || x->fast_ks > 0;
    if (pinned && !x->force_stream) {
        add({0, x->fast_ls > 0 ? x->fast_ls : 4, x->fast_ks > 0 ? x->fast_ks : 1, 0});
        if (!out.empty()) return out;
    }
    if (!x->force_stream) {
        if (x->f64 && !x->no_ei) {
            add({6, 2, 1, 0});
            add({6, 4, 1, 0});
        }
        if (!x->f64 && !x->no_ei) {
            if (x->ei_bc == 1) add({4, 4, 1, 1});
            // form 3|64 (strict short chain, one partial sum) is tuned by
            // default; the plain forms 3, 2 and 0 (faster by ~6 %, with a
            // systematic per-TR BOLD error of up to ~2e-5 against FP64) only
            // when pinned with BNM_EI_FORM
            for (int f : {3 | 64, 3, 2, 0})
                if (x->ei_form < 0 ? f == (3 | 64) : x->ei_form == f) {
                    add({4, 4, 1, f});
                    add({4, 2, 1, f});
                }
        }
        static const int pref[][2] = {{4, 1}, {2, 1}, {1, 1}, {2, 2}, {1, 2}, {4, 2},
                                      {1, 4}, {2, 4}, {1, 8}, {2, 8}, {1, 16}};
        for (const auto& p : pref) add({0, p[0], p[1], 0});
    }
    if (out.empty() || x->force_stream) {
        if (x->stream_ks > 0) {
            add({1, 1, x->stream_ks, 0});
            if (x->f64 && !x->no_hyb) add({7, 1, x->stream_ks, 0});
            if (!x->f64 && !x->no_hyb)
                for (int m : {0, 1, 2})
                    if (x->hyb_bc < 0 || x->hyb_bc == m) add({5, 1, x->stream_ks, m});
        } else {
            for (int k : {16, 8, 4, 2, 1}) add({1, 1, k, 0});
            if (x->f64 && !x->no_hyb)
                for (int k : {8, 4, 2, 1}) add({7, 1, k, 0});
            if (!x->f64 && !x->no_hyb)
                for (int k : {8, 4, 2, 1})
                    for (int m : {0, 1, 2})
                        if (x->hyb_bc < 0 || x->hyb_bc == m) add({5, 1, k, m});
        }
        // sparse row lists (SCs of partial density); beside the streaming
        // kernels only, whose node update it shares (at N <= 128 the tuned
        // default is the strict E/I kernel)
        if (!x->no_sparse && x->stream_ks <= 0) {
            add({8, 1, 1, spm});
            add({8, 2, 1, spm});
        }
    }
    // the MFMA kernel implements FP32 too; the tuner may pick it for large populations
    if (c.kernel == BNM_KERNEL_AUTO && !x->f64 && !pinned && !x->force_stream) add({2, 0, 0, 0});
    return out;
}

static SimArgs make_args(bnm_ctx* x) {
    const bnm_config& c = x->cfg;
    SimArgs a{};
    a.N = x->N;
    a.B = x->B;
    a.time_steps = c.time_steps;
    a.bold_tr = c.bold_tr;
    a.output_ts = x->ots;
    a.n_vols_remove = c.n_vols_remove;
    a.adjust_fic = c.do_fic && c.adjust_fic;
    a.max_fic_trials = c.max_fic_trials;
    a.i_start = c.i_sampling_start;
    a.i_end = c.i_sampling_end;
    a.init_delta = c.init_delta;
    a.noise_period = c.noise_period_ms;
    a.noise = x->noise;
    a.G = x->G;
    a.wEE = x->wEE;
    a.wEI = x->wEI;
    a.wIE_in = c.do_fic ? x->wIEfic : x->wIE;
    a.wIE_out = c.do_fic ? x->wIEadj : x->wIEscratch;
    a.bold = x->bold;
    a.ext = x->ext;
    a.fic_failed = x->failed;
    a.fic_trials = x->trials;
    a.bw_k1 = c.bw_friston2003 ? C64::k1_F : C64::k1;
    a.bw_k2 = c.bw_friston2003 ? C64::k2_F : C64::k2;
    a.bw_k3 = c.bw_friston2003 ? C64::k3_F : C64::k3;
    return a;
}

// Picks the integration variant once per context.  With kernel == AUTO and
// more than one candidate, every candidate integrates the real population for
// a short window (FIC off) and the fastest by HIP-event time is kept; the
// window covers the time-loop steady state, where the workgroups-per-CU and
// per-substep latency trade-off between the variants shows.
static int choose_variant(bnm_ctx* x) {
    if (x->tuned) return BNM_OK;
    std::vector<Variant> cs = candidates(x);
    if (cs.empty()) return fail(BNM_E_UNSUP, "no integration kernel fits this configuration");
    x->variant = cs[0];
    x->tune_log.clear();
    if (cs.size() > 1 && x->cfg.kernel == BNM_KERNEL_AUTO && x->autotune) {
        SimArgs a = make_args(x);
        a.time_steps = std::min(x->cfg.time_steps, x->tune_ms);
        a.adjust_fic = 0;
        a.wIE_out = x->wIEscratch;
        double best = 1e30;
        for (const Variant& v : cs) {
            float ms = 0, mn = 1e30f;
            for (int rep = 0; rep < 2; ++rep) {
                HIPCHK(hipEventRecord(x->ev[0], x->st));
                HIPCHK(launch_variant(x, v, a, x->st));
                HIPCHK(hipEventRecord(x->ev[1], x->st));
                HIPCHK(hipEventSynchronize(x->ev[1]));
                HIPCHK(hipEventElapsedTime(&ms, x->ev[0], x->ev[1]));
                mn = std::min(mn, ms);
            }
            char buf[224];
            std::snprintf(buf, sizeof buf, "%s%s=%.3fms", x->tune_log.empty() ? "" : "; ",
                          variant_name(x, v).c_str(), mn);
            x->tune_log += buf;
            if (mn < best) { best = mn; x->variant = v; }
        }
        x->ms_per_sim_ms = best / std::max(1, (int)a.time_steps + 1);
    }
    x->kname = variant_name(x, x->variant);
    x->tuned = true;
    return BNM_OK;
}

// Enqueues one whole run on stream `st`; with `timed` the phase events are
// recorded (not inside graph capture).
// analyze_only: FC, FCD and fit of the BOLD already in the context (no FIC
// initialisation, no integration, no FIC penalty)
static int enqueue_run(bnm_ctx* x, hipStream_t st, bool timed, bool analyze_only = false) {
    const bnm_config& c = x->cfg;
    if (timed) HIPCHK(hipEventRecord(x->ev[0], st));
    if (c.do_fic && c.analytical_fic && !analyze_only)
        HIPCHK(launch_fic_analytical(x->SC64, x->N, x->B, x->G, x->wEE, x->wEI, x->wIEfic,
                                     x->unstable, st));
    if (timed) HIPCHK(hipEventRecord(x->ev[1], st));
    if (!analyze_only) {
        const SimArgs a = make_args(x);
        HIPCHK(launch_variant(x, x->variant, a, st));
    }
    if (timed) HIPCHK(hipEventRecord(x->ev[2], st));
    if (!c.sim_only) {
        if (x->f64) {
            HIPCHK(launch_seg_moments<double>((const double*)x->bold, x->B, x->N, x->ots, x->nseg,
                                              x->seg_start, x->seg_len, x->mean, x->inv, st));
            if (x->fc64)
                HIPCHK((launch_gram_fc<double, double, double>(
                    (const double*)x->bold, x->B, x->N, x->ots, x->nseg, x->seg_start, x->seg_len,
                    x->mean, x->inv, x->pidx, x->np, (double*)x->fc, (double*)x->wfc, st)));
            else if (x->fcbf16)
                HIPCHK(launch_gram_fc_bf16<double>(
                    (const double*)x->bold, x->B, x->N, x->ots, x->nseg, x->seg_start, x->seg_len,
                    x->mean, x->inv, x->pidx, x->np, (float*)x->fc, (float*)x->wfc, x->fcsplit, st));
            else
                HIPCHK((launch_gram_fc<double, float, float>(
                    (const double*)x->bold, x->B, x->N, x->ots, x->nseg, x->seg_start, x->seg_len,
                    x->mean, x->inv, x->pidx, x->np, (float*)x->fc, (float*)x->wfc, st)));
        } else {
            HIPCHK(launch_seg_moments<float>((const float*)x->bold, x->B, x->N, x->ots, x->nseg,
                                             x->seg_start, x->seg_len, x->mean, x->inv, st));
            if (x->fc64)
                HIPCHK((launch_gram_fc<float, double, double>(
                    (const float*)x->bold, x->B, x->N, x->ots, x->nseg, x->seg_start, x->seg_len,
                    x->mean, x->inv, x->pidx, x->np, (double*)x->fc, (double*)x->wfc, st)));
            else if (x->fcbf16)
                HIPCHK(launch_gram_fc_bf16<float>(
                    (const float*)x->bold, x->B, x->N, x->ots, x->nseg, x->seg_start, x->seg_len,
                    x->mean, x->inv, x->pidx, x->np, (float*)x->fc, (float*)x->wfc, x->fcsplit, st));
            else
                HIPCHK((launch_gram_fc<float, float, float>(
                    (const float*)x->bold, x->B, x->N, x->ots, x->nseg, x->seg_start, x->seg_len,
                    x->mean, x->inv, x->pidx, x->np, (float*)x->fc, (float*)x->wfc, st)));
        }
    }
    if (timed) HIPCHK(hipEventRecord(x->ev[3], st));
    GofArgs g{};
    g.B = x->B;
    g.N = x->N;
    g.np = x->np;
    g.nwp = x->nwp;
    g.m = x->m;
    g.emp_fc_c = x->emp_c;
    g.emp_fc_ssd = x->emp_ssd;
    g.emp_fc_mean = x->emp_mean;
    g.ks_u = x->ks_u;
    g.ks_fe = x->ks_fe;
    g.bins = x->bins;
    g.nan_flag = x->nanf;
    g.ext = x->ext;
    g.fic_failed = x->failed;
    g.sim_only = c.sim_only;
    g.no_fcd = c.no_fcd;
    g.use_fc_diff = c.use_fc_diff;
    g.calc_penalty = c.calculate_fic_penalty && !analyze_only;
    g.fic_reject = c.fic_reject_failed;
    g.max_cost = c.max_cost;
    g.fic_scale = c.fic_penalty_scale;
    g.corr = x->corr;
    g.fic_pen = x->pen;
    if (!c.sim_only && !c.no_fcd) {
        if (x->fc64) {
            HIPCHK(launch_row_moments<double>((const double*)x->wfc, (long long)x->B * x->nw, x->np,
                                              x->wmean, x->winv, st));
            HIPCHK((launch_gram_fcd<double, double, double>((const double*)x->wfc, x->B, x->nw, x->np,
                                                            x->wmean, x->winv, (double*)x->fcd, st)));
        } else {
            HIPCHK(launch_row_moments<float>((const float*)x->wfc, (long long)x->B * x->nw, x->np,
                                             x->wmean, x->winv, st));
            if (x->wz)
                HIPCHK(launch_fcd_z((const float*)x->wfc, x->B, x->nw, x->np, x->wmean, x->winv, x->wz, x->wzlo,
                                    x->fcsplit ? 2 : x->fcbf16 ? 1 : 0, (float*)x->fcd, st, x->fcd_tile));
            else if (x->fcbf16)
                HIPCHK(launch_gram_fcd_bf16((const float*)x->wfc, x->B, x->nw, x->np, x->wmean, x->winv,
                                            (float*)x->fcd, x->fcsplit, st));
            else
                HIPCHK((launch_gram_fcd<float, float, float>((const float*)x->wfc, x->B, x->nw, x->np,
                                                             x->wmean, x->winv, (float*)x->fcd, st)));
        }
        // left / right FCD (bnm.cu:453-495, 519-550): the same windows over the
        // first and the last n_hemi edges of each window FC
        for (int h = 0; h < 2 && c.save_hemis && x->n_hemi > 1; ++h) {
            const int col0 = h == 0 ? 0 : x->np - x->n_hemi;
            const long long rows = (long long)x->B * x->nw;
            if (x->fc64) {
                HIPCHK(launch_row_moments<double>((const double*)x->wfc, rows, x->n_hemi, x->wmean_h[h],
                                                  x->winv_h[h], st, x->np, col0));
                HIPCHK((launch_gram_fcd<double, double, double>((const double*)x->wfc, x->B, x->nw, x->np,
                                                                x->wmean_h[h], x->winv_h[h],
                                                                (double*)x->fcd_h[h], st, col0, x->n_hemi)));
            } else {
                HIPCHK(launch_row_moments<float>((const float*)x->wfc, rows, x->n_hemi, x->wmean_h[h],
                                                 x->winv_h[h], st, x->np, col0));
                if (x->fcbf16)
                    HIPCHK(launch_gram_fcd_bf16((const float*)x->wfc, x->B, x->nw, x->np, x->wmean_h[h],
                                                x->winv_h[h], (float*)x->fcd_h[h], x->fcsplit, st, col0,
                                                x->n_hemi));
                else
                    HIPCHK((launch_gram_fcd<float, float, float>((const float*)x->wfc, x->B, x->nw, x->np,
                                                                 x->wmean_h[h], x->winv_h[h],
                                                                 (float*)x->fcd_h[h], st, col0, x->n_hemi)));
            }
        }
    }
    if (timed) HIPCHK(hipEventRecord(x->ev[4], st));
    if (!c.sim_only && !c.no_fcd) {
        HIPCHK(launch_fill_i32(x->bins, (long long)x->B * (2 * x->m + 1), 0, st));
        HIPCHK(launch_fill_i32(x->nanf, x->B, 0, st));
        if (x->fc64) HIPCHK(launch_ks_bin<double>((const double*)x->fcd, g, st));
        else HIPCHK(launch_ks_bin<float>((const float*)x->fcd, g, st));
    }
    if (x->fc64) HIPCHK(launch_gof<double>((const double*)x->fc, g, st));
    else HIPCHK(launch_gof<float>((const float*)x->fc, g, st));
    if (timed) HIPCHK(hipEventRecord(x->ev[5], st));
    return BNM_OK;
}

static bool ready(const bnm_ctx* x) {
    return x->have_sc && x->have_noise && x->have_params && (x->cfg.sim_only || x->have_emp);
}

int bnm_ctx_set_bold(bnm_ctx* x, const double* bold) {
    if (!x || !bold) return fail(BNM_E_ARG, "null argument");
    const size_t n = (size_t)x->B * x->ots * x->N;
    if (x->f64) {
        HIPCHK(hipMemcpy(x->bold, bold, n * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(bold, bold + n);
        HIPCHK(hipMemcpy(x->bold, f.data(), n * sizeof(float), hipMemcpyHostToDevice));
    }
    x->have_bold = true;
    return BNM_OK;
}

int bnm_ctx_analyze(bnm_ctx* x) {
    if (!x) return fail(BNM_E_ARG, "null context");
    if (x->cfg.sim_only) return fail(BNM_E_ARG, "sim_only context has no FC/FCD to compute");
    if (!x->have_emp || !x->have_bold) return fail(BNM_E_ARG, "empirical target and BOLD must be set first");
    int r = enqueue_run(x, x->st, true, true);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(x->st));
    float ms = 0;
    for (int i = 0; i < 4; ++i) {
        HIPCHK(hipEventElapsedTime(&ms, x->ev[i], x->ev[i + 1]));
        x->timings[i] = ms;
    }
    HIPCHK(hipEventElapsedTime(&ms, x->ev[4], x->ev[5]));
    x->timings[4] = ms;
    HIPCHK(hipEventElapsedTime(&ms, x->ev[0], x->ev[5]));
    x->timings[5] = ms;
    return BNM_OK;
}

int bnm_ctx_run(bnm_ctx* x) {
    if (!x) return fail(BNM_E_ARG, "null context");
    if (!ready(x))
        return fail(BNM_E_ARG, "SC, noise, parameters and the empirical target must be set first");
    int r = choose_variant(x);
    if (r) return r;
    if (x->arena) {
        r = enqueue_run(x, x->st, false);
        if (r) return r;
        HIPCHK(hipStreamSynchronize(x->st));
        for (int i = 0; i < 6; ++i) x->timings[i] = -1;
        x->have_bold = true;
        return BNM_OK;
    }
    if (x->cfg.use_graph) {
        if (!x->graph_exec) {
            hipGraph_t gph = nullptr;
            HIPCHK(hipStreamBeginCapture(x->st, hipStreamCaptureModeThreadLocal));
            r = enqueue_run(x, x->st, false);
            hipError_t e = hipStreamEndCapture(x->st, &gph);
            if (r) { if (gph) (void)hipGraphDestroy(gph); return r; }
            HIPCHK(e);
            e = hipGraphInstantiate(&x->graph_exec, gph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gph);
            HIPCHK(e);
        }
        HIPCHK(hipEventRecord(x->ev[0], x->st));
        HIPCHK(hipGraphLaunch(x->graph_exec, x->st));
        HIPCHK(hipEventRecord(x->ev[5], x->st));
        HIPCHK(hipStreamSynchronize(x->st));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, x->ev[0], x->ev[5]));
        for (int i = 0; i < 5; ++i) x->timings[i] = -1;
        x->timings[5] = ms;
        x->have_bold = true;
        return BNM_OK;
    }
    r = enqueue_run(x, x->st, true);
    if (r) return r;
    HIPCHK(hipStreamSynchronize(x->st));
    float ms = 0;
    for (int i = 0; i < 4; ++i) {
        HIPCHK(hipEventElapsedTime(&ms, x->ev[i], x->ev[i + 1]));
        x->timings[i] = ms;
    }
    HIPCHK(hipEventElapsedTime(&ms, x->ev[4], x->ev[5]));
    x->timings[4] = ms;
    HIPCHK(hipEventElapsedTime(&ms, x->ev[0], x->ev[5]));
    x->timings[5] = ms;
    x->have_bold = true;
    return BNM_OK;
}

// ---- several contexts in one graph launch ----------------------------------------
struct bnm_multi {
    std::vector<bnm_ctx*> ctxs;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr, t0 = nullptr, t1 = nullptr;
    std::vector<hipEvent_t> joins;
    double ms = 0;
    std::string plan;
    std::vector<int> sc_gen;   // each context's sc_gen at capture
    std::vector<hipEvent_t> pair_events;   // fork / join of side-by-side stages
};

// Integration time estimate of a tuned context (ms): the tuner's time per
// simulated ms, or one timed window of the chosen variant when the tuner did
// not time it, times the run's simulated ms (FIC restarts not included).
static int estimate_ms(bnm_ctx* x, double* out) {
    if (x->ms_per_sim_ms < 0) {
        SimArgs a = make_args(x);
        a.time_steps = std::min(x->cfg.time_steps, x->tune_ms);
        a.adjust_fic = 0;
        a.wIE_out = x->wIEscratch;
        float ms = 0;
        HIPCHK(hipEventRecord(x->ev[0], x->st));
        HIPCHK(launch_variant(x, x->variant, a, x->st));
        HIPCHK(hipEventRecord(x->ev[1], x->st));
        HIPCHK(hipEventSynchronize(x->ev[1]));
        HIPCHK(hipEventElapsedTime(&ms, x->ev[0], x->ev[1]));
        x->ms_per_sim_ms = ms / std::max(1, (int)a.time_steps + 1);
    }
    *out = x->ms_per_sim_ms * (x->cfg.time_steps + 1);
    return BNM_OK;
}

// workgroups of the context's integration kernel (one per simulation, or one
// per 16 for the MFMA kernel)
static int integration_wgs(const bnm_ctx* x) { return x->variant.kind == 2 ? (x->B + 15) / 16 : x->B; }

// A "small" integration workgroup: at most 4 waves and 32 KB of LDS.  Two
// such workgroups of different kernels always fit one CU together (8 waves,
// two per SIMD at <= 256 VGPRs each), and a latency-bound small workgroup
// runs about as fast beside another as alone (two 256-thread E/I workgroups
// per CU take the time of one).
static bool small_workgroup(const bnm_ctx* x) {
    const Variant& v = x->variant;
    const size_t el = x->f64 ? 8 : 4;
    int threads = 1024;
    size_t smem = 0;
    switch (v.kind) {
        case 0: { const FastShape s = fast_shape(x->N, v.ls, v.ks, el); threads = s.threads; smem = s.smem; break; }
        case 1: { const FastShape s = stream_shape(x->N, v.ks, el); threads = s.threads; smem = s.smem; break; }
        case 2: { const MfmaShape s = mfma_shape(x->N, v.mode); threads = s.threads; smem = s.smem; break; }
        case 4: { const FastShape s = ei_shape(x->N, v.ls, v.mode == 1); threads = s.threads; smem = s.smem; break; }
        case 5: {
            const FastShape s = hyb_shape(x->N, v.ks, x->lds_max, x->hyb_rr, v.mode);
            threads = s.threads; smem = s.smem; break;
        }
        case 6: { const FastShape s = ei64_shape(x->N, v.ls); threads = s.threads; smem = s.smem; break; }
        case 7: {
            const FastShape s = hyb64_shape(x->N, v.ks, x->lds_max, x->hyb_rr);
            threads = s.threads; smem = s.smem; break;
        }
        case 8: {
            const SparseShape s = sparse_shape(x->N, x->sp_P[v.ls - 1], v.ls, x->lds_max, el);
            threads = s.threads; smem = s.smem; break;
        }
        default: threads = ((x->N + 63) / 64) * 64; smem = (size_t)x->N * el; break;
    }
    return threads <= 256 && smem <= 32 * 1024;
}

bnm_multi* bnm_multi_create(bnm_ctx** ctxs, int n, int* status) {
    auto set = [&](int s) { if (status) *status = s; };
    if (!ctxs || n < 1) { set(fail(BNM_E_ARG, "no contexts")); return nullptr; }
    for (int i = 0; i < n; ++i)
        if (!ctxs[i] || !ready(ctxs[i])) {
            set(fail(BNM_E_ARG, "every context needs SC, noise, parameters and target set"));
            return nullptr;
        }
    auto* m = new bnm_multi;
    m->ctxs.assign(ctxs, ctxs + n);
    auto bad = [&](int code) { bnm_multi_destroy(m); set(code); return (bnm_multi*)nullptr; };
    for (bnm_ctx* x : m->ctxs) {
        const int r = choose_variant(x);
        if (r) return bad(r);
    }
    if (hipEventCreateWithFlags(&m->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&m->t0) != hipSuccess || hipEventCreate(&m->t1) != hipSuccess)
        return bad(fail(BNM_E_HIP, "hipEventCreate failed"));
    m->joins.resize(n, nullptr);
    for (auto& e : m->joins)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return bad(fail(BNM_E_HIP, "hipEventCreate failed"));
    // Packing: the integration kernels hold one CU each per workgroup for
    // the whole run (register and LDS budgets leave no room for a second),
    // so populations started together beyond the CU count share CUs and slow
    // the longest one down.  K = CUs / (largest workgroup count) branches run
    // concurrently; contexts go to branches longest first, each to the branch
    // with the least estimated work (LPT), and run one after another within
    // it, except that two consecutive contexts with small workgroups
    // (small_workgroup) run side by side on the CUs the branch holds (with
    // 2K branches of CU-filling kernels that means: after the large context
    // before them).  BNM_MULTI_BRANCHES overrides K (n: every context its own
    // branch).
    int cus = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1)
        cus = 256;
    int max_wg = 1;
    std::vector<double> est(n);
    for (int i = 0; i < n; ++i) {
        max_wg = std::max(max_wg, integration_wgs(m->ctxs[i]));
        const int re = estimate_ms(m->ctxs[i], &est[i]);
        if (re) return bad(re);
    }
    int K = env_int("BNM_MULTI_BRANCHES", 0);
    if (K <= 0) K = std::max(1, cus / max_wg);
    K = std::min(K, n);
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return est[a] > est[b]; });
    std::vector<std::vector<int>> branch(K);
    std::vector<double> load(K, 0.0);
    for (int i : order) {
        int b = 0;
        for (int q = 1; q < K; ++q)
            if (load[q] < load[b]) b = q;
        branch[b].push_back(i);
        load[b] += est[i];
    }
    // stages of each branch: one context, or two small ones side by side
    std::vector<std::vector<std::vector<int>>> stages(K);
    for (int b = 0; b < K; ++b) {
        const auto& br = branch[b];
        for (size_t q = 0; q < br.size(); ++q) {
            if (q + 1 < br.size() && small_workgroup(m->ctxs[br[q]]) && small_workgroup(m->ctxs[br[q + 1]])) {
                stages[b].push_back({br[q], br[q + 1]});
                ++q;
            } else {
                stages[b].push_back({br[q]});
            }
        }
        load[b] = 0;
        for (const auto& st : stages[b]) load[b] += st.size() == 2 ? std::max(est[st[0]], est[st[1]]) : est[st[0]];
    }
    {
        char buf[96];
        for (int b = 0; b < K; ++b) {
            std::snprintf(buf, sizeof buf, "%sbranch %d (est %.0f ms):", b ? "; " : "", b, load[b]);
            m->plan += buf;
            for (const auto& st : stages[b])
                for (size_t q = 0; q < st.size(); ++q) {
                    const int i = st[q];
                    std::snprintf(buf, sizeof buf, "%s ctx %d (N=%d, %d wg, est %.0f ms)%s",
                                  st.size() == 2 ? (q == 0 ? " [" : " |") : "", i, m->ctxs[i]->N,
                                  integration_wgs(m->ctxs[i]), est[i], st.size() == 2 && q == 1 ? " ]" : "");
                    m->plan += buf;
                }
        }
    }
    // capture: the first context's stream forks to one stream per branch
    // (its first context's), each branch enqueues its contexts in order, and
    // all join back
    hipStream_t s0 = m->ctxs[0]->st;
    hipGraph_t gph = nullptr;
    if (hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return bad(fail(BNM_E_HIP, "hipStreamBeginCapture failed"));
    int r = BNM_OK;
    hipError_t e = hipEventRecord(m->fork, s0);
    std::vector<hipStream_t> bst(K);
    for (int b = 0; b < K; ++b) bst[b] = m->ctxs[branch[b][0]]->st;
    for (int b = 0; b < K && e == hipSuccess; ++b)
        if (bst[b] != s0) e = hipStreamWaitEvent(bst[b], m->fork, 0);
    for (int b = 0; b < K && e == hipSuccess && !r; ++b)
        for (const auto& st : stages[b]) {
            if (r || e != hipSuccess) break;
            if (st.size() == 1) {
                r = enqueue_run(m->ctxs[st[0]], bst[b], false);
                continue;
            }
            // the second context on its own stream, forked from and joined to
            // the branch's stream around the first
            hipEvent_t f = nullptr, jn = nullptr;
            e = hipEventCreateWithFlags(&f, hipEventDisableTiming);
            if (e == hipSuccess) m->pair_events.push_back(f);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&jn, hipEventDisableTiming);
            if (e == hipSuccess) m->pair_events.push_back(jn);
            hipStream_t s2 = m->ctxs[st[1]]->st;
            if (e == hipSuccess) e = hipEventRecord(f, bst[b]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s2, f, 0);
            if (e == hipSuccess) r = enqueue_run(m->ctxs[st[0]], bst[b], false);
            if (e == hipSuccess && !r) r = enqueue_run(m->ctxs[st[1]], s2, false);
            if (e == hipSuccess && !r) e = hipEventRecord(jn, s2);
            if (e == hipSuccess && !r) e = hipStreamWaitEvent(bst[b], jn, 0);
        }
    for (int b = 0; b < K && e == hipSuccess && !r; ++b) {
        if (bst[b] == s0) continue;
        e = hipEventRecord(m->joins[b], bst[b]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s0, m->joins[b], 0);
    }
    const hipError_t e2 = hipStreamEndCapture(s0, &gph);
    if (r || e != hipSuccess || e2 != hipSuccess) {
        if (gph) (void)hipGraphDestroy(gph);
        return bad(r ? r : fail(BNM_E_HIP, "graph capture failed"));
    }
    e = hipGraphInstantiate(&m->exec, gph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gph);
    if (e != hipSuccess) return bad(fail(BNM_E_HIP, "hipGraphInstantiate failed"));
    for (bnm_ctx* x : m->ctxs) m->sc_gen.push_back(x->sc_gen);
    set(BNM_OK);
    return m;
}

int bnm_multi_run(bnm_multi* m) {
    if (!m) return fail(BNM_E_ARG, "null handle");
    for (size_t i = 0; i < m->ctxs.size(); ++i)
        if (m->ctxs[i]->sc_gen != m->sc_gen[i])
            return fail(BNM_E_ARG, "a context's new SC changed its kernel shape after capture; create the multi again");
    hipStream_t s0 = m->ctxs[0]->st;
    HIPCHK(hipEventRecord(m->t0, s0));
    HIPCHK(hipGraphLaunch(m->exec, s0));
    HIPCHK(hipEventRecord(m->t1, s0));
    HIPCHK(hipStreamSynchronize(s0));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, m->t0, m->t1));
    m->ms = ms;
    return BNM_OK;
}

double bnm_multi_last_ms(bnm_multi* m) { return m ? m->ms : -1; }

const char* bnm_multi_plan(bnm_multi* m) { return m ? m->plan.c_str() : ""; }

void bnm_multi_destroy(bnm_multi* m) {
    if (!m) return;
    if (m->exec) (void)hipGraphExecDestroy(m->exec);
    for (hipEvent_t e : {m->fork, m->t0, m->t1})
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : m->pair_events)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : m->joins)
        if (e) (void)hipEventDestroy(e);
    delete m;
}

const char* bnm_ctx_tune_log(bnm_ctx* x) { return x ? x->tune_log.c_str() : ""; }

int bnm_ctx_get_outputs(bnm_ctx* x, double* corr, double* pen, double* wIE_fic, double* BOLD,
                        double* S_E, double* I_E, double* r_E, double* S_I, double* I_I,
                        double* r_I, int* failed, int* trials, double* fc, double* fcd) {
    if (!x) return fail(BNM_E_ARG, "null context");
    const size_t B = x->B, N = x->N;
    if (corr) HIPCHK(hipMemcpy(corr, x->corr, B * sizeof(double), hipMemcpyDeviceToHost));
    if (pen) HIPCHK(hipMemcpy(pen, x->pen, B * sizeof(double), hipMemcpyDeviceToHost));
    if (wIE_fic && x->cfg.do_fic)
        HIPCHK(hipMemcpy(wIE_fic, x->wIEadj, B * N * sizeof(double), hipMemcpyDeviceToHost));
    if (failed) HIPCHK(hipMemcpy(failed, x->failed, B * sizeof(int), hipMemcpyDeviceToHost));
    if (trials) HIPCHK(hipMemcpy(trials, x->trials, B * sizeof(int), hipMemcpyDeviceToHost));
    double* ext_out[6] = {S_E, I_E, r_E, S_I, I_I, r_I};
    bool any_ext = false;
    for (double* p : ext_out) any_ext = any_ext || p;
    if (any_ext) {
        std::vector<double> e(B * 6 * N);
        HIPCHK(hipMemcpy(e.data(), x->ext, e.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int i = 0; i < 6; ++i)
            if (ext_out[i])
                for (size_t b = 0; b < B; ++b)
                    std::memcpy(ext_out[i] + b * N, e.data() + (b * 6 + i) * N, N * sizeof(double));
    }
    auto fetch = [&](double* dst, const void* src, size_t n, bool is64) -> int {
        if (is64) {
            HIPCHK(hipMemcpy(dst, src, n * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(n);
            HIPCHK(hipMemcpy(t.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) dst[i] = t[i];
        }
        return BNM_OK;
    };
    int r;
    if (BOLD && (r = fetch(BOLD, x->bold, B * x->ots * N, x->f64))) return r;
    if (!x->cfg.sim_only) {
        if (fc && (r = fetch(fc, x->fc, B * x->np, x->fc64))) return r;
        if (fcd && !x->cfg.no_fcd && (r = fetch(fcd, x->fcd, B * x->nwp, x->fc64))) return r;
    }
    return BNM_OK;
}

int bnm_ctx_get_hemi_fcd(bnm_ctx* x, double* fcd_left, double* fcd_right) {
    if (!x) return fail(BNM_E_ARG, "null context");
    if (!x->cfg.save_hemis || x->cfg.sim_only || x->cfg.no_fcd || !x->fcd_h[0])
        return fail(BNM_E_ARG, "the context was not created with save_hemis (and FCD)");
    if (x->n_hemi < 2) return fail(BNM_E_ARG, "a hemisphere needs at least 3 nodes");
    const size_t n = (size_t)x->B * x->nwp;
    double* dst[2] = {fcd_left, fcd_right};
    for (int h = 0; h < 2; ++h) {
        if (!dst[h]) continue;
        if (x->fc64) {
            HIPCHK(hipMemcpy(dst[h], x->fcd_h[h], n * sizeof(double), hipMemcpyDeviceToHost));
        } else {
            std::vector<float> t(n);
            HIPCHK(hipMemcpy(t.data(), x->fcd_h[h], n * sizeof(float), hipMemcpyDeviceToHost));
            for (size_t i = 0; i < n; ++i) dst[h][i] = t[i];
        }
    }
    return BNM_OK;
}

int bnm_ctx_timings(bnm_ctx* x, double* ms, int n) {
    if (!x || !ms) return fail(BNM_E_ARG, "null argument");
    const int k = n < 6 ? n : 6;
    for (int i = 0; i < k; ++i) ms[i] = x->timings[i];
    return k;
}

const char* bnm_ctx_kernel_name(bnm_ctx* x) { return x ? x->kname.c_str() : ""; }

namespace {
std::mutex g_scratch_mu;             // serialises no-allocation calls in this process
void* g_scratch_ptr[256] = {};       // scratch address per device ordinal

// Device address of the current device's scratch array (taken once per
// device, with that device current).  Called with g_scratch_mu held.
int scratch_base(void** out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) return fail(BNM_E_UNSUP, "device ordinal beyond 255");
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(BNM_E_HIP, "hipGetSymbolAddress returned a null scratch address");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return BNM_OK;
}
}  // namespace

int bnm_run_simulations_noalloc(const bnm_config* cfg, double* corr, double* fic_penalties,
                                const double* G_list, const double* w_EE_list, const double* w_EI_list,
                                const double* w_IE_list, double* w_IE_fic, const double* SC,
                                const float* noise, long long noise_len, const double* emp_FC_tril,
                                int n_emp_fc, const double* emp_FCD_tril, int n_emp_fcd, double* BOLD,
                                double* S_E, double* I_E, double* r_E, double* S_I, double* I_I,
                                double* r_I, int* FIC_failed, int* fic_n_trials, double* fc_trils,
                                double* fcd_trils) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (int r = check_device()) return r;
    void* base = nullptr;
    if (int r = scratch_base(&base)) return r;
    int st = BNM_OK;
    bnm_ctx* x = ctx_create(cfg, &st, static_cast<unsigned char*>(base), BNM_NOALLOC_SCRATCH_BYTES);
    if (!x) return st;
    int r = bnm_ctx_set_sc(x, SC);
    if (!r) r = bnm_ctx_set_noise(x, noise, noise_len);
    if (!r) r = bnm_ctx_set_empirical(x, emp_FC_tril, n_emp_fc, emp_FCD_tril, n_emp_fcd);
    if (!r) r = bnm_ctx_set_params(x, G_list, w_EE_list, w_EI_list, w_IE_list, w_IE_fic);
    if (!r) r = bnm_ctx_run(x);
    if (!r)
        r = bnm_ctx_get_outputs(x, corr, fic_penalties, w_IE_fic, BOLD, S_E, I_E, r_E, S_I, I_I,
                                r_I, FIC_failed, fic_n_trials, fc_trils, fcd_trils);
    bnm_ctx_destroy(x);
    return r;
}

int bnm_run_simulations(const bnm_config* cfg, double* corr, double* fic_penalties,
                        const double* G_list, const double* w_EE_list, const double* w_EI_list,
                        const double* w_IE_list, double* w_IE_fic, const double* SC,
                        const float* noise, long long noise_len, const double* emp_FC_tril,
                        int n_emp_fc, const double* emp_FCD_tril, int n_emp_fcd, double* BOLD,
                        double* S_E, double* I_E, double* r_E, double* S_I, double* I_I,
                        double* r_I, int* FIC_failed, int* fic_n_trials, double* fc_trils,
                        double* fcd_trils) {
    int st = BNM_OK;
    bnm_ctx* x = bnm_ctx_create(cfg, &st);
    if (!x) return st;
    int r = bnm_ctx_set_sc(x, SC);
    if (!r) r = bnm_ctx_set_noise(x, noise, noise_len);
    if (!r) r = bnm_ctx_set_empirical(x, emp_FC_tril, n_emp_fc, emp_FCD_tril, n_emp_fcd);
    if (!r) r = bnm_ctx_set_params(x, G_list, w_EE_list, w_EI_list, w_IE_list, w_IE_fic);
    if (!r) r = bnm_ctx_run(x);
    if (!r)
        r = bnm_ctx_get_outputs(x, corr, fic_penalties, w_IE_fic, BOLD, S_E, I_E, r_E, S_I, I_I,
                                r_I, FIC_failed, fic_n_trials, fc_trils, fcd_trils);
    bnm_ctx_destroy(x);
    return r;
}

}  // extern "C"
