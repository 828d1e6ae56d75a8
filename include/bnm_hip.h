/*
 * bnm_hip.h -- C ABI of the MI355X (gfx950) batched reduced-Wong-Wang
 * simulation library (libbnm_hip.so).
 *
 * The library integrates reduced Wong-Wang E/I node dynamics (Euler-Maruyama,
 * dt = 0.1 ms, coupling G*SC*S_E), adjusts per-node wIE by numerical feedback
 * inhibition control (FIC), samples Balloon-Windkessel BOLD at TR, and reduces
 * BOLD to FC (Pearson), sliding-window FCD and the KS-based goodness of fit, for
 * a batch of simulations that share one SC and one noise array.
 *
 * Every entry point is synchronous, plain C, takes flat arrays and returns a
 * status code (BNM_OK or a negative BNM_E_*); nothing calls exit().  Array
 * layouts are row-major:
 *   SC            [N*N]        row j = inputs of node j (SC[j*N+k] weights S_E[k])
 *   per-sim lists [B] or [B*N] (sim-major)
 *   noise         [P*10*N*2]   float N(0,1), P ms of draws, order
 *                              [ms][substep][node][E,I]; read at (ms mod P)
 *   BOLD          [B*output_ts*N] (time-major per sim)
 *   FC            [B*n_pairs]  lower triangle i>j row-major, optionally without
 *                              left-right pairs (nodes 0..N/2-1 = left)
 *   FCD           [B*n_window_pairs] lower triangle wi>wj row-major
 */
#ifndef BNM_HIP_H
#define BNM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BNM_OK 0
#define BNM_E_ARG (-1)      /* invalid argument or size                      */
#define BNM_E_HIP (-2)      /* a HIP runtime call failed                     */
#define BNM_E_NODEV (-3)    /* no gfx950 device visible                      */
#define BNM_E_SIZE (-4)     /* empirical FC/FCD size does not fit the config */
#define BNM_E_UNSUP (-5)    /* configuration not supported by the kernel    */

/* integration precision */
#define BNM_PREC_F64 0      /* FP64 state and coupling                      */
#define BNM_PREC_F32 1      /* FP32 state and coupling, FP64 FIC sums       */
#define BNM_PREC_BF16 2     /* FP32 state, BF16 coupling operands, FP32 acc */
#define BNM_PREC_BF16X3 3   /* FP32 state, coupling as three BF16 MFMAs of the
                               hi/lo split operands (hi*hi + lo*hi + hi*lo)  */

/* integration kernel */
#define BNM_KERNEL_AUTO 0   /* autotuned choice                               */
#define BNM_KERNEL_PLAIN 1  /* one simulation per workgroup, one lane per node,
                               SC read from global memory (the reference's
                               GPU design, bnm.cu:72-315)                      */
#define BNM_KERNEL_FAST 2   /* one simulation per workgroup, SC row slices in
                               registers, coupling split over waves            */
#define BNM_KERNEL_MFMA 3   /* 16 simulations per workgroup, coupling
                               SC[NxN] x S_E[Nx16] on MFMA (N <= 128)          */

/* FC / FCD Gram precision */
#define BNM_FC_F64 0
#define BNM_FC_F32 1
#define BNM_FC_BF16 2       /* BF16 operands, FP32 accumulation              */
#define BNM_FC_BF16X3 3     /* hi/lo BF16 split operands, three products     */

typedef struct bnm_config {
    int nodes;                 /* N                                               */
    int n_sims;                /* B                                               */
    int time_steps;            /* simulated ms; the loop runs ts = 0..time_steps  */
    int bold_tr;               /* TR in ms                                        */
    int n_vols_remove;         /* volumes dropped before FC/FCD and the means     */
    int window_size;           /* FCD window (window holds ws/2*2+1 volumes)      */
    int window_step;           /* FCD window step (volumes)                       */
    int exc_interhemispheric;  /* drop left-right pairs from FC/FCD               */
    int do_fic;                /* wIE comes from w_IE_fic (else from w_IE_list)   */
    int adjust_fic;            /* numerical FIC adjustment                        */
    int max_fic_trials;        /* restarts allowed before FIC is declared failed  */
    int i_sampling_start;      /* ms; mean I_E sampling window for FIC            */
    int i_sampling_end;        /* ms                                              */
    double init_delta;         /* initial wIE step                                */
    long long noise_period_ms; /* ms of draws held in the noise array             */
    int calculate_fic_penalty; /* fill fic_penalties                              */
    int sim_only;              /* skip FC/FCD/fit (cost = max_cost)               */
    int no_fcd;                /* skip FCD and KS                                 */
    int use_fc_diff;           /* add |mean(emp FC) - mean(sim FC)| to the cost   */
    double max_cost;           /* cost when FC/FCD has a NaN                      */
    double fic_penalty_scale;
    int fic_reject_failed;     /* +1 penalty when FIC failed                      */
    int analytical_fic;        /* 1: compute the analytical FIC wIE on the device
                                  into w_IE_fic before integrating; 0: take
                                  w_IE_fic as given                               */
    int precision;             /* BNM_PREC_*                                      */
    int fc_precision;          /* BNM_FC_*                                        */
    int kernel;                /* BNM_KERNEL_*                                    */
    int sims_per_block;        /* reserved (the MFMA kernel groups 16)            */
    int autotune;              /* kernel == AUTO: time the candidate kernels on a
                                  short window at the first run and keep the
                                  fastest (env BNM_AUTOTUNE)                      */
    int use_graph;             /* capture the whole run in a hipGraph at the first
                                  run and replay it (per-phase timings are then
                                  not available, only the total)                 */
    int save_hemis;            /* also compute left / right FCD over the first /
                                  last (N/2)(N/2-1)/2 edges of each window FC
                                  (bnm.cu:453-495, 519-550)                       */
    int bw_friston2003;        /* BOLD weights k1 = 7 rho, k2 = 2, k3 = 2 rho - 0.2
                                  (the reference's BW_FRISTON2003 build,
                                  constants.h:82-85) instead of the 3T values
                                  3.72, 0.527, 0.53 (:86-89); env
                                  BNM_BW_FRISTON2003                              */
} bnm_config;

/* Fills `cfg` with the reference defaults (constants.h:93-116), honouring the
 * same BNM_* environment variables (BNM_BOLD_REMOVE, BNM_MAX_COST,
 * BNM_I_SAMPLING_START/END, BNM_NUMERICAL_FIC, BNM_MAX_FIC_TRIALS_CMAES,
 * BNM_INIT_DELTA, BNM_USE_FC_DIFF, BNM_EXC_INTERHEM,
 * BNM_CMAES_FIC_PENALTY_SCALE, BNM_CMAES_FIC_REJECT_FAILED).  The caller then
 * sets nodes, n_sims, time_steps, bold_tr, window_size, window_step and
 * noise_period_ms; n_vols_remove is derived from bold_tr by bnm_finalize_config. */
int bnm_default_config(bnm_config* cfg);

/* Derives n_vols_remove = BNM_BOLD_REMOVE*1000/bold_tr (run_CMAES.cpp:483)
 * when it is negative, and validates the config. */
int bnm_finalize_config(bnm_config* cfg);

/* Derived sizes (run_CMAES.cpp:476-594). */
int bnm_output_ts(int time_steps, int bold_tr);
int bnm_n_pairs(int nodes, int exc_interhemispheric);
int bnm_n_windows(const bnm_config* cfg);
/* Window start/end volume indices (absolute, inclusive); arrays of
 * bnm_n_windows(cfg) ints. */
int bnm_window_bounds(const bnm_config* cfg, int* starts, int* ends);

/* Regional parameters of one simulation (calculate_params, run_CMAES.cpp:61-143):
 *   p_j = base * (1 + sum_m scales[m] * maps[m*N + j]),
 * then, when min_j p_j < 0.001, every p_j is shifted by (0.001 - min_j p_j).
 * maps [n_maps*N], scales [n_maps], out [N]. */
int bnm_regional_params(int nodes, int n_maps, const double* maps, double base,
                        const double* scales, double* out);

/* Analytical FIC wIE (fic.cpp:130-207 of the reference, called per particle by
 * run_CMAES.cpp:144-174) for n_sims parameter sets, on the device: per node the
 * Brent root of the inhibitory steady state on [0, 2] (relative tolerance
 * 1e-3), then wIE_j; an unstable set (J < 0 or no root) gets wIE = 1 at every
 * node and unstable[b] = 1.  SC [N*N], G [B], wEE/wEI/wIE_out [B*N],
 * unstable [B] (optional). */
int bnm_analytical_fic(int nodes, int n_sims, const double* SC, const double* G, const double* wEE,
                       const double* wEI, double* wIE_out, int* unstable);

/* Host noise array as the reference's driver makes it (run_CMAES.cpp:640-657):
 * `count` draws of std::normal_distribution<float>(0,1) over std::mt19937(seed). */
int bnm_make_noise(unsigned seed, long long count, float* out);

/* Large noise arrays (a full-length array holds (T+1)*10*N*2 draws, 902 M at
 * N=100 and 450 s) in parallel: draw i comes from chunk c = i / 2^20, whose
 * std::mt19937 is seeded with std::seed_seq{seed, c}, then
 * std::normal_distribution<float>(0,1).  The result does not depend on
 * `threads` (<= 0: all hardware threads).  No counterpart in the reference,
 * whose driver draws one sequential stream (run_CMAES.cpp:640-657). */
int bnm_make_noise_chunked(unsigned seed, long long count, float* out, int threads);

/* The nonzero lists the sparse integration kernel builds from an N x N SC
 * (row-major, FP64) at bnm_ctx_set_sc, for inspection (host only, no GPU
 * work).  lanes_per_node 1 or 2 (2: each lane takes half of the row's
 * k-range, ceil(N/2) nodes); fp64 selects the FP64 layout (8-byte steps)
 * over the FP32 one (4-byte steps, weights rounded to FP32).  *slots gets P,
 * the slots per lane (0 when the kernel does not apply: N > 512 / lanes,
 * no nonzero, or more nonzeros than it holds).  When P > 0 and W / D are
 * non-null they receive, for p < P and lane column l = sl * N + j (slice sl
 * of row j): W[p * lanes_per_node * N + l] the slot's weight and, in byte
 * p % 4 of D[(p / 4) * lanes_per_node * N + l], its step in bytes from the
 * previous slot's S_E (the first from the slice's first node).  No
 * counterpart in the reference. */
int bnm_sparse_layout(int nodes, const double* SC, int lanes_per_node, int fp64, int* slots, double* W,
                      unsigned* D);

/* Replaces run_simulations (GPU) of bnm.cu:594-617.  Host arrays in, host
 * arrays out; outputs that are NULL are not produced.
 *   corr[B]            goodness-of-fit cost 1 - r(FC) + KS(FCD) [+ |dmean FC|]
 *   fic_penalties[B]   FIC penalty (bnm.cu:752-767)
 *   G_list[B], w_EE_list[B*N], w_EI_list[B*N], w_IE_list[B*N]
 *   w_IE_fic[B*N]      in: FIC initial wIE (unless cfg->analytical_fic);
 *                      out: wIE after numerical adjustment
 *   SC[N*N], noise[noise_len], emp_FC_tril[n_emp_fc], emp_FCD_tril[n_emp_fcd]
 *   BOLD[B*output_ts*N]
 *   S_E, I_E, r_E, S_I, I_I, r_I [B*N]  means over TR samples after n_vols_remove
 *   FIC_failed[B], fic_n_trials[B]
 *   fc_trils[B*n_pairs], fcd_trils[B*n_window_pairs]                          */
int bnm_run_simulations(const bnm_config* cfg, double* corr, double* fic_penalties,
                        const double* G_list, const double* w_EE_list, const double* w_EI_list,
                        const double* w_IE_list, double* w_IE_fic, const double* SC,
                        const float* noise, long long noise_len,
                        const double* emp_FC_tril, int n_emp_fc,
                        const double* emp_FCD_tril, int n_emp_fcd,
                        double* BOLD, double* S_E, double* I_E, double* r_E,
                        double* S_I, double* I_I, double* r_I,
                        int* FIC_failed, int* fic_n_trials,
                        double* fc_trils, double* fcd_trils);

/* No counterpart in the reference.  The same run as bnm_run_simulations for
 * callers that cannot allocate device memory (code that prepares or runs
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc): every device buffer of the run (device copies of SC, noise,
 * parameters and the empirical target; BOLD, moments, FC, window FC, FCD,
 * KS bins, outputs) is carved out of one static device array of
 * BNM_NOALLOC_SCRATCH_BYTES per GPU, held from library load until the process
 * ends; the call makes no hipMalloc / hipFree, creates no stream, event or
 * graph, and picks the first candidate integration kernel without timing
 * windows (set autotune = 0 in bnm_run_simulations for the same choice).  A
 * configuration whose buffers exceed the array returns BNM_E_SIZE before any
 * kernel runs.  Calls from the threads of one process are serialised by a
 * lock; the array keeps the last call's contents and every call rewrites what
 * it reads, so results do not depend on earlier calls.  Calls that overlap on
 * one GPU from outside that lock (another copy of the library in the same
 * process) would race on the array.  Arguments and outputs as
 * bnm_run_simulations. */
#define BNM_NOALLOC_SCRATCH_BYTES (256u << 20)
int bnm_run_simulations_noalloc(const bnm_config* cfg, double* corr, double* fic_penalties,
                                const double* G_list, const double* w_EE_list, const double* w_EI_list,
                                const double* w_IE_list, double* w_IE_fic, const double* SC,
                                const float* noise, long long noise_len,
                                const double* emp_FC_tril, int n_emp_fc,
                                const double* emp_FCD_tril, int n_emp_fcd,
                                double* BOLD, double* S_E, double* I_E, double* r_E,
                                double* S_I, double* I_I, double* r_I,
                                int* FIC_failed, int* fic_n_trials,
                                double* fc_trils, double* fcd_trils);

/* ---- persistent context: device buffers kept across calls (a CMA-ES run
 * calls run_simulations once per generation with the same SC and noise) ---- */
typedef struct bnm_ctx bnm_ctx;

/* Allocates every device buffer for `cfg` on the current HIP device. */
bnm_ctx* bnm_ctx_create(const bnm_config* cfg, int* status);
void bnm_ctx_destroy(bnm_ctx* ctx);
/* Uploads inputs (host pointers). */
int bnm_ctx_set_sc(bnm_ctx* ctx, const double* SC);
int bnm_ctx_set_noise(bnm_ctx* ctx, const float* noise, long long noise_len);
int bnm_ctx_set_empirical(bnm_ctx* ctx, const double* emp_FC_tril, int n_emp_fc,
                          const double* emp_FCD_tril, int n_emp_fcd);
int bnm_ctx_set_params(bnm_ctx* ctx, const double* G_list, const double* w_EE_list,
                       const double* w_EI_list, const double* w_IE_list,
                       const double* w_IE_fic);
/* Runs the whole path on device-resident inputs; returns when done. */
int bnm_ctx_run(bnm_ctx* ctx);
/* Copies outputs to host (NULL outputs skipped). */
int bnm_ctx_get_outputs(bnm_ctx* ctx, double* corr, double* fic_penalties, double* w_IE_fic,
                        double* BOLD, double* S_E, double* I_E, double* r_E, double* S_I,
                        double* I_I, double* r_I, int* FIC_failed, int* fic_n_trials,
                        double* fc_trils, double* fcd_trils);
/* Left and right FCD [B*n_window_pairs] of the last run (save_hemis): the
 * reference's save_hemis outputs L_fcd_trils / R_fcd_trils (bnm.cu:498-550),
 * with n_pairs_hemi = (N/2)(N/2-1)/2.  With exc_interhemispheric on, the
 * first and last n_pairs_hemi edges are exactly the left-left and
 * right-right pairs. */
int bnm_ctx_get_hemi_fcd(bnm_ctx* ctx, double* fcd_left, double* fcd_right);
/* FC, FCD and fit of a given BOLD array (no integration): bnm_ctx_set_bold
 * copies BOLD [B*output_ts*N] (time-major per simulation, the layout
 * bnm_ctx_get_outputs returns) into the context, bnm_ctx_analyze computes
 * the moments, FC, window FC, FCD, KS and cost (no FIC penalty) with the
 * context's fc_precision.  Needs the empirical target.  Also the way to turn
 * empirical BOLD into emp_FC / emp_FCD.  No counterpart in the reference. */
int bnm_ctx_set_bold(bnm_ctx* ctx, const double* BOLD);
int bnm_ctx_analyze(bnm_ctx* ctx);
/* Device time (ms) of the last bnm_ctx_run, per phase:
 * [0] analytical FIC, [1] integration, [2] BOLD moments + FC + window FC,
 * [3] FCD, [4] fit (KS, Pearson, penalty), [5] whole run.
 * Returns the number of values written. */
int bnm_ctx_timings(bnm_ctx* ctx, double* ms, int n);
/* Name of the integration kernel variant the last run used, and its launch
 * shape (for reports). */
const char* bnm_ctx_kernel_name(bnm_ctx* ctx);

/* Integration variants the autotuner timed and their times (empty when it
 * did not run). */
const char* bnm_ctx_tune_log(bnm_ctx* ctx);

/* ---- several populations in one graph launch: the contexts (each with its
 * own SC, node count, duration and parameters) are packed into branches of
 * one captured hipGraph -- as many concurrent branches as the CUs hold
 * integration workgroups, longest populations first, each branch running its
 * contexts one after another -- and bnm_multi_run launches it once and waits.
 * Contexts must have their inputs set; outputs are read per context
 * afterwards.  No counterpart in the reference (one population per call). ---- */
typedef struct bnm_multi bnm_multi;
bnm_multi* bnm_multi_create(bnm_ctx** ctxs, int n, int* status);
int bnm_multi_run(bnm_multi* m);
double bnm_multi_last_ms(bnm_multi* m);
/* The branch plan (contexts per branch, workgroups, time estimates). */
const char* bnm_multi_plan(bnm_multi* m);
void bnm_multi_destroy(bnm_multi* m);

/* ---- CMA-ES over the unit box (cmaes.cpp of the reference: CMAES_Points,
 * Hansen 2006 / 2016 variants, IQR boundary penalty, early stop).  Caller maps
 * the unit box onto parameter ranges (run_CMAES.cpp:727-742). ---- */
typedef struct bnm_cmaes bnm_cmaes;
/* dim free parameters, lambda candidates, seed of the initial mean (SeedMW),
 * initial sigma (BNM_CMAES_SIGMA, 0.5), variant 6 (2006) or 16 (2016). */
bnm_cmaes* bnm_cmaes_create(int dim, int lambda, unsigned seed_mw, double sigma0, int variant);
void bnm_cmaes_destroy(bnm_cmaes* e);
/* current candidates clipped to the unit box, x[j*dim + i] */
int bnm_cmaes_ask(bnm_cmaes* e, double* x_feas);
/* costs of the current candidates (lower is better); penalized (optional)
 * receives cost + boundary penalty.  Returns 1 when the run stopped. */
int bnm_cmaes_tell(bnm_cmaes* e, const double* cost, double* penalized);
int bnm_cmaes_best(bnm_cmaes* e, double* x, double* raw_cost);
double bnm_cmaes_sigma(bnm_cmaes* e);
/* [mueeff, cc, csigma, ccov, dsigma, chiN, c1, cmue, weights[lambda]] */
int bnm_cmaes_strategy(bnm_cmaes* e, double* out, int n);
int bnm_cmaes_iteration(bnm_cmaes* e);
/* Pre-selection (no counterpart in the reference): a pool of n >= lambda
 * candidates from the current distribution, the first lambda being the
 * current sample and the rest drawn with `seed`; x_feas[j*dim + i] clipped.
 * bnm_cmaes_select then makes pool members idx[0..lambda) the sample that the
 * next bnm_cmaes_tell scores. */
int bnm_cmaes_ask_pool(bnm_cmaes* e, int n, unsigned seed, double* x_feas);
int bnm_cmaes_select(bnm_cmaes* e, const int* idx);

/* ---- small MLP trained on the device (dense layers, GELU, mean squared
 * error, Adam), for the learned FIC initialiser and the fit surrogate.  No
 * counterpart in the reference.  Layers: in -> hidden x depth -> out; the
 * flat parameter vector holds, per layer l, W_l [out_l x in_l] row-major then
 * b_l [out_l].  bf16 != 0: GEMM operands rounded to bfloat16 (FP32
 * accumulation, FP32 parameters and Adam state); 0: FP32 operands. ---- */
typedef struct bnm_mlp bnm_mlp;
bnm_mlp* bnm_mlp_create(int in_dim, int hidden, int depth, int out_dim, int bf16, unsigned seed,
                        int* status);
void bnm_mlp_destroy(bnm_mlp* m);
long bnm_mlp_num_params(bnm_mlp* m);
int bnm_mlp_get_params(bnm_mlp* m, float* out);
int bnm_mlp_set_params(bnm_mlp* m, const float* in);
/* Y[n x out] from X[n x in] (host arrays) */
int bnm_mlp_predict(bnm_mlp* m, const float* X, int n, float* Y);
/* mean squared error over n x out and its gradient (flat, host) */
int bnm_mlp_loss_grad(bnm_mlp* m, const float* X, const float* T, int n, float* loss, float* grad);
/* upload a training set, then Adam over shuffled mini-batches; loss[e] is
 * the mean batch loss of epoch e (optional) */
int bnm_mlp_set_data(bnm_mlp* m, const float* X, const float* T, int n);
int bnm_mlp_train(bnm_mlp* m, int epochs, int batch, float lr, unsigned seed, float* loss);
const char* bnm_mlp_last_error(void);

/* Last error message of this thread (static storage). */
const char* bnm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* BNM_HIP_H */
