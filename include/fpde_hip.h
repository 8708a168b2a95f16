/*
 * libfpde_hip.so -- C ABI of the MI355X (gfx950) history-sum solver.
 *
 * Plain C types only (pointers, sizes, ints, doubles); no torch or HIP types.
 * Every function returns an fpde_status; on failure fpde_last_error() holds
 * the message (thread-local).  Sizes are 64-bit throughout (the reference's
 * GPU layer uses int sizes, opencl_cuda_class.h:8,44-45).
 *
 * Three groups of entry points:
 *  1. the device-layer ABI behind the OpenCL_* facade classes of
 *     include/fpde_opencl_facade.h, replacing opencl_cuda_class.h:4-67 /
 *     opencl_cuda_base.cpp:20-826 for the reference's GPU history sum;
 *  2. direct kernels of the hot path (history sums, weights);
 *  3. the solver (1D / 3D / distributed order) and the batched request
 *     server, which keep all state on the device.
 */
#ifndef FPDE_HIP_H
#define FPDE_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FPDE_OK = 0,
    FPDE_ERR_ARG = 1,       /* invalid argument / shape                      */
    FPDE_ERR_HIP = 2,       /* HIP runtime error                             */
    FPDE_ERR_NOMEM = 3,     /* device allocation failed / over budget        */
    FPDE_ERR_NOTFOUND = 4,  /* unknown kernel name                           */
    FPDE_ERR_STATE = 5      /* call not valid in the object's current state  */
} fpde_status;

const char *fpde_last_error(void);
int fpde_version(void);
/* Release what the library itself holds across calls (the per-device
 * argument tables of fpde_solver_advance_batch) and wait for the devices it
 * used.  Call it after destroying every solver, server and facade context,
 * before the process exits, so that no device resource is left to the HIP
 * runtime's own teardown.  The library stays usable afterwards (the tables
 * are allocated again on demand).  No counterpart in the reference, whose
 * OpenCL_program destructor releases its objects (opencl_cuda_base.cpp:418-434). */
fpde_status fpde_shutdown(void);
/* Number of visible devices (0 when no GPU); never fails. */
int fpde_device_count(void);

/* ======================================================================
 * 1. Device layer behind the OpenCL_* facade (opencl_cuda_class.h:4-67).
 *    Kernels are precompiled for gfx950; a "program" is created from the
 *    reference's kernel source text only to read its #define constants
 *    (sum_alg, OS, BS: fpde_hpc_1d.cpp:1771-1775); kernels bind by name.
 * ====================================================================== */
typedef struct fpde_cl_ctx fpde_cl_ctx;
typedef struct fpde_cl_buf fpde_cl_buf;
typedef struct fpde_cl_prog fpde_cl_prog;
typedef struct fpde_cl_kernel fpde_cl_kernel;

/* OpenCL_program(int gpu) + create_queue (opencl_cuda_base.cpp:356-406, :448-452). */
fpde_status fpde_cl_ctx_create(int device, fpde_cl_ctx **out);
fpde_status fpde_cl_ctx_destroy(fpde_cl_ctx *ctx);
/* OpenCL_buffer(ct,flags,size,mem) (opencl_cuda_base.cpp:342-351); host_init may be NULL. */
fpde_status fpde_cl_buf_alloc(fpde_cl_ctx *ctx, size_t bytes, const void *host_init,
                              fpde_cl_buf **out);
fpde_status fpde_cl_buf_free(fpde_cl_buf *buf);
size_t fpde_cl_buf_size(const fpde_cl_buf *buf);
/* Device address of a buffer (for direct kernel calls below). */
void *fpde_cl_buf_device_ptr(const fpde_cl_buf *buf);
/* EnqueueWriteBuffer / EnqueueBuffer (opencl_cuda_base.cpp:210-228): blocking;
 * bytes == 0 means "whole buffer from offset". */
fpde_status fpde_cl_buf_write(fpde_cl_ctx *ctx, fpde_cl_buf *buf, size_t offset,
                              size_t bytes, const void *src);
fpde_status fpde_cl_buf_read(fpde_cl_ctx *ctx, fpde_cl_buf *buf, size_t offset,
                             size_t bytes, void *dst);
/* Device-to-device copy between buffers (the clEnqueueCopyBuffer analogue;
 * blocking).  Also the known-size copy that calibrates HBM byte counters. */
fpde_status fpde_cl_buf_copy(fpde_cl_ctx *ctx, fpde_cl_buf *dst, size_t dst_offset,
                             fpde_cl_buf *src, size_t src_offset, size_t bytes);
/* OpenCL_prg(ct,did,source) (opencl_cuda_base.cpp:55-179): parses #defines. */
fpde_status fpde_cl_program_create(fpde_cl_ctx *ctx, const char *source, fpde_cl_prog **out);
fpde_status fpde_cl_program_destroy(fpde_cl_prog *prog);
/* OpenCL_kernel(ct,prog,name) (opencl_cuda_base.cpp:181-197); name "Sum". */
fpde_status fpde_cl_kernel_create(fpde_cl_prog *prog, const char *name, fpde_cl_kernel **out);
fpde_status fpde_cl_kernel_destroy(fpde_cl_kernel *k);
/* SetArg(idx,size,ptr) / SetBufferArg(buf,idx) (opencl_cuda_base.cpp:198-209). */
fpde_status fpde_cl_kernel_set_arg(fpde_cl_kernel *k, int idx, size_t size, const void *value);
fpde_status fpde_cl_kernel_set_buffer_arg(fpde_cl_kernel *k, int idx, fpde_cl_buf *buf);
/* ExecuteKernel(k,ndims,dims,ldims) (opencl_cuda_base.cpp:229-321).  "Sum"
 * runs with the argument ABI of fr_ocl_call_sum (fpde_hpc_1d.cpp:1851-1868);
 * sum_alg 0/1 read 0 S, 1 old, 2 old_size, 3 tau, 6 kb_3_cache, 7 sum_param;
 * sum_alg 2 also 4 Htmp [nodes][sum_param], 5 row2_precalc, 8 gtj, 9 gt,
 * 10 alpha, 11 tstep, with the series coefficients of time_summing
 * (:4137-4151, _kb_row2_fixed_coefs :1159-1180). */
fpde_status fpde_cl_launch(fpde_cl_ctx *ctx, fpde_cl_kernel *k, int ndims,
                           const size_t *gdims, const size_t *ldims);
/* Finish() (opencl_cuda_base.cpp:322-341). */
fpde_status fpde_cl_finish(fpde_cl_ctx *ctx);

/* ======================================================================
 * 2. Direct hot-path kernels on device pointers (stream = default stream of
 *    the calling thread; all synchronous on return).
 * ====================================================================== */
/* The reference "Sum" kernel, sum_alg 0/1 (fpde_hpc_1d.cpp:1247-1284):
 * S[id] = sum_{t=old_size-2..0} kb[t]*(old[(t+1)*os+id]-old[t*os+id])/tau[t],
 * for id < nodes; sum_alg 1 stops at the first 16-level chunk whose weights
 * are all below sum_param.  old: [old_size][os] doubles, level-major. */
fpde_status fpde_reference_sum(double *S, const double *old, int64_t old_size,
                               int64_t os, int64_t nodes, const double *tau,
                               const double *kb, int sum_alg, double sum_param);

/* Plain FP64 streaming history sum on differences (the solver's plain path):
 * S[i] = sum_{j=1..n} omega[j-1]*D[(n-j)*ld + i], i < nodes. */
fpde_status fpde_history_sum_f64(double *S, const double *D, int64_t n, int64_t nodes,
                                 int64_t ld, const double *omega);

/* fpde_history_sum_f64 without any device allocation: the same kernels and
 * bitwise the same S, with the partial rows in a fixed 32 MiB device array
 * of the library (one per GPU, held until the process ends) instead of a
 * buffer allocated per call.  ld must be at most 65536 (else FPDE_ERR_ARG).
 * Like fpde_history_sum_f64 it returns after the device has finished.
 * No counterpart in the reference: it is for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid hipMalloc).
 * All calls on one GPU share the array, so calls that overlap on one GPU
 * (other threads or streams) race. */
fpde_status fpde_history_sum_f64_noalloc(double *S, const double *D, int64_t n, int64_t nodes,
                                         int64_t ld, const double *omega);

/* Space-fractional sum of the explicit scheme, calc_v (fpde_hpc_1d.cpp:2303-2335,
 * one-sided, with_i = 0), on the uniform node grid Z_k = k h:
 * V[i] = Dg * sum_{j=1}^{i-1} kb(gamma, j-1, j, i; Z) * F[j], i = 1..nodes-2,
 * V[0] = V[nodes-1] = 0, Dg = 1/Gamma(1-gamma).  weight_mode 1 takes the
 * reference's series for the intervals j = 1, 2 (stop at |term| <= series_eps,
 * kb dispatch :2265), 0 the closed form everywhere.  F, V: device arrays of
 * `nodes` doubles (F[0] and F[nodes-1] are not read); 0 < gamma < 1. */
fpde_status fpde_space_sum(double *V, const double *F, int64_t nodes, double h, double gamma,
                           int weight_mode, double series_eps);

/* Time-block history GEMM of the blocked solver paths, for B future steps:
 * P[r*ld + i] = sum_{t<tc} omega[n0+r-t-1] * D[t*ld + i], r < block, i < nodes,
 * computed in `mode` (1 FP64 MFMA, 2 split bf16, 3 split fp16+scaling,
 * 4 split bf16 x3; split modes need tc % 32 == 0).  D: [levels][ld] device
 * FP64 (levels >= tc rounded up to 32), omega: >= n0+block+512 entries, P:
 * [block][ld].  Temporary split storage is allocated internally. */
fpde_status fpde_history_block_gemm(int mode, const double *D, int64_t ld, int64_t nodes,
                                    int64_t levels, int n0, int tc, int block,
                                    const double *omega, double *P);

/* ======================================================================
 * 3. Solver
 * ====================================================================== */
typedef struct {
    int dims;            /* 1 or 3                                               */
    int nodes;           /* 1D: nodes on [0,length]; 3D: nodes per axis (n^3)    */
    int steps;           /* M                                                    */
    double tau;          /* uniform time step                                    */
    double alpha;        /* Caputo order (single order)                          */
    double length;       /* domain length                                        */
    double ksat;         /* conductivity scale K = ksat*Kr                       */
    int coef_mode;       /* 0 VGM (cv_form 4, k_form 1), 1 linear (K=1, 1/C=1)   */
    int picard_max;      /* >=1; 1 = lagged coefficients (the reference scheme)  */
    double picard_tol;   /* stop when max|U^k-U^{k-1}| <= tol (checked on device) */
    int weight_mode;     /* 0 closed-form L1, 1 reference series for t0<=1       */
    double series_eps;   /* the reference's global_eps (1e-5)                    */
    int do_q;            /* 0 single order; >0 distributed order with q nodes    */
    int do_rule;         /* 0 Gauss-Legendre, 1 trapezoid (fr_do.cpp)            */
    double do_a0, do_a1; /* order interval for distributed order                 */
    int hist_mode;       /* 0 plain FP64, 1 FP64 MFMA GEMM, 2 split bf16 hi+lo,
                            3 split fp16 hi+lo + scaling, 4 split bf16 x3        */
    int block;           /* future steps per GEMM block (multiple of 16, <=128)  */
    int step_mode;       /* 1D step launches: 0 = cooperative blocks (one launch
                            per history block, every Picard iteration inside,
                            grid barriers between iterations) for the GEMM and
                            series history forms when the device supports
                            cooperative launch, else per iteration;
                            1 = one launch per Picard iteration                 */
    int sum_alg;         /* time_summing form (fpde_hpc_1d.cpp:4086-4164):
                            0 full history sum;
                            1 restricted: newest intervals only, down to the
                              first weight with |kb| < sum_param (:4117-4131;
                              per order for distributed order);
                            2 series moments: P = sum_param terms of the
                              binomial series kept per node instead of the
                              levels (:4137-4151); single order, hist_mode 0,
                              P in [1, 64]                                      */
    double sum_param;    /* sum_alg 1: threshold on kb; sum_alg 2: terms P       */
    const double *taus;  /* NULL: uniform tau.  Else a prescribed non-uniform
                            grid: the M step lengths tau_0..tau_{M-1}
                            (Time[k+1] = Time[k] + tau_k, init_time_step
                            fpde_hpc_1d.cpp:3942-3949), copied at create; the
                            weights are then per step (not Toeplitz): 1D,
                            sum_alg 0; the GEMM forms take dense weight rows  */
    int var_tau;         /* 1: the step length is given step by step at run time
                            (fpde_solver_advance_tau, fpde_varT_step); 1D,
                            hist_mode 0, sum_alg 0, taus NULL; steps = capacity */
    int split3d;         /* 3D, the reference's split3d flag (rp_split_3d): even =
                            the history term only in the first (y, m3d_c=2) sweep;
                            odd = a third of it in each of the three sweeps
                            (Om1 fpde_hpc_1d.cpp:4172-4173, :4258-4263); 1D
                            ignores it                                          */
    double gamma;        /* space-fractional order (CLI "bb", H_solver gamma,
                            fpde_hpc_1d.cpp:3510): 0 or 1 = classical diffusion;
                            in (0,1) the reference's explicit one-sided scheme
                            (implicit=0, space_der=0): the rows carry
                            Dg kb(gamma,i-1,i,i) (A_/B/R, :3095-3142) and the
                            right-hand side -Dg inv_dw_dh calc_v(F(U^n))
                            (Om1 :4165-4189, calc_v :2303-2335); 1D only       */
    int fk;              /* kernel function of the generalized Caputo derivative
                            (CLI "FK", func_in_kernel, g() fpde_hpc_1d.cpp:
                            1889-1898): 0 g(t)=t, 1 sqrt(t), 2 t^2, 3 t^fp; the
                            L1 weights are then int (g(T_a)-g(s))^-alpha ds over
                            each interval, evaluated exactly (weight_mode 0,
                            sum_alg 0), one weight row per step                */
    double fp;           /* fk 3: the power (CLI "FP", func_power)             */
    int chunk_rows;      /* 1D: nodes per chunk of the partitioned line solve
                            (one workgroup each), 0 or 32..256.  0: 192, or
                            N/256 rounded up where 192-node chunks would number
                            more than 256 (the fastest single solve: three
                            waves of rows per workgroup).  256 packs more solves
                            into one batched launch (fpde_solver_advance_batch,
                            the server): ceil(N/chunk_rows) workgroups per
                            solve, at most 256; a batch holds 512.  Results
                            depend on the chunking
                            in the last bits only.  No counterpart in the
                            reference, whose line solve is one serial Thomas
                            sweep (fpde_hpc_1d.cpp:4672-4725)                 */
} fpde_spec;

typedef struct fpde_solver fpde_solver;

typedef struct {
    double ms_total;          /* device time of the last run() (hipEvents)    */
    double ms_history;        /* summed device time of history kernels        */
    double ms_step;           /* summed device time of step kernels           */
    int64_t history_launches;
    int64_t step_launches;
    double history_alg_bytes; /* algorithmic bytes of the history kernels (8 per term) */
    double history_alg_flops; /* algorithmic flops (2 per term)               */
    size_t device_bytes;      /* device memory held by the solver             */
    int picard_iters_total;   /* Picard iterations actually performed         */
    int reserved;
    double step_alg_bytes;    /* algorithmic bytes of the 1D step kernels: per
                                 node and step 100 B (coefficients, rows and
                                 the line solve) + 8 B per history term the
                                 step kernel sums itself (near field)          */
    int64_t steps_done;       /* time steps completed by the counted launches */
} fpde_stats;

/* Bytes of device memory a solve of this spec will hold (admission). */
size_t fpde_solver_footprint(const fpde_spec *spec);
fpde_status fpde_solver_create(const fpde_spec *spec, fpde_solver **out);
fpde_status fpde_solver_destroy(fpde_solver *s);
/* u0: host array of nodes (1D) or nodes^3 (3D, depth index fastest). */
fpde_status fpde_solver_set_initial(fpde_solver *s, const double *u0);
/* Restart the solve from u0 at step 0, keeping every device allocation. */
fpde_status fpde_solver_reset(fpde_solver *s, const double *u0);
/* Autotuner of the history block size B (the level count of one history
 * GEMM block and of the near field summed inside the step: 16, 32, 64 or 128)
 * for a GEMM-form spec (hist_mode 1-4; spec->block may be 0 here): each
 * candidate runs a trial solve from u0 of the first `trial_steps` steps
 * (<= 0: 2048; at least 257, at most spec->steps) on the device, timed from
 * enqueue to completion; the fastest B goes to *block_out and its trial time
 * per step (ms) to *ms_per_step_out (may be NULL).  Small B costs more
 * launches and GEMM blocks, large B a longer near-field sum per step, and the
 * history share grows with the step count: a trial of the whole run
 * (trial_steps = spec->steps) measures the real balance.  hist_mode 0 has no
 * block: FPDE_OK without a trial.  No counterpart in the reference: its
 * OpenCL block size is a command-line option (oclBS, fpde_hpc_1d.cpp:9713-9714)
 * compiled into the kernel text (:1771-1775). */
fpde_status fpde_autotune_block(const fpde_spec *spec, const double *u0, int trial_steps,
                                int *block_out, double *ms_per_step_out);

/* Advance `count` (1..32) independent solvers together, each by nsteps
 * (<= 0: to its last step), and return when all are done.  Solvers may differ
 * in N, M, alpha, weights, history form and block size.  The 1D solvers that
 * run cooperative block launches (step_mode 0 with a GEMM or series history)
 * share ONE cooperative launch per history block: their far-part kernels run
 * on their own streams, then a single k_step1d_batch launch runs the block's
 * steps of every one of them (each with its own barriers and Picard
 * decisions, so results equal those of fpde_solver_advance).  The others
 * advance on their own streams meanwhile.  FPDE_ERR_ARG when the batch's chunk
 * workgroups do not fit on the GPU at once.  For ensemble runs and parameter
 * sweeps, which the reference runs as one process per solve, configured on
 * the command line (mpi_test.sh:12-23). */
fpde_status fpde_solver_advance_batch(fpde_solver *const *solvers, int count, int nsteps);

/* Enqueue nsteps steps (or all remaining if nsteps<=0); asynchronous. */
fpde_status fpde_solver_advance(fpde_solver *s, int nsteps);
/* The same steps as fpde_solver_advance, enqueued as hipGraphs: every
 * `steps_per_graph` steps of launches (history kernels, step launches,
 * space-fractional sums) are stream-captured into one graph, the solver's
 * executable graph is updated in place when the topology repeats
 * (hipGraphExecUpdate) and launched.  Results are identical to
 * fpde_solver_advance.  For solvers whose steps launch per Picard iteration
 * (1D with step_mode 1 or hist_mode 0, and 3D); the cooperative block launches
 * of step_mode 0 are not captured (FPDE_ERR_STATE).  After a failure inside a
 * capture the solver needs fpde_solver_reset.  This is the graph-captured
 * step loop of the device layer; the reference launches kernel by kernel
 * (fr_ocl_call_sum, fpde_hpc_1d.cpp:1819-1883). */
fpde_status fpde_solver_advance_graph(fpde_solver *s, int nsteps, int steps_per_graph);
/* var_tau solvers: one step of length tau (Time[n+1] = Time[n] + tau,
 * init_time_step fpde_hpc_1d.cpp:3942-3949), completed before returning to the
 * stream (U^{n+1} and D[n] in memory). */
fpde_status fpde_solver_advance_tau(fpde_solver *s, double tau);
/* H_solver::copy (fpde_hpc_1d.cpp:5805-5986): dst takes src's state (solution,
 * history levels, time grid, step count).  Both var_tau solvers of the same
 * nodes; dst's capacity must hold src's steps. */
fpde_status fpde_solver_copy(fpde_solver *dst, const fpde_solver *src);
/* One outer step of the reference's variable time step (fpde_hpc_1d.cpp:
 * 8978-9022 with solve_T_half_step_and_check :8778-8793): s3 keeps s0's state,
 * s0 takes one step of *tau, s2 two steps of *tau/2, err = sum_i (U0_i - U2_i)^2
 * (solution_difference :8723-8749); while err > eps both restart from s3 with
 * *tau halved (below 1e-10 the first tau is taken as it is); else, while
 * err <= eps and *tau < tau_max, both restart with *tau doubled.  On return s0
 * holds the last step taken, *tau its length and *err its error. */
fpde_status fpde_varT_step(fpde_solver *s0, fpde_solver *s2, fpde_solver *s3, double *tau,
                           double eps, double tau_max, double *err);
fpde_status fpde_solver_sync(fpde_solver *s);
/* 1 when every enqueued launch of the solver has completed, 0 otherwise
 * (non-blocking), -1 on error. */
int fpde_solver_query(fpde_solver *s);
/* Steps completed so far. */
int fpde_solver_step(const fpde_solver *s);
fpde_status fpde_solver_get_solution(fpde_solver *s, double *u_out);
/* Output files of the reference driver (fpde_hpc_1d.cpp:9498-9526, opened
 * at :8945-8947), for the current solution at time step*tau: appends to
 * path_txt one row per depth node, "t z u 0 0 theta" (1D) or
 * "t z x y u 0 0 theta" along the depth line through the reference's centre
 * x = y = (nodes-1)/2 (3D), printed with %g; the two zeros stand for the C
 * and T fields the moisture-only model does not have, theta is the VGM
 * wetness at P = z - u.  3D also appends the field to path_3d (NULL:
 * skipped) as the reference's b_U: (nodes+1)^3 FP64 values, depth index
 * fastest, then x, then y, with one halo slot at the end of each axis that
 * holds the initial value of the nearest node. */
fpde_status fpde_solver_write_results(fpde_solver *s, const char *path_txt, const char *path_3d);
/* Effective weights (local, omega[M], corr[2M]) the solver uses (host copies). */
fpde_status fpde_solver_get_weights(fpde_solver *s, double *local, double *omega, double *corr);
/* Per-kernel timing with hipEvents on the solver's stream: bit 0 = history
 * kernels (plain stream / pack+GEMM / combine), bit 1 = step kernels. */
fpde_status fpde_solver_set_profiling(fpde_solver *s, int mask);
fpde_status fpde_solver_get_stats(fpde_solver *s, fpde_stats *out);

/* Initial conditions of the workloads (host; same draws as the oracle's). */
fpde_status fpde_init_1d(int nodes, double length, uint64_t seed, double *u);
fpde_status fpde_init_3d(int nodes, double length, uint64_t seed, double *u);
/* ksat giving the mesh Fourier number K*tau/(C*h^2) = fo at suction -u_top. */
double fpde_ksat_for_fourier(double u_top, double h, double tau, double fo);

/* ---------------------------------------------------------------------
 * Batched request server (continuous batching on one GPU).
 * --------------------------------------------------------------------- */
typedef struct fpde_server fpde_server;
typedef struct {
    int64_t id;
    double submit_s, start_s, finish_s;   /* server clock (seconds)            */
    int status;                           /* 0 done, 1 rejected, 2 cancelled,
                                             3 failed (a device error or a grid-
                                             barrier time-out in its solve)    */
    int steps_done;
} fpde_result_info;

fpde_status fpde_server_create(size_t mem_budget_bytes, int max_active, fpde_server **out);
fpde_status fpde_server_destroy(fpde_server *srv);
/* Queue a request; u0 is copied.  Rejected (status 1) if its footprint
 * exceeds the whole budget. */
fpde_status fpde_server_submit(fpde_server *srv, const fpde_spec *spec, const double *u0,
                               int64_t *id_out);
fpde_status fpde_server_cancel(fpde_server *srv, int64_t id);
/* Run the scheduler, in rounds of 256 steps, until every queued request has
 * finished (or until the number of finished requests reaches stop_after, if
 * > 0). */
fpde_status fpde_server_run(fpde_server *srv, int64_t stop_after);
/* One scheduling round: admit, advance each active solve by up to `quantum`
 * steps, retire finished ones.  Returns number of active+queued requests. */
int fpde_server_poll(fpde_server *srv, int quantum);
fpde_status fpde_server_result(fpde_server *srv, int64_t id, fpde_result_info *info,
                               double *u_out, size_t u_len);
double fpde_server_clock(fpde_server *srv);

#ifdef __cplusplus
}
#endif
#endif
