/*
 * FP64 CPU oracle for the Caputo L1 history-sum path of sevaboh/FPDE_HPC.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * entry and bench.py's cpu_baseline leg may load it; the product library
 * (libfpde_hip.so) never links, loads or calls it.
 *
 * It restates, in plain C (C99 + OpenMP), the reference algorithm:
 *   - L1 weights b(t0,t1;a) = int_{x_t0}^{x_t1} (x_a - s)^(-alpha) ds
 *     (fpde_hpc_1d.cpp:1933-1996 closed form for g(t)=t, :2000-2131 Newton
 *     binomial series, :2167-2301 dispatch, :3942-3949 kkk0, :4069-4085
 *     kb_3_cache);
 *   - the history sum sum_t kb_3_cache[t]*(old[t+1]-old[t])/tau[t]
 *     (fpde_hpc_1d.cpp:4086-4164: sum_alg 0 full, node-outer / t-inner;
 *     sum_alg 1 restricted; sum_alg 2 series moments, :4137-4151 with
 *     _kb_row2_fixed_coefs :1159-1180);
 *   - van Genuchten-Mualem coefficients (fpde_hpc_1d.cpp:2493-2575, :2625-2657);
 *   - the tridiagonal step A_/B/R + al1/Om1/bt1/back-substitution
 *     (fpde_hpc_1d.cpp:3095-3142, :4002-4027, :4165-4268, :4491-4590, :4672-4725);
 *   - the distributed-order derivative (fr_do.cpp:51-113) with either the
 *     reference trapezoid in alpha or Gauss-Legendre nodes;
 *   - the 3D locally-one-dimensional step (fpde_hpc_1d.cpp:5144-5626, sweep
 *     order K -> M -> N, history only in the first sweep).
 *
 * Pinning status of each function is listed in the header of fpde_oracle.c.
 */
#ifndef FPDE_ORACLE_H
#define FPDE_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers: MT19937-64 + libstdc++ uniform_real_distribution ---- */
typedef struct { uint64_t mt[312]; int mti; } orc_mt64;
void     orc_mt64_seed(orc_mt64 *s, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *s);
double   orc_mt64_uniform(orc_mt64 *s, double a, double b);

/* ---- L1 weights (g(t)=t, FK=0) ---- */
double orc_kb_closed(double alpha, double p0, double p1, double pa);
double orc_kb_series(double alpha, double p0, double p1, double pa,
                     double eps, int max_niter, int *niter_out);
/* kb(alpha,t0,t0+1,a) over a time grid, reference dispatch (a3k == 1):
 * weight_mode 0 = closed form everywhere, 1 = reference (series for t0<=1). */
double orc_kb(double alpha, int t0, int a, const double *Time,
              int weight_mode, double eps);
/* kb_3_cache for step n (levels 0..n), n entries; returns kkk0. */
double orc_l1_weights(double alpha, int n, const double *Time,
                      int weight_mode, double eps, double *kb3);

/* ---- generalized kernel g (func_in_kernel fk: 0 t, 1 sqrt t, 2 t^2, 3 t^fp) ---- */
double orc_g(int fk, double fp, double t);
double orc_inv_g_der(int fk, double fp, double t, int n, double der_nm1);
/* the reference's Taylor series _kb_row (branch g(t1) <= gtj) */
double orc_kb_row_taylor(double alpha, int fk, double fp, double t0, double t1, double gtj,
                         double eps, int max_niter, int *niter_out);
/* int_{p0}^{p1} (pa^gp - s^gp)^-alpha ds by tanh-sinh quadrature */
double orc_kb_general(double alpha, double gp, double p0, double p1, double pa);
double orc_g_exponent(int fk, double fp);
double orc_kb_g(double alpha, int fk, double fp, int t0, int a, const double *Time,
                int weight_mode, double eps);
double orc_l1_weights_g(double alpha, int fk, double fp, int n, const double *Time,
                        int weight_mode, double eps, double *kb3);

/* ---- history sum: S[i] = sum_t kb3[t]*(old[t+1][i]-old[t][i])/tau[t];
 * sum_alg 1 stops (newest first) at the first |kb3[t]| < sum_param. */
void orc_time_summing(int nodes, int levels, const double *old,
                      const double *kb3, const double *tau, double *S,
                      int nthreads, int sum_alg, double sum_param);

/* ---- series-compressed history (time_summing sum_alg 2) ----
 * Coefficients the reference caches for the newest interval [p0,p1] at
 * a = pa (row2_precalc[time_i1][0], filled by _kb_row2): v3[i] for every
 * term the series evaluates (it stops after the first |term| <= eps).
 * Returns the number of terms K; the first min(K, cap) go to v3. */
int orc_series_row(double alpha, double p0, double p1, double pa, double eps,
                   int cap, double *v3);
/* One step of sum_alg 2 for every node (fpde_hpc_1d.cpp:4137-4151):
 * H[id*P+j] += p1^(j+1) * v3[j] * (u1[id]-u0[id]) for j < min(P, K), then
 * S[id] = _kb_row2_fixed_coefs(gtj, gtj^-alpha, H[id*P..], P) / tau. */
void orc_series_step(int nodes, int P, double *H, const double *v3, int K,
                     double p1, double gtj, double alpha, double tau,
                     const double *u0, const double *u1, double *S);

/* ---- VGM coefficients, P = suction head in metres ---- */
double orc_wetness(double P);
double orc_inv_dw_dh(double P);
double orc_kr(double P);

/* ---- Gauss-Legendre nodes/weights on [a,b] ---- */
void orc_gauss_legendre(int n, double a, double b, double *x, double *w);

/* ---- solver parameters (shared by the 1D and 3D drivers) ---- */
typedef struct {
    int    nodes;        /* 1D: nodes on [0,length]; 3D: nodes per axis   */
    int    steps;        /* M                                             */
    double tau;          /* uniform time step                             */
    double alpha;        /* Caputo order (single-order runs)              */
    double length;       /* domain length (1D depth, 3D cube side)        */
    double ksat;         /* saturated conductivity scale (nondimensional) */
    int    coef_mode;    /* 0 = VGM (cv_form 4, k_form 1), 1 = linear K=1, 1/C=1 */
    int    picard_max;   /* >=1; 1 = lagged coefficients (reference scheme) */
    double picard_tol;
    int    weight_mode;  /* 0 closed form, 1 reference series for t0<=1  */
    double series_eps;   /* global_eps of the reference (1e-5)           */
    int    do_q;         /* 0 = single order; >0 = distributed order, q nodes */
    int    do_rule;      /* 0 = Gauss-Legendre, 1 = trapezoid (fr_do.cpp)  */
    double do_a0, do_a1; /* order interval                                 */
    int    nthreads;
    int    sum_alg;      /* 0 full sum, 1 truncated (sum_alg 1), 2 series (sum_alg 2) */
    double sum_param;    /* sum_alg 1: threshold on |kb|; sum_alg 2: series terms P  */
    const double *taus;  /* NULL: uniform tau; else the M step lengths tau_0..tau_{M-1}
                            (Time[k+1] = Time[k] + tau_k, init_time_step :3942-3949) */
    double gamma;        /* space-fractional order: 0 or 1 = classical diffusion,
                            (0,1) = the explicit one-sided scheme (calc_v :2303-2335),
                            in 3D along every sweep line */
    int    split3d;      /* 3D, the reference's split3d flag (rp_split_3d): even = the
                            history term only in the first (y, m3d_c=2) sweep; odd =
                            a third of it in each of the three sweeps (Om1 :4172-4173,
                            :4258-4263) */
    int    fk;           /* kernel function g of the generalized Caputo derivative
                            (func_in_kernel, g() :1889-1898): 0 t, 1 sqrt(t), 2 t^2,
                            3 t^fp */
    double fp;           /* func_power of fk 3 */
} orc_params;

/* 1D solve: u (nodes) holds u0 on entry, u^M on exit.  u_top = u0[0]. */
int orc_solve_1d(const orc_params *p, double *u);
/* 3D LOD solve on nodes^3 (depth index fastest); u holds u0 / u^M. */
int orc_solve_3d(const orc_params *p, double *u);
/* The same solve; sweeps (3 * nodes^3, may be NULL) receives the field after
 * the y, the x and the z sweep of the last step, in that order. */
int orc_solve_3d_trace(const orc_params *p, double *u, double *sweeps);

/* Space-fractional sum of the explicit scheme (calc_v, :2303-2335, one-sided,
 * with_i = 0) on the uniform grid Z_k = k h of `nodes` nodes:
 *   V[i] = Dg * sum_{j=1}^{i-1} kb(gamma, j-1, j, i; Z) * F[j],  i = 1..nodes-2,
 * Dg = 1/Gamma(1-gamma); V[0] = V[nodes-1] = 0.  kb dispatch as orc_kb. */
void orc_space_v(int nodes, double h, double gamma, int weight_mode, double eps,
                 const double *F, double *V);
/* F[j] of Om1 equation 0 (:4178-4179) on a uniform grid, j = 1..nodes-2:
 * (K[j] (u[j+1]-u[j]) - K[j-1] (u[j]-u[j-1])) / h^2; F[0] = F[nodes-1] = 0. */
void orc_space_flux(int nodes, double h, const double *K, const double *u, double *F);

/* ---- fr_do.cpp: the distributed-order diffusion equation ----
 *   int_0^1 w(a) D^a u da = D u_xx + f(t,x),  x in [0, max_x], t in [0, max_t],
 *   u(0,t) = u(max_x,t) = 0, u(x,0) = u0(x)
 * (eq_solve, fr_do.cpp:733-775): Crank-Nicolson in space (eq_Ff :656-665,
 * eq_one_step :724-732), the alpha integral as the trapezoid over na nodes
 * a_i = i/(na-1) with WA_i = w(a_i)/(na-1) (fill_arrays :51-72, DO_Cder
 * :89-113), L1 weights bk(d1,d2,a) = d1^(1-a) - d2^(1-a) over Gamma(2-a),
 * local coefficient eq_b0 (:688-690), and the reference's block scheme
 * (history before the block at its start, in-block intervals added step by
 * step, :739-761).  example 0: w = Gamma(5-a), u0 = 0, f = sin(2 pi x)
 * (800 pi^2 t^4 + 2400 (t^4 - t^3)/ln t), exact u = 100 t^4 sin(2 pi x)
 * (:458-465, :626-633, eq_anal_u); example 1: w = a^p (1+p), u0 = x (1-x),
 * f = 0.  C: [nx][nt] node-major as the reference's eq_C (level j of node i
 * at C[i*nt + j]); returns the RMS error against the exact solution of
 * example 0 at the last level (the figure eq_solve prints), 0 for example 1. */
typedef struct {
    int    nx, na, tblock, nt;  /* eq_Nx, eq_Na, eq_tblock, eq_Nt (a multiple of tblock) */
    int    example;             /* eq_ex */
    double power;               /* eq_ex1_power */
    double D, max_x, max_t;     /* eq_D, eq_max_x, eq_max_t */
    int    gpu_pi;              /* 1: the source as the GPU program text forms it (eq_ff) */
} orc_doeq_params;
double orc_doeq_solve(const orc_doeq_params *p, double *C);
/* The distributed-order derivative of fr_do's DO_Cder on a grid function F
 * (levels 0..ntau-1, spacing tau): DF[t] for t in [t0, ntau) with history
 * terms j in [j0, min(t, maxt)) (maxt < 0: no limit), trapezoid over na
 * orders with weight w (example as above; example < 0: Gamma(3-a)). */
void orc_do_cder(int ntau, double tau, int na, int example, double power, const double *F, int maxt,
                 int t0, int j0, double *DF);

/* Effective (folded) weights used by the product for uniform tau:
 * local = coefficient of (U^{n+1}-U^n); omega[j-1] for j=1..M (Toeplitz),
 * corr[2n], corr[2n+1] = corrections of the t=0,1 weights at step n.       */
void orc_folded_weights(const orc_params *p, double *local, double *omega,
                        double *corr);

/* Thomas solve of a_i x_{i-1} + b_i x_i + c_i x_{i+1} = d_i (tests). */
void orc_thomas(int n, const double *a, const double *b, const double *c,
                const double *d, double *x);

/* Initial conditions of the workloads. */
void orc_init_1d(int nodes, double length, uint64_t seed, double *u);
void orc_init_3d(int nodes, double length, uint64_t seed, double *u);
/* ksat giving mesh Fourier number K*tau/(C*h^2) = fo at the top node. */
double orc_ksat_for_fourier(double u_top, double h, double tau, double fo);

#ifdef __cplusplus
}
#endif
#endif
