/*
 * FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY (see fpde_oracle.h).
 *
 * Every function restates the reference algorithm and cites the reference
 * file:line it follows.  How each part is pinned (tests/test_oracle.py):
 *   orc_mt64_*          : C++ standard known answer (10000th output of the
 *                         default-seeded mt19937_64) + libstdc++ canonical form.
 *   orc_kb_closed       : telescoping identity, alpha->1 limit, exact integral
 *                         by scipy.integrate.quad.
 *   orc_kb_series       : converges to orc_kb_closed as eps->0 (both equal the
 *                         integral); truncation at |term|<=eps checked against
 *                         the series' own tail bound.
 *   orc_time_summing    : L1 is exact for u linear in t (D^a t = t^(1-a)/G(2-a)),
 *                         O(tau^(2-a)) convergence for u=t^2, brute force loop.
 *   orc_wetness         : the van Genuchten retention curve evaluated in numpy.
 *   orc_inv_dw_dh       : a numerical derivative of the wetness curve and the
 *                         analytic derivative of the van Genuchten curve.
 *   orc_kr              : the Mualem integral Se^(1/2) [int_0^Se dx/h /
 *                         int_0^1 dx/h]^2 over the inverted retention curve,
 *                         by scipy.integrate.quad.
 *   orc_thomas          : numpy.linalg.solve on random diagonally dominant systems.
 *   orc_gauss_legendre  : numpy.polynomial.legendre.leggauss.
 *   orc_solve_1d        : linear mode vs the Mittag-Leffler eigen-solution
 *                         (convergence in tau); a lagged VGM step satisfies
 *                         the rows at U^n; a Picard step run to tol 1e-14
 *                         satisfies the rows assembled in numpy with K and 1/C
 *                         at its own result (and the lagged step does not).
 *   orc_solve_3d        : a 3D field constant in x,y reduces to orc_solve_1d
 *                         (LOD sweeps along x,y are then identities); every
 *                         sweep of a VGM step with a non-zero history term
 *                         satisfies its own rows assembled in numpy (which
 *                         sweep carries the history, split3d even and odd, the
 *                         field each sweep takes its coefficients from).
 *   DO path             : closed form int rho(a) D^a t^2 da = (t^1.9-t)/ln t.
 *   orc_space_v         : for F = 1 the closed-form weights telescope to
 *                         Dg (Z_i^(1-g) - h^(1-g))/(1-g); for random F it equals
 *                         Dg * int_0^{Z_{i-1}} (Z_i-s)^(-g) f(s) ds of the
 *                         piecewise-constant f = F[j] on (Z_{j-1}, Z_j) by
 *                         scipy.integrate.quad; the series branch (j <= 2)
 *                         agrees with the same integrals to the series tolerance.
 *   space-fractional solve : one step equals the dense system assembled and
 *                         solved independently (numpy, quad weights); as
 *                         gamma -> 1 the scheme tends to the classical one
 *                         (Dg kb(g,i-1,i,i) -> 1, the explicit sum -> 0).
 *   reference-series mode weights (t0<=1): parity unpinned against the
 *                         compiled reference (it needs fftpack.h, absent in this
 *                         image); pinned only by the series/closed-form limit.
 */
#include "fpde_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* MT19937-64 (Matsumoto & Nishimura 2004), the engine behind std::mt19937_64 */
#define MT_NN 312
#define MT_MM 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

void orc_mt64_seed(orc_mt64 *s, uint64_t seed)
{
    s->mt[0] = seed;
    for (int i = 1; i < MT_NN; i++)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = MT_NN;
}

uint64_t orc_mt64_next(orc_mt64 *s)
{
    static const uint64_t mag01[2] = {0ULL, MT_MATRIX_A};
    uint64_t x;
    if (s->mti >= MT_NN) {
        int i;
        for (i = 0; i < MT_NN - MT_MM; i++) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + MT_MM] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        for (; i < MT_NN - 1; i++) {
            x = (s->mt[i] & MT_UM) | (s->mt[i + 1] & MT_LM);
            s->mt[i] = s->mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        }
        x = (s->mt[MT_NN - 1] & MT_UM) | (s->mt[0] & MT_LM);
        s->mt[MT_NN - 1] = s->mt[MT_MM - 1] ^ (x >> 1) ^ mag01[(int)(x & 1ULL)];
        s->mti = 0;
    }
    x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

/* std::uniform_real_distribution<double>(a,b) as libstdc++ evaluates it:
 * canonical = double(x) / 2^64 (clamped below 1), value = canonical*(b-a)+a. */
double orc_mt64_uniform(orc_mt64 *s, double a, double b)
{
    double c = (double)orc_mt64_next(s) / 18446744073709551616.0;
    if (c >= 1.0) c = nextafter(1.0, 0.0);
    return c * (b - a) + a;
}

/* ------------------------------------------------------------------------ */
/* L1 weights for the kernel function g(t)=t (func_in_kernel 0).            */

/* _kb_row, fpde_hpc_1d.cpp:1933-1996, branch g(t1)<=g(a).  For g(t)=t the
 * Taylor series stops after its first term (inv_g_der(.,n>=2)=0, :1929), so
 * the value is a^(1-alpha)*(|1-p0/a|^(1-alpha)-|1-p1/a|^(1-alpha))/(1-alpha). */
double orc_kb_closed(double alpha, double p0, double p1, double pa)
{
    if (alpha == 1.0) return (p1 == pa) ? 1.0 : 0.0;          /* :1939-1942 */
    double v = pow(pa, 1.0 - alpha);                         /* :1945-1946 */
    double v3 = pow(fabs(1.0 - p1 / pa), 1.0 - alpha);       /* :1948 */
    double v4 = pow(fabs(1.0 - p0 / pa), 1.0 - alpha);       /* :1949 */
    double v2 = -(v3 - v4) / (1.0 - alpha);                  /* :1950 */
    return v2 * v;
}

/* _kb_row2, fpde_hpc_1d.cpp:2000-2131, mode 0 (g(t1)<=g(a)), g(t)=t:
 * Newton binomial series of (a-s)^(-alpha) integrated over [p0,p1]; term i is
 * (alpha)_i/i! * a^(-alpha) * p1*(p1/a)^i * (1-(p0/p1)^(i+1))/(i+1); the loop
 * stops after the first term with |term| <= eps (:2128). */
double orc_kb_series(double alpha, double p0, double p1, double pa,
                     double eps, int max_niter, int *niter_out)
{
    double sum = 0.0, v1 = 1.0, v2, v3, v5, v6, term;
    double i = 0.0;
    int niter = 0;
    if (alpha == 1.0) { if (niter_out) *niter_out = 0; return (p1 == pa) ? 1.0 : 0.0; }
    v2 = pow(pa, -alpha) * p1;                                /* :2031 */
    v6 = p0 / p1;                                             /* :2032 */
    do {
        if (i != 0.0) {
            v1 *= -(-alpha - i + 1.0) / i;                    /* :2045 */
            v2 *= p1 / pa;                                    /* :2049 */
            v6 *= p0 / p1;                                    /* :2050 */
        }
        if (i == 0.0)
            v3 = (p1 - p0) / p1;                              /* :2070 */
        else {
            /* inv_g_der(g(t1),1,.)=1 and all higher derivatives are 0 for
             * g(t)=t, so the inner series (:2082-2113) keeps its first term
             * and runs one iteration whenever that term is non-zero. */
            v5 = (1.0 / (i + 1.0)) * (1.0 - v6);              /* :2078 */
            v3 = v5;
            if (fabs(v3) > fabs(v3) * eps) niter++;
        }
        term = v1 * v2 * v3;
        sum += term;                                          /* :2123 */
        i += 1.0;
        niter++;
        if (niter > max_niter) break;
    } while (fabs(term) > eps);
    if (niter_out) *niter_out = niter;
    return sum;
}

/* kb(alpha,t0,t0+1,a,&Time[0],&Time), fpde_hpc_1d.cpp:2167-2301, called with
 * a time grid: never cached (:2292 needs time==NULL); with sum_alg!=2 the
 * series is used when t0<=a3k && t1!=a && t0!=a (:2265).  a3k starts at 1
 * (:3457) and, for g(t)=t, is never changed: _kb_row reports niter2<=1 while
 * _kb_row2 reports niter>=1, so the test at :2272 never increments it and the
 * decrement at :2275 requires a3k!=1. */
double orc_kb(double alpha, int t0, int a, const double *Time,
              int weight_mode, double eps)
{
    int t1 = t0 + 1;
    if (alpha == 1.0) return (t1 == a) ? 1.0 : 0.0;           /* :2169-2175 */
    double p0 = Time[t0], p1 = Time[t1], pa = Time[a];
    if (weight_mode == 1 && t0 <= 1 && t1 != a && t0 != a)
        return orc_kb_series(alpha, p0, p1, pa, eps, 100000, NULL);
    return orc_kb_closed(alpha, p0, p1, pa);
}

/* ------------------------------------------------------------------------ */
/* Generalized Caputo kernel (func_in_kernel 1/2/3).                        */

/* g() of fpde_hpc_1d.cpp:1889-1898: t, sqrt(t), t^2, t^fp. */
double orc_g(int fk, double fp, double t)
{
    if (fk == 1) return sqrt(t);
    if (fk == 2) return t * t;
    if (fk == 3) return pow(t, fp);
    return t;
}

/* inv_g_der of :1901-1930: the n-th derivative of f = g^-1 at t, given the
 * (n-1)-th as der_nm1 (the recurrence the Taylor series runs on). */
double orc_inv_g_der(int fk, double fp, double t, int n, double der_nm1)
{
    if (fk == 1) {                                   /* f = t^2 */
        if (n == 0) return t * t;
        if (n == 1) return 2 * t * (der_nm1 / (t * t));
        if (n == 2) return 2 * (der_nm1 / (2 * t));
        return 0.0;
    }
    if (fk == 2) {                                   /* f = sqrt(t) */
        if (n == 0) return sqrt(t);
        if (n == 1) return (0.5 / sqrt(t)) * (der_nm1 / sqrt(t));
        return der_nm1 * (0.5 - n + 1) / t;
    }
    if (fk == 3) {                                   /* f = t^(1/fp) */
        if (n == 0) return pow(t, 1.0 / fp);
        if (n == 1) return ((1.0 / fp) * pow(t, (1.0 / fp) - 1)) * (der_nm1 / pow(t, 1.0 / fp));
        return der_nm1 * ((1.0 / fp) - n + 1) / t;
    }
    return (n == 0) ? t : ((n == 1) ? (der_nm1 / t) : 0.0);
}

/* _kb_row, :1933-1996, branch g(t1) <= gtj: Taylor series of f' = (g^-1)'
 * around gtj = g(a), integrated term by term over [g(t0), g(t1)]; stops after
 * the first |term| <= eps (the reference's global_eps) or max_niter terms. */
double orc_kb_row_taylor(double alpha, int fk, double fp, double t0, double t1, double gtj,
                         double eps, int max_niter, int *niter_out)
{
    double sum = 0.0, v, v2, v3, v4, i = 0.0;
    int niter = 0;
    double g0 = orc_g(fk, fp, t0), g1 = orc_g(fk, fp, t1);
    if (alpha == 1.0) { if (niter_out) *niter_out = 0; return (g1 == gtj) ? 1.0 : 0.0; }   /* :1939-1942 */
    v = orc_inv_g_der(fk, fp, gtj, 1, orc_inv_g_der(fk, fp, gtj, 0, 1));                  /* :1945 */
    v *= pow(gtj, 1 - alpha);
    v3 = pow(fabs(1.0 - g1 / gtj), 1 - alpha);
    v4 = pow(fabs(1.0 - g0 / gtj), 1 - alpha);
    v2 = -(v3 - v4) / (1 - alpha);
    sum += v2 * v;
    while (fabs(v * v2) > eps) {                                                          /* :1951 */
        i += 1.0;
        v = orc_inv_g_der(fk, fp, gtj, (int)(i + 1.0), v);
        v *= -gtj / i;
        v3 *= fabs(1.0 - g1 / gtj);
        v4 *= fabs(1.0 - g0 / gtj);
        v2 = -(v3 - v4) / (i - alpha + 1.0);
        sum += v2 * v;
        niter++;
        if (niter > max_niter) break;
    }
    if (niter_out) *niter_out = niter;
    return sum;
}

/* g(pa) - g(pa - d) for g = t^gp, exact for small d. */
static double gap_pow(double pa, double gp, double d, double G) { return -G * expm1(gp * log1p(-d / pa)); }

/* b(t0,t1;a) = int_{p0}^{p1} (pa^gp - s^gp)^-alpha ds, the integral the
 * reference's series approximate, by tanh-sinh quadrature (step 1/64, |t| <=
 * 4): for p1 < pa directly in s, with pa - s formed from the distance to p1;
 * for the newest interval (p1 = pa) in w = (pa - s)^(1-alpha), which takes
 * the (pa - s)^-alpha singularity into the measure and leaves a bounded
 * integrand. */
double orc_kb_general(double alpha, double gp, double p0, double p1, double pa)
{
    if (alpha == 1.0) return (p1 == pa) ? 1.0 : 0.0;
    if (gp == 1.0) return orc_kb_closed(alpha, p0, p1, pa);
    const double G = pow(pa, gp), h = 1.0 / 64.0;
    double sum = 0.0;
    if (p1 < pa) {
        const double L = p1 - p0, gr = pa - p1;
        for (int j = -256; j <= 256; j++) {
            double t = j * h, u = 0.5 * M_PI * sinh(t);
            double dr = L / (1.0 + exp(2.0 * u));                 /* p1 - s */
            double eu = exp(u), ch = eu + 1.0 / eu;
            double wt = h * 0.5 * L * (0.5 * M_PI * cosh(t)) * 4.0 / (ch * ch);
            sum += wt * pow(gap_pow(pa, gp, gr + dr, G), -alpha);
        }
        return sum;
    }
    const double e1 = 1.0 - alpha, w0 = pow(pa - p0, e1);
    for (int j = -256; j <= 256; j++) {
        double t = j * h, u = 0.5 * M_PI * sinh(t);
        double w = w0 / (1.0 + exp(-2.0 * u));
        double eu = exp(u), ch = eu + 1.0 / eu;
        double wt = h * 0.5 * w0 * (0.5 * M_PI * cosh(t)) * 4.0 / (ch * ch);
        /* (pa-p0)^e1^(1/e1) can round above pa-p0; past pa it is outside the domain */
        double d = fmin(pow(w, 1.0 / e1), pa - p0);
        /* below d = 1e-16 pa the ratio gap/d equals its limit g'(pa) in double;
         * formed from subnormal d it would lose its digits or divide 0 by d */
        double phi = (d > 1e-16 * pa) ? pow(gap_pow(pa, gp, d, G) / d, -alpha) : pow(gp * pow(pa, gp - 1.0), -alpha);
        sum += wt * phi;
    }
    return sum / e1;
}

double orc_g_exponent(int fk, double fp)
{
    return fk == 1 ? 0.5 : (fk == 2 ? 2.0 : (fk == 3 ? fp : 1.0));
}

/* kb over a grid with kernel fk: fk 0 = orc_kb; else the integral. */
double orc_kb_g(double alpha, int fk, double fp, int t0, int a, const double *Time,
                int weight_mode, double eps)
{
    if (fk == 0) return orc_kb(alpha, t0, a, Time, weight_mode, eps);
    return orc_kb_general(alpha, orc_g_exponent(fk, fp), Time[t0], Time[t0 + 1], Time[a]);
}

double orc_l1_weights_g(double alpha, int fk, double fp, int n, const double *Time,
                        int weight_mode, double eps, double *kb3)
{
    for (int t = 0; t < n; t++)
        kb3[t] = orc_kb_g(alpha, fk, fp, t, n + 1, Time, weight_mode, eps);
    return orc_kb_g(alpha, fk, fp, n, n + 1, Time, weight_mode, eps);
}

/* init_time_step (:3942-3949) + precalc_kb_3 (:4069-4085) for step n:
 * kkk0 = kb(alpha,n,n+1,n+1); kb3[t] = kb(alpha,t,t+1,n+1), t<n. */
double orc_l1_weights(double alpha, int n, const double *Time,
                      int weight_mode, double eps, double *kb3)
{
    for (int t = 0; t < n; t++)
        kb3[t] = orc_kb(alpha, t, n + 1, Time, weight_mode, eps);
    return orc_kb(alpha, n, n + 1, Time, weight_mode, eps);
}

/* ------------------------------------------------------------------------ */
/* time_summing (fpde_hpc_1d.cpp:4110-4136): node-outer, t-inner, newest
 * interval first; old is level-major [levels][nodes].  sum_alg 1 (restricted
 * summing, :4117-4131) breaks before the first weight with |kb| < sum_param. */
void orc_time_summing(int nodes, int levels, const double *old,
                      const double *kb3, const double *tau, double *S,
                      int nthreads, int sum_alg, double sum_param)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
    for (int id = 0; id < nodes; id++) {
        double time_sum = 0.0;
        for (int t = levels - 2; t >= 0; t--) {
            double kv = kb3[t];
            if (sum_alg == 1 && fabs(kv) < sum_param) break;
            double diff = old[(size_t)(t + 1) * nodes + id] - old[(size_t)t * nodes + id];
            time_sum += kv * diff / tau[t];
        }
        S[id] = time_sum;
    }
    (void)nthreads;
}

/* ------------------------------------------------------------------------ */
/* sum_alg 2.  Under sum_alg 2, precalc_kb_3 (fpde_hpc_1d.cpp:4076-4081) calls
 * kb(alpha, tstep-1, tstep, tstep+1, .., time_i1, 0), which takes the series
 * branch for every t1 != a (:2280-2285) and caches v3 of each evaluated term
 * in row2_precalc[time_i1][0] (:2114-2119; eps = global_eps for idx ==
 * time_i1, :2015).  alpha == 1 returns before the cache (:2011-2014). */
int orc_series_row(double alpha, double p0, double p1, double pa, double eps,
                   int cap, double *v3out)
{
    double v1 = 1.0, v2, v3, v6, term, i = 0.0;
    int niter = 0, k = 0;
    if (alpha == 1.0) return 0;
    v2 = pow(pa, -alpha) * p1;                                /* :2031 */
    v6 = p0 / p1;                                             /* :2032 */
    do {
        if (i != 0.0) {
            v1 *= -(-alpha - i + 1.0) / i;                    /* :2045 */
            v2 *= p1 / pa;                                    /* :2049 */
            v6 *= p0 / p1;                                    /* :2050 */
        }
        if (i == 0.0)
            v3 = (p1 - p0) / p1;                              /* :2070 */
        else {
            v3 = (1.0 / (i + 1.0)) * (1.0 - v6);              /* :2078 */
            if (fabs(v3) > fabs(v3) * eps) niter++;           /* one inner pass */
        }
        if (k < cap) v3out[k] = v3;
        k++;
        term = v1 * v2 * v3;
        i += 1.0;
        niter++;
        if (niter > 100000) break;                            /* integr_max_niter */
    } while (fabs(term) > eps);
    return k;
}

/* _kb_row2_fixed_coefs, fpde_hpc_1d.cpp:1159-1180 (right == 0). */
static double series_fixed_coefs(double gtj, double gtjma, const double *c, int n, double alpha)
{
    double sum = 0.0, v1 = gtjma, i = 0.0;
    int ii = 0;
    do {
        sum += v1 * c[ii];
        i += 1.0;
        ii++;
        v1 *= -(-alpha - i + 1.0) / (i * gtj);
    } while (ii != n);
    return sum;
}

void orc_series_step(int nodes, int P, double *H, const double *v3, int K,
                     double p1, double gtj, double alpha, double tau,
                     const double *u0, const double *u1, double *S)
{
    double gtjma = pow(gtj, -alpha);
    for (int id = 0; id < nodes; id++) {
        double *h = H + (size_t)id * P;
        double diff = u1[id] - u0[id];
        double mult = p1;                                     /* :4143 */
        for (int j = 0; j < P; j++)
            if (K > j) {                                      /* :4145 */
                h[j] += mult * v3[j] * diff;                  /* :4147 */
                mult *= p1;
            }
        S[id] = series_fixed_coefs(gtj, gtjma, h, P, alpha) / tau;   /* :4150 */
    }
}

/* ------------------------------------------------------------------------ */
/* van Genuchten-Mualem soil, fpde_hpc_1d.cpp:2513-2524 (cv_form 4),
 * :2558-2571 (1/C), :2634-2635 (k_form 1).  P = suction head [m]. */
#define VGM_N 1.25
#define VGM_S0 0.0736
#define VGM_S1 0.55
#define VGM_A 0.066

double orc_wetness(double P)
{
    double n = VGM_N, s0 = VGM_S0, s1 = VGM_S1, a = VGM_A;
    P *= 100.0;
    if (P <= 0.0) return s1;
    return s0 + ((s1 - s0) / pow(1 + pow(a * P, n), (1 - 1 / n)));
}

double orc_inv_dw_dh(double P)
{
    double n = VGM_N, s0 = VGM_S0, s1 = VGM_S1, a = VGM_A, Ch;
    if (P <= 0.0) Ch = 0.0;
    else {
        double hn = pow(100.0, n), h2n = pow(100.0, 2.0 * n);
        Ch = ((((1 - n) * hn * s0 + (n - 1) * hn * s1) * pow(a * P, n) * pow(1 + hn * pow(a * P, n), (1 / n))) /
              (h2n * P * pow(a * P, 2 * n) + 2 * hn * P * pow(a * P, n) + P));
    }
    if (Ch != 0.0) return 1.0 / Ch;
    return 1e20;
}

double orc_kr(double P)
{
    double n = VGM_N, s0 = VGM_S0, s1 = VGM_S1;
    double w = orc_wetness(P);
    return pow(((w - s0) / (s1 - s0)), 0.5) *
           pow(1.0 - pow(1.0 - pow(((w - s0) / (s1 - s0)), (1.0 / (1.0 - 1.0 / n))), (1.0 - 1.0 / n)), 2.0);
}

/* ------------------------------------------------------------------------ */
void orc_gauss_legendre(int n, double a, double b, double *x, double *w)
{
    for (int i = 0; i < n; i++) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; it++) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
        }
        /* ascending order */
        x[n - 1 - i] = 0.5 * (a + b) + 0.5 * (b - a) * z;
        w[n - 1 - i] = (b - a) / ((1.0 - z * z) * pp * pp);
    }
}

/* ------------------------------------------------------------------------ */
void orc_thomas(int n, const double *a, const double *b, const double *c,
                const double *d, double *x)
{
    double *cp = (double *)malloc(sizeof(double) * n), *dp = (double *)malloc(sizeof(double) * n);
    cp[0] = c[0] / b[0]; dp[0] = d[0] / b[0];
    for (int i = 1; i < n; i++) {
        double m = b[i] - a[i] * cp[i - 1];
        cp[i] = c[i] / m;
        dp[i] = (d[i] - a[i] * dp[i - 1]) / m;
    }
    x[n - 1] = dp[n - 1];
    for (int i = n - 2; i >= 0; i--) x[i] = dp[i] - cp[i] * x[i + 1];
    free(cp); free(dp);
}

/* ------------------------------------------------------------------------ */
/* Orders and quadrature weights of a run: q=0 -> the single order alpha with
 * weight 1; otherwise c_q*rho(a_q) with rho(a)=Gamma(3-a)/2 (Gauss-Legendre)
 * or the fr_do.cpp trapezoid (:67-71, :107-108) with w(a)=Gamma(3-a)/2.     */
static int run_orders(const orc_params *p, double *al, double *cw)
{
    if (p->do_q <= 0) { al[0] = p->alpha; cw[0] = 1.0; return 1; }
    int q = p->do_q;
    if (p->do_rule == 0) {
        orc_gauss_legendre(q, p->do_a0, p->do_a1, al, cw);
        for (int i = 0; i < q; i++) cw[i] *= tgamma(3.0 - al[i]) / 2.0;
    } else {
        /* trapezoid: sum_i 0.5*(WA_i CD_i + WA_{i+1} CD_{i+1}), WA_i = w(a_i)*h */
        double hstep = (p->do_a1 - p->do_a0) / (q - 1);
        for (int i = 0; i < q; i++) {
            al[i] = p->do_a0 + hstep * i;
            cw[i] = tgamma(3.0 - al[i]) / 2.0 * hstep * ((i == 0 || i == q - 1) ? 0.5 : 1.0);
        }
    }
    return p->do_q;
}

/* Da = 1/Gamma(1-alpha) (fpde_hpc_1d.cpp:3598-3601); tgamma agrees with the
 * reference Gamma (Cody's rational form on (1,2)) to ~1e-15. */
static double da_of(double alpha) { return (alpha != 1.0) ? 1.0 / tgamma(1.0 - alpha) : 1.0; }

typedef struct {
    int q;
    double al[256], cw[256];
    double *Time;   /* cumulative time grid, as init_time_step builds it */
    double *tau;
    double *kb3;    /* per-order scratch */
    double *Hs;     /* sum_alg 2: moments [q][nodes][P] (Htmp, :3342-3348) */
    double *v3;     /* sum_alg 2: coefficient row scratch */
} weights_ctx;

static void wctx_init(weights_ctx *w, const orc_params *p)
{
    w->q = run_orders(p, w->al, w->cw);
    w->Time = (double *)malloc(sizeof(double) * (p->steps + 2));
    w->tau = (double *)malloc(sizeof(double) * (p->steps + 1));
    w->kb3 = (double *)malloc(sizeof(double) * (p->steps + 1));
    w->Hs = NULL;
    w->v3 = NULL;
    w->Time[0] = 0.0;
    for (int k = 0; k <= p->steps; k++)
        w->tau[k] = p->taus ? p->taus[k < p->steps ? k : p->steps - 1] : p->tau;
    for (int k = 1; k <= p->steps + 1; k++) w->Time[k] = w->Time[k - 1] + w->tau[k - 1];  /* :3944 */
}
static void wctx_free(weights_ctx *w) { free(w->Time); free(w->tau); free(w->kb3); free(w->Hs); free(w->v3); }

/* ------------------------------------------------------------------------ */
/* Tridiagonal line solve of the reference step for one line of nodes.
 * Rows (A_,B,R) follow fpde_hpc_1d.cpp:3095-3142 with Dg=1 and kb(gamma=1)=1
 * (gamma=1: classical diffusion, calc_v returns 0, :2307-2313).
 *   A_i = K_{i-1} invC_i/(h h), B_i = K_i invC_i/(h h),
 *   R_i = invC_i (K_i/(h h) + K_{i-1}/(h h)) + L,    L = Da kkk0/tau,
 *   Om_i = -L U^n_i + H_i  (H = Da * history sum, 0 off the history sweep).
 * al1/bt1 (:4002-4027, :4585-4590) with Al[1], Bt[1] from the boundary kind;
 * bottom U_N from :4684-4703 (bcN()=0 for the zero-flux bottom); back
 * substitution :4723-4724.
 * top: 0 = Dirichlet value utop (al1 analytic-test branch, Al[1]=0, Bt[1]=utop),
 *      1 = zero flux (3D lateral sweeps, Al[1]=1, Bt[1]=0, :4021/:4511). */
typedef struct { int coef_mode; double ksat, zoff, zstep; } coef_ctx;

static inline double c_invC(const coef_ctx *c, double u, double z)
{
    if (c->coef_mode == 1) return 1.0;
    return orc_inv_dw_dh(z - u);            /* P = -(U - Z), :2530 */
}
static inline double c_K(const coef_ctx *c, double u, double z)
{
    if (c->coef_mode == 1) return 1.0;
    return c->ksat * orc_kr(z - u);         /* KK = Kr*k, :2656 */
}

/* One line: n nodes; un = U^n (rhs), uit = coefficient iterate, unew = out.
 * zc[i] = depth of node i (for the suction head), h = spacing. */
static void line_solve(int n, const double *un, const double *uit, const double *H,
                       const double *zc, double h, double L, int top, double utop,
                       const coef_ctx *cc, double *Al, double *Bt, double *K,
                       double *iC, double *unew, const double *kap, const double *V)
{
    int N = n - 1;   /* reference N: nodes 0..N */
    for (int i = 0; i < n; i++) {
        K[i] = c_K(cc, uit[i], zc[i]);
        iC[i] = c_invC(cc, uit[i], zc[i]);
    }
    double hh = h * h;
    if (top == 0) { Al[1] = 0.0; Bt[1] = utop; } else { Al[1] = 1.0; Bt[1] = 0.0; }
    for (int i = 1; i < N; i++) {
        /* space-fractional factor Dg kb(gamma,i-1,i,i) of A_/B/R (:3095-3120)
         * and the explicit sum -Dg inv_dw_dh calc_v of Om1 (:4182-4189) */
        double kp = kap ? kap[i] : 1.0;
        double A_ = K[i - 1] * iC[i] / hh * kp;
        double B_ = K[i] * iC[i] / hh * kp;
        double R_ = iC[i] * ((K[i] / hh) + (K[i - 1] / hh)) * kp + L;
        double Om = -L * un[i] + H[i];
        if (V) Om -= iC[i] * V[i];
        Al[i + 1] = B_ / (R_ - A_ * Al[i]);
        Bt[i + 1] = (Al[i + 1] / B_) * (A_ * Bt[i] - Om);
    }
    /* zero-flux bottom, :4688-4703 with bcN()=0 */
    if (fabs(Bt[N]) > 1e-15) {
        if (fabs(1.0 - Al[N]) > 1e-15) unew[N] = Bt[N] / (1.0 - Al[N]);
        else unew[N] = uit[N - 1];
    } else {
        if (fabs(1.0 - Al[N]) > 1e-15) unew[N] = 0.0;
        else unew[N] = uit[N - 1];
    }
    for (int i = N - 1; i >= 0; i--) unew[i] = Al[i + 1] * unew[i + 1] + Bt[i + 1];
}

/* History term Da*time_summing for every node: per order, the reference
 * weights of step n and the sum_alg-0 loop; orders combined with cw. */
static void history_term(weights_ctx *w, const orc_params *p, int n, int nodes,
                         const double *old, double *H, double *tmp, double *Lout)
{
    double L = 0.0;
    int P = (int)p->sum_param;
    for (int i = 0; i < nodes; i++) H[i] = 0.0;
    if (p->sum_alg == 2 && w->Hs == NULL) {
        w->Hs = (double *)calloc((size_t)w->q * nodes * P, sizeof(double));
        w->v3 = (double *)malloc(sizeof(double) * P);
    }
    for (int q = 0; q < w->q; q++) {
        double a = w->al[q];
        double kkk0 = orc_l1_weights_g(a, p->fk, p->fp, n, w->Time, p->weight_mode, p->series_eps, w->kb3);
        double Da = da_of(a);
        L += w->cw[q] * Da * kkk0 / w->tau[n];
        if (n >= 1 && p->sum_alg == 2) {
            /* moments of the newest interval [n-1, n], then the series at n+1 */
            const double *T = w->Time;
            int K = orc_series_row(a, T[n - 1], T[n], T[n + 1], p->series_eps, P, w->v3);
            orc_series_step(nodes, P, w->Hs + (size_t)q * nodes * P, w->v3, K, T[n], T[n + 1], a,
                            w->tau[n], old + (size_t)(n - 1) * nodes, old + (size_t)n * nodes, tmp);
            for (int i = 0; i < nodes; i++) H[i] += w->cw[q] * Da * tmp[i];
        } else if (n >= 1) {
            orc_time_summing(nodes, n + 1, old, w->kb3, w->tau, tmp, p->nthreads, p->sum_alg, p->sum_param);
            for (int i = 0; i < nodes; i++) H[i] += w->cw[q] * Da * tmp[i];
        }
    }
    *Lout = L;
}

/* ------------------------------------------------------------------------ */
/* Space-fractional explicit scheme (implicit=0, space_der=0).  The weights
 * Dg*kb(gamma,j-1,j,i;Z) of calc_v (:2315-2322) over the uniform grid
 * Z_k = k h (:2386-2387), row-packed: row i (1..n-2) holds j = 1..i-1. */
static int space_on(const orc_params *p) { return p->gamma > 0.0 && p->gamma < 1.0; }

static double *space_table(int n, double h, double gamma, int weight_mode, double eps)
{
    double *Z = (double *)malloc(sizeof(double) * n);
    size_t rows = n > 2 ? (size_t)(n - 2) : 0;
    double *T = (double *)malloc(sizeof(double) * (rows * (rows + 1) / 2 + 1));
    if (!Z || !T) { free(Z); free(T); return NULL; }
    for (int k = 0; k < n; k++) Z[k] = k * h;
    double Dg = 1.0 / tgamma(1.0 - gamma);                        /* :3602-3603 */
    for (int i = 1; i <= n - 2; i++) {
        double *row = T + (size_t)(i - 1) * i / 2;               /* row i starts at (i-1)i/2 */
        for (int j = 1; j <= i - 1; j++)
            row[j - 1] = Dg * orc_kb(gamma, j - 1, i, Z, weight_mode, eps);
    }
    free(Z);
    return T;
}

static void space_apply(int n, const double *T, const double *F, double *V)
{
    for (int i = 0; i < n; i++) V[i] = 0.0;
    for (int i = 1; i <= n - 2; i++) {
        const double *row = T + (size_t)(i - 1) * i / 2;
        double v = 0.0;
        for (int j = 1; j <= i - 1; j++) v += row[j - 1] * F[j];  /* :2319-2323 */
        V[i] = v;
    }
}

void orc_space_v(int nodes, double h, double gamma, int weight_mode, double eps,
                 const double *F, double *V)
{
    double *T = space_table(nodes, h, gamma, weight_mode, eps);
    if (!T) { for (int i = 0; i < nodes; i++) V[i] = 0.0; return; }
    space_apply(nodes, T, F, V);
    free(T);
}

void orc_space_flux(int nodes, double h, const double *K, const double *u, double *F)
{
    double hh = h * h;
    F[0] = 0.0;
    F[nodes - 1] = 0.0;
    for (int j = 1; j <= nodes - 2; j++)                          /* :4176-4179 */
        F[j] = K[j] * u[j + 1] / hh - (K[j] / hh + K[j - 1] / hh) * u[j] + K[j - 1] * u[j - 1] / hh;
}

/* the generalized kernel is restated with its exact weights (weight_mode 0)
 * and the full history sum only */
static int fk_supported(const orc_params *p)
{
    return p->fk == 0 || (p->fk >= 1 && p->fk <= 3 && p->weight_mode == 0 && p->sum_alg == 0);
}

int orc_solve_1d(const orc_params *p, double *u)
{
    if (!fk_supported(p)) return -2;
    int n = p->nodes, M = p->steps;
    double h = p->length / (n - 1);
    weights_ctx w; wctx_init(&w, p);
    coef_ctx cc = {p->coef_mode, p->ksat, 0.0, h};
    double *old = (double *)malloc(sizeof(double) * (size_t)n * (M + 1));
    double *H = (double *)malloc(sizeof(double) * n), *tmp = (double *)malloc(sizeof(double) * n);
    double *Al = (double *)malloc(sizeof(double) * (n + 1)), *Bt = (double *)malloc(sizeof(double) * (n + 1));
    double *K = (double *)malloc(sizeof(double) * n), *iC = (double *)malloc(sizeof(double) * n);
    double *zc = (double *)malloc(sizeof(double) * n);
    double *uit = (double *)malloc(sizeof(double) * n), *unew = (double *)malloc(sizeof(double) * n);
    if (!old || !H || !tmp || !Al || !Bt || !K || !iC || !zc || !uit || !unew) return -1;
    for (int i = 0; i < n; i++) zc[i] = i * h;
    double utop = u[0];
    /* space-fractional scheme: kap[i] = Dg kb(gamma,i-1,i,i) (the newest
     * interval, closed form: t1 == a), V = calc_v of F(U^n) with K at U^n */
    double *kap = NULL, *V = NULL, *Tsp = NULL, *Fsp = NULL, *Ksp = NULL;
    if (space_on(p)) {
        kap = (double *)malloc(sizeof(double) * n);
        V = (double *)malloc(sizeof(double) * n);
        Fsp = (double *)malloc(sizeof(double) * n);
        Ksp = (double *)malloc(sizeof(double) * n);
        Tsp = space_table(n, h, p->gamma, p->weight_mode, p->series_eps);
        if (!kap || !V || !Fsp || !Ksp || !Tsp) return -1;
        double Dg = 1.0 / tgamma(1.0 - p->gamma);
        for (int i = 0; i < n; i++) kap[i] = (i >= 1) ? Dg * orc_kb(p->gamma, i - 1, i, zc, p->weight_mode, p->series_eps) : 1.0;
    }
    memcpy(old, u, sizeof(double) * n);                        /* :5104-5113 */
    for (int s = 0; s < M; s++) {
        double L;
        history_term(&w, p, s, n, old, H, tmp, &L);
        const double *un = old + (size_t)s * n;
        if (V) {
            for (int i = 0; i < n; i++) Ksp[i] = c_K(&cc, un[i], zc[i]);
            orc_space_flux(n, h, Ksp, un, Fsp);
            space_apply(n, Tsp, Fsp, V);
        }
        memcpy(uit, un, sizeof(double) * n);
        /* Picard (picard_max > 1) starts from the linear extrapolation
         * 2 U^n - U^{n-1}; a single pass keeps the lagged coefficients at U^n
         * (the reference scheme, KK() inside A_/B/R, :3095-3142). */
        if (p->picard_max > 1 && s >= 1)
            for (int i = 0; i < n; i++) uit[i] = 2.0 * un[i] - un[i - n];
        for (int k = 0; k < p->picard_max; k++) {
            line_solve(n, un, uit, H, zc, h, L, 0, utop, &cc, Al, Bt, K, iC, unew, kap, V);
            double dmax = 0.0;
            for (int i = 0; i < n; i++) { double d = fabs(unew[i] - uit[i]); if (d > dmax) dmax = d; }
            memcpy(uit, unew, sizeof(double) * n);
            if (dmax <= p->picard_tol) break;
        }
        memcpy(old + (size_t)(s + 1) * n, uit, sizeof(double) * n);   /* :4793-4805 */
    }
    memcpy(u, old + (size_t)M * n, sizeof(double) * n);
    free(old); free(H); free(tmp); free(Al); free(Bt); free(K); free(iC); free(zc); free(uit); free(unew);
    free(kap); free(V); free(Tsp); free(Fsp); free(Ksp);
    wctx_free(&w);
    return 0;
}

/* 3D LOD (fpde_hpc_1d.cpp:5144-5626): layout idx = z + n*(x + n*y) with the
 * depth axis z fastest (the reference's N axis, idx() at :1181).  Sweep order
 * y (m3d_c=2, with the history term), x (m3d_c=1), z (m3d_c=0, Dirichlet top,
 * zero-flux bottom); the lateral sweeps have zero flux at both ends.  Every
 * sweep has the L*(U - U_prev) term (R at :3120; implicit3d=0) and evaluates
 * the coefficients at the field left by the previous sweep. */
int orc_solve_3d(const orc_params *p, double *u) { return orc_solve_3d_trace(p, u, NULL); }

int orc_solve_3d_trace(const orc_params *p, double *u, double *sweeps)
{
    if (!fk_supported(p)) return -2;
    int n = p->nodes, M = p->steps;
    size_t nn = (size_t)n * n * n;
    double h = p->length / (n - 1);
    weights_ctx w; wctx_init(&w, p);
    coef_ctx cc = {p->coef_mode, p->ksat, 0.0, h};
    double *old = (double *)malloc(sizeof(double) * nn * (M + 1));
    double *H = (double *)malloc(sizeof(double) * nn), *tmp = (double *)malloc(sizeof(double) * nn);
    double *cur = (double *)malloc(sizeof(double) * nn);
    double *utop = (double *)malloc(sizeof(double) * n * n);
    if (!old || !H || !tmp || !cur || !utop) return -1;
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++) utop[y * n + x] = u[(size_t)n * (x + (size_t)n * y)];
    memcpy(old, u, sizeof(double) * nn);
    /* space-fractional scheme: every sweep applies the 1D operator along its
     * own axis (the y/x/z sweeps switch Z, dL, N to that axis, :5262-5265) */
    double *Tsp = NULL, kap[1024];
    if (space_on(p)) {
        double zl[1024];
        for (int i = 0; i < n; i++) zl[i] = i * h;
        Tsp = space_table(n, h, p->gamma, p->weight_mode, p->series_eps);
        if (!Tsp) return -1;
        double Dg = 1.0 / tgamma(1.0 - p->gamma);
        for (int i = 0; i < n; i++) kap[i] = (i >= 1) ? Dg * orc_kb(p->gamma, i - 1, i, zl, p->weight_mode, p->series_eps) : 1.0;
    }
    for (int s = 0; s < M; s++) {
        double L;
        history_term(&w, p, s, (int)nn, old, H, tmp, &L);
        memcpy(cur, old + (size_t)s * nn, sizeof(double) * nn);
        for (int axis = 2; axis >= 0; axis--) {       /* y, x, z */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(p->nthreads > 0 ? p->nthreads : omp_get_max_threads())
#endif
            for (int line = 0; line < n * n; line++) {
                double lun[1024], luit[1024], lH[1024], lz[1024], lAl[1025], lBt[1025], lK[1024], liC[1024], lnew[1024];
                int a0 = line % n, a1 = line / n;
                size_t base, stride;
                if (axis == 0) { base = (size_t)n * (a0 + (size_t)n * a1); stride = 1; }          /* z line at (x=a0,y=a1) */
                else if (axis == 1) { base = a0 + (size_t)n * n * a1; stride = n; }           /* x line at (z=a0,y=a1) */
                else { base = a0 + (size_t)n * a1; stride = (size_t)n * n; }                   /* y line at (z=a0,x=a1) */
                for (int i = 0; i < n; i++) {
                    size_t id = base + stride * i;
                    lun[i] = cur[id];
                    /* Om += mult Da time_summing: mult = 1 in the m3d_c=2 sweep
                     * only, or 1/3 in every sweep when split3d is odd (the
                     * cached per-node sum, :4106-4108, :4172-4173, :4258-4263) */
                    lH[i] = (p->split3d & 1) ? H[id] / 3.0 : ((axis == 2) ? H[id] : 0.0);
                    lz[i] = (axis == 0) ? i * h : a0 * h;
                }
                memcpy(luit, lun, sizeof(double) * n);
                double top_val = (axis == 0) ? utop[a1 * n + a0] : 0.0;
                double lF[1024], lV[1024];
                if (Tsp) {
                    for (int i = 0; i < n; i++) lK[i] = c_K(&cc, lun[i], lz[i]);
                    orc_space_flux(n, h, lK, lun, lF);
                    space_apply(n, Tsp, lF, lV);
                }
                line_solve(n, lun, luit, lH, lz, h, L, axis == 0 ? 0 : 1, top_val, &cc, lAl, lBt, lK, liC, lnew,
                           Tsp ? kap : NULL, Tsp ? lV : NULL);
                for (int i = 0; i < n; i++) cur[base + stride * i] = lnew[i];
            }
            if (sweeps && s == M - 1) memcpy(sweeps + (size_t)(2 - axis) * nn, cur, sizeof(double) * nn);
        }
        memcpy(old + (size_t)(s + 1) * nn, cur, sizeof(double) * nn);
    }
    memcpy(u, old + (size_t)M * nn, sizeof(double) * nn);
    free(old); free(H); free(tmp); free(cur); free(utop); free(Tsp);
    wctx_free(&w);
    return 0;
}

/* Effective weights for a uniform step, in the form the product folds them:
 * step n's history term is sum_{j=1..n} omega[j-1]*D[n-j] + corr[2n]*D[0] +
 * corr[2n+1]*D[1] with D[t] = U^{t+1}-U^t.  Built here from orc_kb over an
 * exact grid Time[k] = k*tau. */
void orc_folded_weights(const orc_params *p, double *local, double *omega, double *corr)
{
    double al[256], cw[256];
    int q = run_orders(p, al, cw), M = p->steps;
    double *Time = (double *)malloc(sizeof(double) * (M + 2));
    for (int k = 0; k <= M + 1; k++) Time[k] = k * p->tau;
    for (int j = 0; j < M; j++) omega[j] = 0.0;
    for (int k = 0; k < 2 * M; k++) corr[k] = 0.0;
    *local = 0.0;
    for (int iq = 0; iq < q; iq++) {
        double Da = da_of(al[iq]) * cw[iq] / p->tau;
        *local += Da * orc_kb_closed(al[iq], 0.0, p->tau, p->tau);
        for (int j = 1; j <= M; j++)           /* interval [n-j, n-j+1] at a=n+1 */
            omega[j - 1] += Da * orc_kb_closed(al[iq], Time[M - j], Time[M - j + 1], Time[M + 1]);
        if (p->weight_mode == 1)
            for (int n = 1; n < M; n++)
                for (int t = 0; t <= 1 && t < n; t++) {
                    double s = orc_kb_series(al[iq], Time[t], Time[t + 1], Time[n + 1], p->series_eps, 100000, NULL);
                    double c = orc_kb_closed(al[iq], Time[t], Time[t + 1], Time[n + 1]);
                    corr[2 * n + t] += Da * (s - c);
                }
    }
    free(Time);
}

/* ------------------------------------------------------------------------ */
/* Workload initial conditions.  1D: u0 = -1 + sum_{k=1..4} a_k sin(k pi z + phi_k),
 * drawn from mt19937_64(seed) in the order a_1, phi_1, a_2, phi_2, ...      */
void orc_init_1d(int nodes, double length, uint64_t seed, double *u)
{
    orc_mt64 s; orc_mt64_seed(&s, seed);
    double a[4], ph[4];
    for (int k = 0; k < 4; k++) { a[k] = orc_mt64_uniform(&s, 0.0, 0.05); ph[k] = orc_mt64_uniform(&s, 0.0, 2.0 * M_PI); }
    double h = length / (nodes - 1);
    for (int i = 0; i < nodes; i++) {
        double z = i * h, v = -1.0;
        for (int k = 0; k < 4; k++) v += a[k] * sin((k + 1) * M_PI * z + ph[k]);
        u[i] = v;
    }
}

/* 3D: u0 = -1 + sum_{b=1..8} A_b exp(-|x-c_b|^2/(2 s_b^2)); per blob the draws
 * are c_x, c_y, c_z ~ U(0.2,0.8)*length, s ~ U(0.05,0.15)*length, A ~ U(0,0.05). */
void orc_init_3d(int n, double length, uint64_t seed, double *u)
{
    orc_mt64 s; orc_mt64_seed(&s, seed);
    double c[8][3], sg[8], A[8];
    for (int b = 0; b < 8; b++) {
        for (int d = 0; d < 3; d++) c[b][d] = orc_mt64_uniform(&s, 0.2, 0.8) * length;
        sg[b] = orc_mt64_uniform(&s, 0.05, 0.15) * length;
        A[b] = orc_mt64_uniform(&s, 0.0, 0.05);
    }
    double h = length / (n - 1);
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++)
            for (int z = 0; z < n; z++) {
                double pz = z * h, px = x * h, py = y * h, v = -1.0;
                for (int b = 0; b < 8; b++) {
                    double dx = px - c[b][0], dy = py - c[b][1], dz = pz - c[b][2];
                    v += A[b] * exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sg[b] * sg[b]));
                }
                u[z + (size_t)n * (x + (size_t)n * y)] = v;
            }
}

double orc_ksat_for_fourier(double u_top, double h, double tau, double fo)
{
    double P = 0.0 - u_top;
    return fo * h * h / (tau * orc_kr(P) * orc_inv_dw_dh(P));
}

/* ------------------------------------------------------------------------ */
/* fr_do.cpp: the distributed-order diffusion equation (see fpde_oracle.h). */

/* bk of the fr_do kernels (fr_do.cpp:129-134): d1^(1-b) - d2^(1-b), with
 * either distance 0 giving the other's power alone. */
static double frdo_bk(double d1, double d2, double beta)
{
    if (d1 == 0.0) return pow(d2, 1.0 - beta);
    if (d2 == 0.0) return pow(d1, 1.0 - beta);
    return pow(d1, 1.0 - beta) - pow(d2, 1.0 - beta);
}

/* weight functions: w (:40-43) of the derivative test, eq_w (:467-474) */
static double frdo_w(int example, double power, double a)
{
    if (example < 0) return tgamma(3.0 - a);
    if (example == 0) return tgamma(5.0 - a);
    if (example == 1) return pow(a, power) * (1.0 + power);
    return 0.0;
}

/* eq_f (:458-465); gpu_pi: eq_ff of the GPU program text (:256-263), where
 * 2 pi and 800 pi pi are formed in float from a float constant for pi */
static double frdo_f(int example, int gpu_pi, double t, double x)
{
    if (example != 0) return 0.0;
    if (gpu_pi) {
        const float pf = (float)M_PI;
        const double two_pi = (double)(2.0f * pf), k800 = (double)(800.0f * pf * pf);
        return sin(two_pi * x) * (k800 * t * t * t * t + (2400 * (t * t * t * t - t * t * t) / log(t)));
    }
    return sin(2 * M_PI * x) * (800 * M_PI * M_PI * t * t * t * t + (2400 * (t * t * t * t - t * t * t) / log(t)));
}

/* fill_arrays (:51-72) with F given: bkk[i na + j] = bk((i+1) tau, i tau, a_j)
 * for ntau intervals and WA_j = w(a_j)/(na-1); gs_j = Gamma(2 - a_j). */
typedef struct { int ntau, na; double *bkk, *WA, *gs, *CD; } frdo_tables;

static void frdo_fill(frdo_tables *T, int ntau, double tau, int na, int example, double power)
{
    T->ntau = ntau;
    T->na = na;
    T->bkk = (double *)malloc(sizeof(double) * (size_t)ntau * na);
    T->WA = (double *)malloc(sizeof(double) * na);
    T->gs = (double *)malloc(sizeof(double) * na);
    T->CD = (double *)malloc(sizeof(double) * na);
    for (int i = 0; i < ntau; i++)
        for (int j = 0; j < na; j++)
            T->bkk[(size_t)i * na + j] = frdo_bk((i + 1) * tau, i * tau, (na == 1) ? 0.0 : j / (double)(na - 1));
    for (int j = 0; j < na; j++) {
        T->WA[j] = frdo_w(example, power, j / (double)(na - 1)) * (1.0 / (na - 1));
        T->gs[j] = tgamma(2.0 - (j / (double)(na - 1)));
    }
}
static void frdo_free(frdo_tables *T) { free(T->bkk); free(T->WA); free(T->gs); free(T->CD); }

/* DO_Cder (:89-113) with F given: CD_i = sum_j f_j bkk[(t-j-1) na + i] /
 * Gamma(2 - a_i), f_j = (F[j+1]-F[j])/tau, then the trapezoid in a. */
static void frdo_cder(const frdo_tables *T, int ntau, double tau, const double *F, int maxt, int t0, int j0,
                      double *DF)
{
    const int na = T->na;
    for (int t = t0; t < ntau; t++) {
        double sum = 0.0;
        int jhi = (maxt < 0) ? t : ((t > maxt) ? maxt : t);
        for (int i = 0; i < na; i++) {
            double sum2 = 0.0;
            for (int j = j0; j < jhi; j++)
                if (j >= 0) sum2 += ((F[j + 1] - F[j]) / tau) * T->bkk[(size_t)(t - j - 1) * na + i];
            sum2 /= T->gs[i];
            T->CD[i] = sum2;
        }
        for (int i = 0; i < na - 1; i++) sum += 0.5 * (T->WA[i] * T->CD[i] + T->WA[i + 1] * T->CD[i + 1]);
        DF[t] = sum;
    }
}

void orc_do_cder(int ntau, double tau, int na, int example, double power, const double *F, int maxt,
                 int t0, int j0, double *DF)
{
    frdo_tables T;
    frdo_fill(&T, ntau, tau, na, example, power);
    frdo_cder(&T, ntau, tau, F, maxt, t0, j0, DF);
    frdo_free(&T);
}

double orc_doeq_solve(const orc_doeq_params *p, double *C)
{
    const int Nx = p->nx, Nt = p->nt, Na = p->na, tb = p->tblock;
    const double h = p->max_x / (Nx - 1), tau = p->max_t / Nt;
    const double Aa = 0.5 * p->D / (h * h), Bb = Aa;
    double *A = (double *)calloc(Nx + 1, sizeof(double)), *B = (double *)calloc(Nx + 1, sizeof(double));
    double *DFb = (double *)calloc((size_t)Nx * tb, sizeof(double));   /* eq_DF [node][block step] */
    double *DF = (double *)calloc((size_t)Nt + 1, sizeof(double));
    frdo_tables T;
    frdo_fill(&T, Nt, tau, Na, p->example, p->power);                        /* eq_init :685 */
    for (int i = 0; i < Nx * Nt; i++) C[i] = 0.0;                         /* eq_init :678-681 */
    for (int i = 0; i < Nx; i++) C[(size_t)i * Nt] = (p->example == 1) ? (i * h) * (1.0 - i * h) : 0.0;
    /* eq_b0: the trapezoid of WA bkk[0] / Gamma(2-a), bkk[0] = tau^(1-a) (:687-690) */
    double b0 = 0.0;
    for (int a = 0; a < Na - 1; a++) {
        double a0 = a / (double)(Na - 1), a1 = (a + 1) / (double)(Na - 1);
        double w0 = frdo_w(p->example, p->power, a0) / (Na - 1), w1 = frdo_w(p->example, p->power, a1) / (Na - 1);
        b0 += 0.5 * (w0 * frdo_bk(tau, 0.0, a0) / tgamma(2.0 - a0) + w1 * frdo_bk(tau, 0.0, a1) / tgamma(2.0 - a1));
    }
    const double Ss = (p->D / (h * h)) + (b0 / tau);
    A[1] = 0.0;                                                              /* u(t,0) = 0 (:693-695) */
    for (int i = 1; i < Nx - 1; i++) A[i + 1] = Bb / (Ss - Aa * A[i]);
    for (int i = 0; i < Nt; i += tb) {
        /* the history before the block, every node (:742-747) */
        for (int j = 0; j < Nx; j++) {
            frdo_cder(&T, i + tb, tau, C + (size_t)j * Nt, (i == 0) ? i : i - 1, i, 0, DF);
            for (int k = 0; k < tb; k++) DFb[(size_t)j * tb + k] = DF[i + k];
        }
        for (int j = (i == 0) ? 1 : 0; j < tb; j++) {
            int lev = i + j;
            /* eq_one_step (:724-732) with eq_Ff (:656-665) of level lev-1 */
            B[1] = 0.0;
            for (int x = 1; x < Nx - 1; x++) {
                const double *c = C + (size_t)x * Nt;
                double rp = -0.5 * (p->D / (h * h)) * (C[(size_t)(x - 1) * Nt + lev - 1] - 2.0 * c[lev - 1] + C[(size_t)(x + 1) * Nt + lev - 1]);
                double lp = -(b0 * c[lev - 1] / tau) - 0.5 * (frdo_f(p->example, p->gpu_pi, (lev - 1) * tau, x * h) +
                                                              frdo_f(p->example, p->gpu_pi, lev * tau, x * h));
                double nlp = DFb[(size_t)x * tb + (lev % tb)];
                if (!isfinite(lp)) lp = -(b0 * c[lev - 1] / tau) - frdo_f(p->example, p->gpu_pi, lev * tau, x * h);
                double F = rp + lp + nlp;
                B[x + 1] = (A[x + 1] / Bb) * (Aa * B[x] - F);
            }
            C[(size_t)(Nx - 1) * Nt + lev] = 0.0;                          /* u(max_x,t) = 0 */
            for (int x = Nx - 2; x >= 0; x--) C[(size_t)x * Nt + lev] = A[x + 1] * C[(size_t)(x + 1) * Nt + lev] + B[x + 1];
            /* the in-block interval [lev-1, lev] joins the next level's history (:751-756) */
            if (j != tb - 1)
                for (int x = 0; x < Nx; x++) {
                    frdo_cder(&T, lev + 2, tau, C + (size_t)x * Nt, lev, lev + 1, i - 1, DF);
                    DFb[(size_t)x * tb + j + 1] += DF[lev + 1];
                }
        }
    }
    double err = 0.0;
    if (p->example == 0) {
        int i = Nt - 1;
        for (int x = 0; x < Nx; x++) {
            double a = 100 * pow(i * tau, 4) * sin(2 * M_PI * x * h);          /* eq_anal_u :635-642 */
            double d = C[(size_t)x * Nt + i] - a;
            err += d * d;
        }
        err = sqrt(err / Nx);
    }
    free(A); free(B); free(DFb); free(DF);
    frdo_free(&T);
    return err;
}
