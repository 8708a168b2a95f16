// FP64 L1 / distributed-order weight generation on the device.
//
// For a uniform step tau the weight of the history difference D[t] =
// U^{t+1}-U^t at step n (unknown U^{n+1}) is, per order a,
//   Da(a)/tau * b(t,t+1;n+1),   b = int_{t tau}^{(t+1) tau} ((n+1)tau - s)^(-a) ds
// (precalc_kb_3 + time_summing, fpde_hpc_1d.cpp:4069-4136, Da = 1/Gamma(1-a)
// at :3598-3601).  With g(t)=t the closed form depends on j = n-t only:
//   b = tau^(1-a)/(1-a) * ((j+1)^(1-a) - j^(1-a)),
// evaluated here as j^(1-a)*expm1((1-a)*log1p(1/j)) to avoid the cancellation
// of the plain difference at large j.  The reference replaces the two oldest
// intervals (t0 <= a3k = 1) by the Newton-binomial series of _kb_row2
// (fpde_hpc_1d.cpp:2000-2131, stop at the first |term| <= global_eps); those
// enter as per-step corrections corr[2n+t] = series - Toeplitz weight.
// Orders are folded: omega_j = sum_q c_q Da(a_q)/tau b_j(a_q) (distributed
// order, fr_do.cpp:89-113 with the alpha-integral as a quadrature).
#include "fpde_internal.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace fpde {

struct OrderSet {
    int q;
    double a[64];
    double c[64];    // c_q * Da(a_q) / tau
};

__device__ inline double l1_closed_lag(double a, double tau, int j) {
    // b for lag j >= 1: tau^(1-a)/(1-a) * ((j+1)^(1-a) - j^(1-a));
    // a == 1: the reference returns 0 for every interval but the newest (:2169-2175)
    if (a == 1.0) return 0.0;
    double e = 1.0 - a;
    double jd = (double)j;
    return pow(tau, e) / e * pow(jd, e) * expm1(e * log1p(1.0 / jd));
}

// _kb_row2 mode 0 with g(t)=t (see oracle/fpde_oracle.c orc_kb_series).
__device__ inline double l1_series(double a, double p0, double p1, double pa, double eps) {
    if (a == 1.0) return (p1 == pa) ? 1.0 : 0.0;
    double sum = 0.0, v1 = 1.0, v2 = pow(pa, -a) * p1, v6 = p0 / p1, v3, term;
    double i = 0.0;
    int niter = 0;
    do {
        if (i != 0.0) {
            v1 *= -(-a - i + 1.0) / i;
            v2 *= p1 / pa;
            v6 *= p0 / p1;
            v3 = (1.0 / (i + 1.0)) * (1.0 - v6);
            if (fabs(v3) > fabs(v3) * eps) niter++;
        } else {
            v3 = (p1 - p0) / p1;
        }
        term = v1 * v2 * v3;
        sum += term;
        i += 1.0;
        niter++;
        if (niter > 100000) break;
    } while (fabs(term) > eps);
    return sum;
}

// Restricted summing (time_summing sum_alg 1, fpde_hpc_1d.cpp:4117-4131)
// walks the intervals newest first and stops at the first weight with
// |b| < cut.  The closed-form b decreases with the lag, so a Toeplitz lag j
// is kept exactly when b(j) >= cut; cut = 0 keeps every lag.
__device__ inline bool kept(double b, double cut) { return !(fabs(b) < cut); }

__global__ void k_weights_omega(OrderSet os, double tau, int M, double cut, double* __restrict__ omega) {
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (j > M) return;
    double s = 0.0;
    for (int q = 0; q < os.q; q++) {
        double b = l1_closed_lag(os.a[q], tau, j);
        if (kept(b, cut)) s += os.c[q] * b;
    }
    omega[j - 1] = s;
}

// corr[2n+t] for n in [1,M), t in {0,1}, t < n (after omega is complete).
// Interval t in {0,1} uses the series weight (weight_mode 1) or the closed one
// (weight_mode 0).  Under restricted summing it is reached only when every
// newer interval was kept: t=1 needs lag n-2 kept (or n < 3), t=0 needs t=1.
__global__ void k_weights_corr(OrderSet os, double tau, int M, double eps, int series, double cut,
                               const double* __restrict__ omega, double* __restrict__ corr) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * M) return;
    int n = idx >> 1, t = idx & 1;
    if (t >= n) { corr[idx] = 0.0; return; }
    double s = 0.0;
    for (int q = 0; q < os.q; q++) {
        double a = os.a[q];
        bool reach = (n < 3) || kept(l1_closed_lag(a, tau, n - 2), cut);
        // interval 1 exists for n >= 2 only (at n = 1 it is the newest one,
        // whose series would not converge: t1 = a)
        double b1 = 0.0;
        if (n >= 2) b1 = series ? l1_series(a, tau, 2.0 * tau, (n + 1) * tau, eps) : l1_closed_lag(a, tau, n - 1);
        if (t == 0) {
            reach = reach && (n < 2 || kept(b1, cut));
            double b0 = series ? l1_series(a, 0.0, tau, (n + 1) * tau, eps) : l1_closed_lag(a, tau, n);
            if (reach && kept(b0, cut)) s += os.c[q] * b0;
        } else if (reach && kept(b1, cut)) {
            s += os.c[q] * b1;
        }
    }
    corr[idx] = s - omega[n - t - 1];
}

// Series-compressed history (time_summing sum_alg 2, fpde_hpc_1d.cpp:4137-4151).
// With g(t)=t the kernel (a-s)^-a = sum_i (a)_i/i! a^{-a-i} s^i turns the
// history into moments H_j = sum_t int_{t_t}^{t_{t+1}} s^j ds D_t, updated
// once per step with the newest interval and summed with the coefficients
// of _kb_row2_fixed_coefs (:1159-1180).  Row n (unknown U^{n+1}):
//   ser_c[n][j] = t_n^{j+1} v3_j  for j < min(P, K_n), else 0, where v3_j and
//     K_n are the cached terms of the _kb_row2 series of [t_{n-1}, t_n] at
//     a = t_{n+1} (eps = series_eps, :2015, :2114-2119; the series stops
//     after the first |term| <= eps);
//   ser_v[n][j] = Da/tau * t_{n+1}^-a prod_{i=1..j} (a+i-1)/(i t_{n+1}).
// One thread per step n; the recurrences are the reference's.
__global__ void k_series_tables(double a, double tau, int M, int P, double eps, double scale,
                                double* __restrict__ ser_c, double* __restrict__ ser_v) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= M) return;
    double* c = ser_c + (size_t)n * P;
    double* v = ser_v + (size_t)n * P;
    if (n == 0 || a == 1.0) {                     // a == 1: no cached terms (:2011-2014)
        for (int j = 0; j < P; j++) { c[j] = 0.0; v[j] = 0.0; }
        if (n == 0) return;
    }
    const double p0 = (n - 1) * tau, p1 = n * tau, pa = (n + 1) * tau;
    if (a != 1.0) {
        double v1 = 1.0, v2 = pow(pa, -a) * p1, v6 = p0 / p1, mult = p1;
        bool live = true;                         // the series has not stopped yet
        int niter = 0;
        for (int j = 0; j < P; j++) {
            double v3 = 0.0;
            if (live) {
                if (j != 0) {
                    double i = (double)j;
                    v1 *= -(-a - i + 1.0) / i;
                    v2 *= p1 / pa;
                    v6 *= p0 / p1;
                    v3 = (1.0 / (i + 1.0)) * (1.0 - v6);
                    if (fabs(v3) > fabs(v3) * eps) niter++;
                } else {
                    v3 = (p1 - p0) / p1;
                }
                c[j] = mult * v3;
                mult *= p1;
                niter++;
                if (!(fabs(v1 * v2 * v3) > eps) || niter > 100000) live = false;
            } else {
                c[j] = 0.0;
            }
        }
    }
    double gtj = pa, w = scale * pow(gtj, -a);
    for (int j = 0; j < P; j++) {
        v[j] = w;
        double i = (double)(j + 1);
        w *= -(-a - i + 1.0) / (i * gtj);
    }
}

// ---------------------------------------------------------------------------
// Generalized Caputo kernel (func_in_kernel 1/2/3, g() and inv_g_der() at
// fpde_hpc_1d.cpp:1889-1930): g(t) = t^gp with gp = 1/2, 2 or func_power.
// The L1 weight of an interval is
//   b(t0,t1;a) = int_{T_t0}^{T_t1} (g(T_a) - g(s))^(-alpha) ds,
// which the reference evaluates by truncated series (_kb_row, :1933-1996,
// _kb_row2, :2000-2131, stop at |term| <= global_eps).  Here it is the
// integral itself, to ~1e-14 relative.  The integrand has its branch points
// at s = T_a and (gp not an integer) s = 0 only:
//  - a history interval away from 0 (t0 >= 1): Gauss-Legendre in s, 16, 8 or
//    4 nodes by the distance of the nearer branch point over the width (at
//    least 1: convergence factor >= 5.8 per node pair);
//  - the oldest interval (t0 = 0): the binomial series
//    (G - s^gp)^-alpha = G^-alpha sum_i (alpha)_i/i! (s/T_a)^(gp i), whose
//    terms integrate in closed form (ratio (T_1/T_a)^gp <= 2^-gp);
//  - the newest interval (t1 = a, the local coefficient kkk0, host):
//    w = (T_a - s)^(1-alpha) moves the singular end into the measure, the
//    bounded rest is integrated by tanh-sinh; at n = 0 the closed form
//    T_1^(1-alpha gp) Gamma(1+1/gp) Gamma(1-alpha) / Gamma(1+1/gp-alpha).
// g(T_a) - g(T_a - d) is formed as -G expm1(gp log1p(-d/T_a)), exact for
// small d.
__constant__ double kGL16x[8] = {0.095012509837637454, 0.28160355077925892, 0.45801677765722737,
                                 0.61787624440264377, 0.755404408355003, 0.86563120238783176,
                                 0.9445750230732326, 0.98940093499164994};
__constant__ double kGL16w[8] = {0.18945061045506859, 0.18260341504492361, 0.16915651939500262,
                                 0.14959598881657676, 0.12462897125553403, 0.095158511682492591,
                                 0.062253523938647706, 0.027152459411754037};
__constant__ double kGL8x[4] = {0.18343464249564978, 0.52553240991632899, 0.79666647741362673,
                                0.96028985649753618};
__constant__ double kGL8w[4] = {0.36268378337836177, 0.31370664587788705, 0.22238103445337434,
                                0.10122853629037669};
__constant__ double kGL4x[2] = {0.33998104358485626, 0.86113631159405257};
__constant__ double kGL4w[2] = {0.65214515486254621, 0.34785484513745368};

__host__ __device__ inline double g_gap(double pa, double gp, double d, double G) {
    return -G * expm1(gp * log1p(-d / pa));      // pa^gp - (pa - d)^gp
}

// b(t0,t1;a) of a history interval, 0 <= p0 < p1 < pa.
__device__ inline double kb_gen(double a, double gp, double p0, double p1, double pa) {
    const double G = pow(pa, gp);
    if (p0 == 0.0) {
        const double r = pow(p1 / pa, gp);
        double c = 1.0, sum = 0.0;                // c = (a)_i / i! r^i
        for (int i = 0; i < 20000; i++) {
            const double term = c / (gp * i + 1.0);
            sum += term;
            if (term <= 1e-17 * sum) break;
            c *= (a + i) / (i + 1.0) * r;
        }
        return pow(G, -a) * p1 * sum;
    }
    const double hw = 0.5 * (p1 - p0), gapr = pa - p1;
    const double dist = fmin(gapr, p0) / (p1 - p0);
    const double* xs = dist < 4.0 ? kGL16x : (dist < 32.0 ? kGL8x : kGL4x);
    const double* ws = dist < 4.0 ? kGL16w : (dist < 32.0 ? kGL8w : kGL4w);
    const int half = dist < 4.0 ? 8 : (dist < 32.0 ? 4 : 2);
    double sum = 0.0;
    for (int k = 0; k < half; k++) {
        // nodes m + hw x and m - hw x; pa - s measured from the right end
        const double dl = gapr + hw * (1.0 - xs[k]), dr = gapr + hw * (1.0 + xs[k]);
        sum += ws[k] * (pow(g_gap(pa, gp, dl, G), -a) + pow(g_gap(pa, gp, dr, G), -a));
    }
    return hw * sum;
}

// The local coefficient's interval b(n,n+1;n+1) = int_{p0}^{pa} (G - s^gp)^-a ds (host).
double kb_newest(double a, double gp, double p0, double pa) {
    if (a == 1.0) return 1.0;                    // kb() at alpha == 1, t1 == a (:2169-2175)
    if (gp == 1.0) return std::pow(pa - p0, 1.0 - a) / (1.0 - a);
    if (p0 == 0.0)
        return std::pow(pa, 1.0 - a * gp) *
               std::exp(std::lgamma(1.0 + 1.0 / gp) + std::lgamma(1.0 - a) - std::lgamma(1.0 + 1.0 / gp - a));
    const double G = std::pow(pa, gp), e1 = 1.0 - a, w0 = std::pow(pa - p0, e1);
    const double h = 1.0 / 32.0;
    double sum = 0.0;
    for (int j = -112; j <= 112; j++) {          // t in [-3.5, 3.5]
        const double t = j * h, u = 0.5 * M_PI * std::sinh(t);
        const double w = w0 / (1.0 + std::exp(-2.0 * u));                    // distance from w = 0
        const double eu = std::exp(u), ch = eu + 1.0 / eu;
        const double wt = h * w0 * 0.5 * (0.5 * M_PI * std::cosh(t)) * 4.0 / (ch * ch);
        const double d = std::fmin(std::pow(w, 1.0 / e1), pa - p0);          // pa - s, rounded into range
        // below d = 1e-16 pa, gap/d is its limit g'(pa) in double; from a subnormal d it is not
        const double phi = d > 1e-16 * pa ? std::pow(g_gap(pa, gp, d, G) / d, -a)
                                          : std::pow(gp * std::pow(pa, gp - 1.0), -a);
        sum += wt * phi;
    }
    return sum / e1;
}

// Prescribed non-uniform grid: the weight of interval t at step n (unknown
// U^{n+1}) is w_n(t) = sum_q cw_q Da_q kb_q(T_t, T_{t+1}; T_{n+1}) / tau_t, the
// kb_3_cache of precalc_kb_3 over Time[] (fpde_hpc_1d.cpp:4069-4085) divided by
// tau[t] in time_summing (:4131).  The closed form ((a-t0)^e - (a-t1)^e)/e,
// e = 1-alpha, is evaluated as (a-t1)^e expm1(e log1p((t1-t0)/(a-t1)))/e; the
// two oldest intervals take the _kb_row2 series in weight_mode 1 (:2265).
// With a generalized kernel (gp != 1) the weights are kb_gen's integrals
// (uniform steps ride on the same per-step rows: they are not Toeplitz).
__device__ inline double kb_grid(double a, double gp, double p0, double p1, double pa, int t, int series, double eps) {
    if (a == 1.0) return 0.0;                    // history intervals have t1 < a
    if (gp != 1.0) return kb_gen(a, gp, p0, p1, pa);
    if (series && t <= 1) return l1_series(a, p0, p1, pa, eps);
    double e = 1.0 - a, x = pa - p1;
    return pow(x, e) * expm1(e * log1p((p1 - p0) / x)) / e;
}

__global__ void k_weights_grid(OrderSet os, const double* __restrict__ Time, const double* __restrict__ taus,
                               int M, int n_first, int rows, int len, int t_lo, int reversed, int series,
                               double eps, double gp, double* __restrict__ W, int64_t ldw) {
    int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)rows * len) return;
    int r = (int)(idx / len), j = (int)(idx % len);
    int n = n_first + r;
    int t = reversed ? n - 1 - j : j;
    double w = 0.0;
    // rows of a block past the grid (n > M) stay zero: Time holds the M + 2
    // points Time[0..M+1]
    if (n <= M && t >= t_lo && t >= 0 && t < n) {
        for (int q = 0; q < os.q; q++)
            w += os.c[q] * kb_grid(os.a[q], gp, Time[t], Time[t + 1], Time[n + 1], t, series, eps);
        w /= taus[t];
    }
    W[(int64_t)r * ldw + j] = w;
}

// one point of a grid built step by step (var_tau)
__global__ void k_set_grid_point(double* Time, double* taus, double* Lv, int n, double t_next, double tau, double L) {
    Time[n + 1] = t_next;
    taus[n] = tau;
    Lv[n] = L;
}

fpde_status launch_set_grid_point(const DeviceWeights& w, int n, double t_next, double tau, double L, hipStream_t st) {
    hipLaunchKernelGGL(k_set_grid_point, dim3(1), dim3(1), 0, st, w.Time, w.taus, w.Lv, n, t_next, tau, L);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

static void gauss_legendre(int n, double a, double b, double* x, double* w) {
    for (int i = 0; i < n; i++) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0, p1 = 1.0, p2 = 0.0;
        for (int it = 0; it < 100; it++) {
            p1 = 1.0; p2 = 0.0;
            for (int j = 1; j <= n; j++) {
                double p3 = p2; p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        p1 = 1.0; p2 = 0.0;
        for (int j = 1; j <= n; j++) {
            double p3 = p2; p2 = p1;
            p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1.0);
        x[n - 1 - i] = 0.5 * (a + b) + 0.5 * (b - a) * z;
        w[n - 1 - i] = (b - a) / ((1.0 - z * z) * pp * pp);
    }
}

static double da_of(double a) { return (a != 1.0) ? 1.0 / std::tgamma(1.0 - a) : 1.0; }

// Orders of a run and their combination weights (host side).
static fpde_status build_orders(const fpde_spec& s, OrderSet& os) {
    if (s.do_q <= 0) {
        os.q = 1;
        os.a[0] = s.alpha;
        os.c[0] = da_of(s.alpha) / s.tau;
        return FPDE_OK;
    }
    if (s.do_q > 64 || (s.do_rule == 1 && s.do_q < 2)) {
        set_error("distributed order: do_q must be in [1,64] (>=2 for the trapezoid)");
        return FPDE_ERR_ARG;
    }
    os.q = s.do_q;
    if (s.do_rule == 0) {
        double x[64], w[64];
        gauss_legendre(s.do_q, s.do_a0, s.do_a1, x, w);
        for (int i = 0; i < s.do_q; i++) {
            os.a[i] = x[i];
            os.c[i] = w[i] * std::tgamma(3.0 - x[i]) / 2.0 * da_of(x[i]) / s.tau;
        }
    } else {
        // fr_do.cpp trapezoid: sum_i 0.5*(WA_i CD_i + WA_{i+1} CD_{i+1}) (:107-108)
        double hstep = (s.do_a1 - s.do_a0) / (s.do_q - 1);
        for (int i = 0; i < s.do_q; i++) {
            double a = s.do_a0 + hstep * i;
            double edge = (i == 0 || i == s.do_q - 1) ? 0.5 : 1.0;
            os.a[i] = a;
            os.c[i] = std::tgamma(3.0 - a) / 2.0 * hstep * edge * da_of(a) / s.tau;
        }
    }
    return FPDE_OK;
}

// orders with c_q = weight_q * Da(a_q) (no 1/tau: a grid divides per interval)
static fpde_status build_orders_grid(const fpde_spec& s, OrderSet& os) {
    fpde_status rc = build_orders(s, os);
    if (rc != FPDE_OK) return rc;
    for (int q = 0; q < os.q; q++) os.c[q] *= s.tau;
    return FPDE_OK;
}

double kernel_exponent(const fpde_spec& s) {
    switch (s.fk) {
        case 1: return 0.5;
        case 2: return 2.0;
        case 3: return s.fp;
        default: return 1.0;
    }
}

// local coefficient of step n on a grid: sum_q c_q Da_q kb_q(n,n+1;n+1)/tau_n
// (kkk0 of init_time_step, fpde_hpc_1d.cpp:3942-3949), Tn = T_n, Tn1 = T_{n+1}
double grid_local(const fpde_spec& s, double Tn, double Tn1) {
    OrderSet og;
    if (build_orders_grid(s, og) != FPDE_OK) return 0.0;
    const double gp = kernel_exponent(s), tn = Tn1 - Tn;
    double L = 0.0;
    for (int q = 0; q < og.q; q++) {
        double a = og.a[q];
        double kkk0 = (a == 1.0) ? 1.0 : (gp == 1.0 ? std::pow(tn, 1.0 - a) / (1.0 - a) : kb_newest(a, gp, Tn, Tn1));
        L += og.c[q] * kkk0 / tn;
    }
    return L;
}

fpde_status launch_weights_grid(const fpde_spec& s, const DeviceWeights& w, int n_first, int rows, int len,
                                int t_lo, int reversed, double* W, int64_t ldw, hipStream_t st) {
    if (rows <= 0 || len <= 0) return FPDE_OK;
    OrderSet os;
    fpde_status rc = build_orders_grid(s, os);
    if (rc != FPDE_OK) return rc;
    int64_t total = (int64_t)rows * len;
    hipLaunchKernelGGL(k_weights_grid, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, os,
                       (const double*)w.Time, (const double*)w.taus, w.M, n_first, rows, len, t_lo, reversed,
                       s.weight_mode == 1 ? 1 : 0, s.series_eps, kernel_exponent(s), W, ldw);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_weights(const fpde_spec& s, DeviceWeights& w, hipStream_t st) {
    OrderSet os;
    fpde_status rc = build_orders(s, os);
    if (rc != FPDE_OK) return rc;
    w.jmax = w.M;                                // every lag kept (restricted summing narrows it below)
    if (w.grid) {
        // prescribed grid: the local coefficient of every step (kkk0 =
        // b(n,n+1;n+1) = tau_n^(1-a)/(1-a), divided by tau_n, :3942-3949);
        // the weights themselves are generated per step or per block
        int M = w.M;
        w.hLv.assign(M, 0.0);
        if (!w.hTime.empty()) {                  // a grid known in advance (taus, or uniform with a generalized kernel)
            for (int n = 0; n < M; n++) w.hLv[n] = grid_local(s, w.hTime[n], w.hTime[n + 1]);
            FPDE_HIP_TRY(hipMemcpyAsync(w.Lv, w.hLv.data(), sizeof(double) * M, hipMemcpyHostToDevice, st));
            FPDE_HIP_TRY(hipStreamSynchronize(st));
        }
        w.local = w.hLv[0];
        FPDE_HIP_TRY(hipMemsetAsync(w.corr, 0, sizeof(double) * 2 * (size_t)M, st));
        return FPDE_OK;
    }
    // local coefficient: Da/tau * kkk0, kkk0 = b(n,n+1;n+1) = tau^(1-a)/(1-a) (alpha=1: 1)
    double local = 0.0;
    for (int q = 0; q < os.q; q++) {
        double a = os.a[q];
        double kkk0 = (a == 1.0) ? 1.0 : std::pow(s.tau, 1.0 - a) / (1.0 - a);
        local += os.c[q] * kkk0;
    }
    w.local = local;
    int M = w.M;
    // Largest kept lag, for skipping history work below the cut.  Bisection
    // on the decreasing closed form, with a relative margin of 1e-9 and one
    // extra lag so that rounding never drops a lag the device kernel keeps.
    w.jmax = M;
    if (s.sum_alg == 1 && s.sum_param > 0.0) {
        int jm = 0;
        for (int q = 0; q < os.q; q++) {
            double a = os.a[q], e = 1.0 - a;
            if (a == 1.0) continue;
            auto b = [&](int j) {
                double jd = (double)j;
                return std::pow(s.tau, e) / e * std::pow(jd, e) * std::expm1(e * std::log1p(1.0 / jd));
            };
            double cut = s.sum_param * (1.0 - 1e-9);
            if (b(1) < cut) continue;
            int lo = 1, hi = M;                      // b(lo) >= cut
            if (b(hi) >= cut) lo = hi;
            while (hi - lo > 1) {
                int mid = lo + (hi - lo) / 2;
                if (b(mid) >= cut) lo = mid; else hi = mid;
            }
            jm = std::max(jm, lo);
        }
        w.jmax = std::min(M, jm + 1);
    }
    int tb = 256;
    double cut = (s.sum_alg == 1) ? s.sum_param : 0.0;
    hipLaunchKernelGGL(k_weights_omega, dim3((M + tb - 1) / tb), dim3(tb), 0, st, os, s.tau, M, cut, w.omega);
    if (s.weight_mode == 1 || s.sum_alg == 1) {
        hipLaunchKernelGGL(k_weights_corr, dim3((2 * M + tb - 1) / tb), dim3(tb), 0, st, os, s.tau, M,
                           s.series_eps, s.weight_mode == 1 ? 1 : 0, cut, (const double*)w.omega, w.corr);
    } else {
        FPDE_HIP_TRY(hipMemsetAsync(w.corr, 0, sizeof(double) * 2 * (size_t)M, st));
    }
    if (s.sum_alg == 2) {
        if (!w.ser_c || !w.ser_v || w.ser_P < 1) { set_error("series tables not allocated"); return FPDE_ERR_ARG; }
        hipLaunchKernelGGL(k_series_tables, dim3((M + 63) / 64), dim3(64), 0, st, os.a[0], s.tau, M, w.ser_P,
                           s.series_eps, os.c[0], w.ser_c, w.ser_v);
    }
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

// Space-fractional scheme: calc_v (fpde_hpc_1d.cpp:2303-2335) weighs F[j] by
// Dg kb(gamma, j-1, j, i; Z) on the node grid Z_k = k h.  That is the time
// weight of interval t = j-1 at step n = i-1 with tau -> h, alpha -> gamma,
// Da -> Dg = 1/Gamma(1-gamma) (:3602-3603) and no division by a step: the
// same closed-form lag vector (lag d = i-j) and the same series corrections
// of the two intervals next to the surface (j = 1, 2; kb's series branch
// for t0 <= 1, :2265).  The row factor Dg kb(gamma, i-1, i, i) of A_/B/R
// (:3095-3120) is the closed form of the newest interval (t1 == a).
fpde_status launch_space_weights(const fpde_spec& s, DeviceWeights& w, double h, hipStream_t st) {
    w.sp_on = false;
    w.sp_kap = 1.0;
    if (!(s.gamma > 0.0 && s.gamma < 1.0)) return FPDE_OK;
    if (!w.sp_omega || !w.sp_corr) { set_error("space-fractional weights not allocated"); return FPDE_ERR_ARG; }
    OrderSet os;
    os.q = 1;
    os.a[0] = s.gamma;
    os.c[0] = 1.0 / std::tgamma(1.0 - s.gamma);
    const int N = s.nodes, tb = 256;
    hipLaunchKernelGGL(k_weights_omega, dim3((N + tb - 1) / tb), dim3(tb), 0, st, os, h, N, 0.0, w.sp_omega);
    if (s.weight_mode == 1) {
        hipLaunchKernelGGL(k_weights_corr, dim3((2 * N + tb - 1) / tb), dim3(tb), 0, st, os, h, N, s.series_eps, 1,
                           0.0, (const double*)w.sp_omega, w.sp_corr);
    } else {
        FPDE_HIP_TRY(hipMemsetAsync(w.sp_corr, 0, sizeof(double) * 2 * (size_t)N, st));
    }
    FPDE_HIP_TRY(hipGetLastError());
    const double e = 1.0 - s.gamma;
    w.sp_kap = os.c[0] * std::pow(h, e) / e;
    w.sp_on = true;
    return FPDE_OK;
}

}  // namespace fpde
