// Internal declarations shared by the HIP translation units of libfpde_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <map>
#include <vector>

#include "../../include/fpde_hip.h"

namespace fpde {

// ---------------------------------------------------------------------------
// error plumbing: every C-ABI entry returns an fpde_status and records text.
void set_error(const std::string& msg);
#define FPDE_HIP_TRY(expr)                                                     \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::fpde::set_error(std::string(#expr " failed: ") +                 \
                              hipGetErrorString(e_));                          \
            return FPDE_ERR_HIP;                                               \
        }                                                                      \
    } while (0)

// ---------------------------------------------------------------------------
// van Genuchten-Mualem soil constants (fpde_hpc_1d.cpp:2513-2524).
constexpr double kVgmN = 1.25;
constexpr double kVgmS0 = 0.0736;
constexpr double kVgmS1 = 0.55;
constexpr double kVgmA = 0.066;

// Conductivity K = ksat*Kr and the inverse water capacity 1/C at suction head
// P = z - u (metres).  Same quantities as KK()/inv_dw_dh() of the reference
// (fpde_hpc_1d.cpp:2527-2575, :2625-2657, cv_form 4, k_form 1), rewritten with
// y = 100 a P and x = y^n:  theta - s0 = (s1-s0) Se,  Se = (1+x)^(-m),
// Kr = Se^(1/2) (1 - (x/(1+x))^m)^2,  C = d theta/dh = (n-1)(s1-s0) x Se / (P (1+x)),
// m = 1 - 1/n.  With the reference's n = 5/4 (m = 1/5) every power is a root:
//   b = y^(1/4) = x^(1/5) (two square roots),  x = y b,  w = 1 + x,
//   Se = r = w^(-1/5),  c = (x/(1+x))^(1/5) = r b,
// and since (1 - c)(1 + c + c^2 + c^3 + c^4) = 1 - c^5 = 1/w,
//   1 - c = 1 / (w (1 + c + c^2 + c^3 + c^4)),
// a sum of positive terms: no digits cancel where c -> 1 (dry soil), where
// the difference of logarithms loses them (1e-10 relative in Kr at P ~ 100 m
// for the log/expm1 form).  r comes from a single-precision seed and two
// Newton steps for the inverse fifth root, r <- r + (r/5)(1 - w r^5) (error
// e -> 3 e^2: 1e-7 -> 3e-14 -> 1e-27), multiplications only; w beyond the
// single-precision range takes exp(-log(w)/5).  Three square roots and two
// divisions in all; within 3e-15 of the exact K and 1/C for P in [1e-8, 1e3].
static_assert(kVgmN == 1.25, "vgm_coefs takes fifth and fourth roots: n = 5/4");
__host__ __device__ inline void vgm_coefs(double u, double z, double ksat, int mode,
                                          double& K, double& invC) {
    if (mode == 1) { K = 1.0; invC = 1.0; return; }
    double P = z - u;
    if (P <= 0.0) {               // saturated: wetness = s1, Se = 1, Kr = 1, C = 0
        K = ksat;
        invC = 1e20;
        return;
    }
    const double y = 100.0 * kVgmA * P;
    const double b = sqrt(sqrt(y));                 // y^(1/4)
    const double x = y * b;                         // y^(5/4)
    const double w = 1.0 + x;
    double r;                                       // w^(-1/5) = Se
    if (w < 1e30) {
        r = (double)exp2f(-0.2f * log2f((float)w));
        for (int it = 0; it < 2; it++) {
            const double r2 = r * r;
            r = fma(r * 0.2, fma(-w, r2 * r2 * r, 1.0), r);
        }
    } else {
        r = exp(-0.2 * log(w));
    }
    const double c = r * b;                         // (x/(1+x))^(1/5)
    const double T = (((c + 1.0) * c + 1.0) * c + 1.0) * c + 1.0;
    const double t = 1.0 / (w * T);                 // 1 - c
    K = ksat * sqrt(r) * t * t;
    invC = (x != 0.0) ? 4.0 * w * P / ((kVgmS1 - kVgmS0) * x * r) : 1e20;   // (n-1) = 1/4
}

// 1/x for the normalised pivots of the tridiagonal reductions: the hardware
// reciprocal refined by two Newton steps (within an ulp or two of 1/x; the
// pivots are O(1) or well inside the normal range).
__device__ inline double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Near-field history folded ahead in groups of 16 levels (1D cooperative
// step, 3D history combine): the group [lo, lo + 16) of node i goes into
// mid for the steps p in [p_lo, p_hi) of the history block that starts at
// step n0:
//   mid[p - n0][i] = (first ? 0 : mid[p - n0][i]) + sum_k w_p[p - lo - k - 1] D[lo + k][i],
// w_p = omega + (p - n0) * omega_stride (one Toeplitz vector, stride 0, or
// the step's own row).  The steps go in 8-wide batches: their mid values
// load together, then the sums.
constexpr int kNearGroup = 16;
__device__ inline void fold_near_group(const double* __restrict__ D, int64_t ld, const double* __restrict__ omega,
                                       int64_t omega_stride, int n0, double* __restrict__ mid, int64_t i, int lo,
                                       int p_lo, int p_hi, bool first) {
    double dv[kNearGroup];
#pragma unroll
    for (int k = 0; k < kNearGroup; k++) dv[k] = D[(int64_t)(lo + k) * ld + i];
    constexpr int PB = 8;
    for (int p0 = p_lo; p0 < p_hi; p0 += PB) {
        double acc[PB];
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int p = p0 + j;
            acc[j] = (!first && p < p_hi) ? mid[(int64_t)(p - n0) * ld + i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int p = p0 + j;
            if (p < p_hi) {
                const double* om = omega + (int64_t)(p - n0) * omega_stride + (p - lo - 1);
#pragma unroll
                for (int k = 0; k < kNearGroup; k++) acc[j] = fma(om[-k], dv[k], acc[j]);
            }
        }
#pragma unroll
        for (int j = 0; j < PB; j++) {
            const int p = p0 + j;
            if (p < p_hi) mid[(int64_t)(p - n0) * ld + i] = acc[j];
        }
    }
}

// one DPP move of a double (both 32-bit halves) inside the rows of 16 lanes
template <int CTRL>
__device__ inline double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}

// max over the 64 lanes of a wave, in every lane (called with all lanes
// active): rotations inside each row of 16 lanes (DPP row_ror 8, 4, 2, 1)
// leave every lane with its row's maximum; the four row maxima are then read
// from lanes 0, 16, 32 and 48
__device__ inline double wave_max(double v) {
    v = fmax(v, dpp_move<0x128>(v));
    v = fmax(v, dpp_move<0x124>(v));
    v = fmax(v, dpp_move<0x122>(v));
    v = fmax(v, dpp_move<0x121>(v));
    auto lane_val = [&](int l) {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                                __builtin_amdgcn_readlane(__double2loint(v), l));
    };
    return fmax(fmax(lane_val(0), lane_val(16)), fmax(lane_val(32), lane_val(48)));
}

// ---------------------------------------------------------------------------
// Weight sets on the device for a uniform step (see weights.hip):
//   local   : coefficient of (U^{n+1}-U^n) in the discrete derivative
//   omega[j-1], j=1..M : Toeplitz weight of D[n-j] (already x Da/tau, folded over orders)
//   corr[2n+t], t=0,1  : reference-series correction of the weight of D[t] at step n
struct DeviceWeights {
    double local = 0.0;
    double* omega = nullptr;   // [M]
    double* corr = nullptr;    // [2*M]
    int M = 0;
    int jmax = 0;              // no weight beyond this lag is non-zero (restricted summing)
    // series-compressed history (sum_alg 2): per step n, the moment update
    // coefficients ser_c[n][j] and the series coefficients ser_v[n][j] (Da/tau
    // folded in), j < ser_P
    int ser_P = 0;
    double* ser_c = nullptr;   // [M][P]
    double* ser_v = nullptr;   // [M][P]
    // prescribed non-uniform grid (spec.taus): cumulative times, step lengths,
    // the local coefficient of every step (host and device copies)
    bool grid = false;
    std::vector<double> hTime; // host copy of Time when the grid is known at create
    double* Time = nullptr;    // [M+2]
    double* taus = nullptr;    // [M+1]
    double* Lv = nullptr;      // [M]
    std::vector<double> hLv;
    // space-fractional scheme (spec.gamma in (0,1)), on the node grid Z_k = k h:
    // sp_omega[d-1] = Dg kb(gamma, j-1, j, i) for d = i-j >= 1 (closed form,
    // Toeplitz), sp_corr[2(i-1)+t] = series - Toeplitz weight of the interval
    // j = t+1 in row i (the reference's series for t0 <= 1), sp_kap = the row
    // factor Dg kb(gamma, i-1, i, i) = Dg h^(1-gamma)/(1-gamma)
    bool sp_on = false;
    double sp_kap = 1.0;
    double* sp_omega = nullptr;  // [N]
    double* sp_corr = nullptr;   // [2N]
};

// Dense weight rows of a prescribed grid (weights.hip): rows r < rows are
// the steps n = n_first + r; entry j of row r is w_n(t) for t = n-1-j
// (reversed: omega-style, w_n(t) at index n-t-1) or t = j (direct), and 0
// outside t in [t_lo, n).
fpde_status launch_weights_grid(const fpde_spec& s, const DeviceWeights& w, int n_first, int rows, int len,
                                int t_lo, int reversed, double* W, int64_t ldw, hipStream_t st);
// local coefficient of the step from Tn to Tn1: sum_q cw_q Da_q kkk0 / (Tn1 - Tn),
// kkk0 = b(n,n+1;n+1) (init_time_step, fpde_hpc_1d.cpp:3942-3949): (Tn1-Tn)^(1-a)/(1-a)
// for g(t) = t, the integral of the generalized kernel otherwise
double grid_local(const fpde_spec& s, double Tn, double Tn1);
// exponent gp of the kernel function g(t) = t^gp of spec.fk (1 for fk 0)
double kernel_exponent(const fpde_spec& s);
// int_{p0}^{pa} (pa^gp - s^gp)^-a ds (host; the newest interval's weight)
double kb_newest(double a, double gp, double p0, double pa);
fpde_status launch_set_grid_point(const DeviceWeights& w, int n, double t_next, double tau, double L, hipStream_t st);

constexpr int kMaxSeriesTerms = 64;

// sum_alg 2 moment update and series of one node (time_summing,
// fpde_hpc_1d.cpp:4143-4150): H[j][i] += c_j d, S = sum_j v_j H[j][i] in the
// reference's order (j ascending).  Loads are issued 8 at a time.
__device__ inline double series_update(double* __restrict__ H, int64_t ld, int64_t i, int P, double d,
                                       const double* __restrict__ c, const double* __restrict__ v) {
    double s = 0.0;
    for (int j0 = 0; j0 < P; j0 += 8) {
        double h[8];
#pragma unroll
        for (int j = 0; j < 8; j++) h[j] = (j0 + j < P) ? H[(int64_t)(j0 + j) * ld + i] : 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            if (j0 + j < P) {
                double hn = h[j] + c[j0 + j] * d;
                H[(int64_t)(j0 + j) * ld + i] = hn;
                s += v[j0 + j] * hn;
            }
        }
    }
    return s;
}

// history storage modes
enum HistMode : int {
    HIST_PLAIN_F64 = 0,   // FP64 streaming sum over every level, every step
    HIST_MFMA_F64 = 1,    // time-block GEMM on v_mfma_f64_16x16x4
    HIST_SPLIT_BF16 = 2,  // time-block GEMM, bf16 hi+lo, fp32 accumulate
    HIST_SPLIT_F16 = 3,   // time-block GEMM, fp16 hi+lo with per-group scaling
    HIST_SPLIT_BF16X3 = 4 // time-block GEMM, bf16 hi+mid+lo
};

constexpr int kGroup = 32;     // history levels per packed group (K of one 16x16x32 MFMA)


// 1D chunks of the partitioned solve (256 nodes each, N <= 65536)
constexpr int kMaxChunks = 256;
// grid exchange of the cooperative 1D launches: per barrier parity, 8 words
// per reduced row (2 per chunk) and 2 per chunk measure, 8 bytes each
constexpr size_t kXchBytes = 2 * (size_t)(2 * (8 * kMaxChunks + kMaxChunks)) * 8;

// Arguments of one 1D step launch (step1d.hip)
struct Step1DArgs {
    int N, C;
    double h, h2inv, ksat, L, utop;
    int coef_mode;
    int n;          // step: unknown is U^{n+1}
    int k;          // Picard iteration (0 first), -1 = finalize only
    int recover;    // a reduced system from the previous launch is pending
    int parity;     // reduced-system buffer written by this launch
    double* U;      // U^n
    double* D;      // FP64 history [levels][ld]
    int64_t ld;
    const double* omega;
    const double* corr;
    const double* part;       // far partial sums
    int nsplit;
    int64_t part_stride, part_off;
    int near_lo;              // near terms t in [near_lo, n)
    double* Sbuf;             // history term of step n, reused by Picard iterations
    double* xyz;              // [3][ld] chunk factors
    double* red;              // [2][2*C][4]
    double* Uit;              // Picard iterate (for the convergence test)
    double* dmax;             // [2][C] per-chunk max |U^k - U^{k-1}|
    int* flags;               // unused
    double tol;
    int64_t omega_stride;     // per-step weight rows (prescribed grid): omega + (n - n0) * stride
    const double* Lv;         // per-step local coefficient (prescribed grid), else L
    int prev_k;               // Picard index of the previous launch (-1: none)
    int extrap;               // Picard starts from 2 U^n - U^{n-1} (picard_max > 1)
    double* Uit0;             // that initial iterate (convergence test of iteration 1)
    double* uend;             // [2][2C] U^n at the chunk ends, by step parity
    // cooperative block form (k_step1d_block): steps [n, n_end) in one launch
    int n_end;
    int picard_max;
    int n0;                   // first step of the history block (far partial row = n - n0)
    int blk;                  // rows of the history block (far partial and mid rows)
    double* mid;              // near-field groups folded ahead: mid[p - n0][i] = sum over the
                              // folded 16-level groups of the step-p weights times D (null: none)
    unsigned long long* xch;  // grid exchange words [2 parities][kXchBytes / 16] (tagged halves)
    unsigned xtag;            // this launch's tag base: (launch count mod 65535 + 1) << 16
    int* err;                 // [0] set when a grid barrier times out; [1] Picard solves (cooperative launches)
    int64_t wait_ticks;       // a grid barrier gives up after this many ticks of the device's wall clock
    // series-compressed history (ser_P > 0): moments H [P][ld], coefficient
    // rows of step n; D is then a ring of two levels (row t & 1)
    int ser_P;
    double* H;
    const double* ser_c;
    const double* ser_v;
    // space-fractional scheme (sp_on): rows scaled by sp_kap and the explicit
    // sum E_i = sum_y sp_P[y][i] + sp_corr-weighted F_1, F_2 subtracted as
    // inv_dw_dh E_i; E of step n is kept in sp_E by the launch that forms it
    int sp_on;
    double sp_kap;
    double* sp_P;             // [chunks][sp_ldp] partial sums of the Toeplitz part
    int64_t sp_ldp;
    double* sp_F;             // F of U^n
    const double* sp_corr;
    double* sp_E;
};

// The fields of a batched cooperative launch that change from block to block:
// solve j runs steps [n[j], n_end[j]) on workgroups [first[j], first[j+1]).
constexpr int kMaxBatch = 32;
struct BatchTick {
    int count;
    int first[kMaxBatch + 1];
    int n[kMaxBatch], n_end[kMaxBatch], n0[kMaxBatch], nsplit[kMaxBatch], near_lo[kMaxBatch];
    unsigned xtag[kMaxBatch];
    int64_t part_stride[kMaxBatch];
    const double* part[kMaxBatch];    // the block's partial rows (a solve's buffer alternates by block)
};

// ---------------------------------------------------------------------------
// Space-fractional sum (space.hip): rows are cut into tiles of kSpTile rows
// (16 blocks of 16) and the 16-wide block diagonals d of the lower-triangular
// Toeplitz matrix into chunks of kSpDiag; a tile t has block diagonals
// d < 16 (t + 1) and space_chunks(t) partial rows.
constexpr int kSpTile = 256;
constexpr int kSpDiag = 64;
__host__ __device__ inline int space_chunks(int t) { return (16 * (t + 1) + kSpDiag - 1) / kSpDiag; }

typedef double sd4_t __attribute__((ext_vector_type(4)));

// LDS of one work unit: the F blocks J in [I0 - d_hi + 1, I0 + 15 - d_lo]
// with a stride of 17 per block (the B reads of one 16-lane row fall into
// distinct banks), the weights w(m), m in [16 d_lo - 15, 16 d_hi - 1], and
// the four waves' results.
struct SpaceUnitLds {
    double Fs[(kSpDiag + 15) * 17];
    double Ws[16 * kSpDiag + 15];
    double red[4][4][64];
};

// One work unit of the spatial sum: P[y][i] = sum over the block diagonals d
// in [kSpDiag y, kSpDiag y + kSpDiag) of row tile t (rows i = 256 t + 16 I'
// + p) of sum_q w(16 d + p - q) F[16 (I0 + I' - d) + q], I0 = 16 t, with
// w(m) = om[m-1] for 1 <= m <= N, else 0 (om: the Toeplitz lag vector).
// Wave w takes 16 of the diagonals; the four waves' results are added
// through LDS.  v_mfma_f64_16x16x4f64 layout (as in the history GEMM): A
// lane l = A[l%16][l/16], B lane l = B[l/16][l%16], result v of lane l =
// D[4v + l/16][l%16]; M index p, N index I', K index q.  Called by all 256
// threads of a workgroup (it synchronises the workgroup).
__device__ inline void space_unit(const double* __restrict__ F, int N, const double* __restrict__ om,
                                  double* __restrict__ P, int64_t ldp, int t, int y, SpaceUnitLds& L) {
    const int tid = threadIdx.x;
    const int I0 = 16 * t;
    const int d_lo = kSpDiag * y;
    const int d_hi = min(d_lo + kSpDiag, I0 + 16);
    const int Jb = I0 - d_hi + 1;
    const int nfb = d_hi - d_lo + 15;
    const int mb = 16 * d_lo - 15;
    const int nw = 16 * (d_hi - d_lo) + 15;
    __syncthreads();                                   // LDS free (a previous unit may still read it)
    for (int q = tid; q < nfb * 16; q += 256) {
        const int b = q >> 4, c = q & 15;
        const int64_t j = (int64_t)(Jb + b) * 16 + c;
        L.Fs[b * 17 + c] = (j >= 0 && j < ldp) ? F[j] : 0.0;
    }
    for (int q = tid; q < nw; q += 256) {
        const int m = mb + q;
        L.Ws[q] = (m >= 1 && m <= N) ? om[m - 1] : 0.0;
    }
    __syncthreads();
    const int lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    sd4_t acc = {0.0, 0.0, 0.0, 0.0};
    const int dw_lo = d_lo + 16 * wave, dw_hi = min(dw_lo + 16, d_hi);
    for (int d = dw_lo; d < dw_hi; d++) {
        const int fb = (r16 + d_hi - 1 - d) * 17 + k4;   // F block I0 + I' - d, element 4 kq + k4
        const int wb = 16 * d + r16 - k4 - mb;           // w(16 d + p - q), q = 4 kq + k4
#pragma unroll
        for (int kq = 0; kq < 4; kq++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(L.Ws[wb - 4 * kq], L.Fs[fb + 4 * kq], acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; v++) L.red[wave][v][lane] = acc[v];
    __syncthreads();
    const int v = tid >> 6, l = tid & 63;
    const double sum = (L.red[0][v][l] + L.red[1][v][l]) + (L.red[2][v][l] + L.red[3][v][l]);
    const int64_t i = (int64_t)kSpTile * t + 16 * (l & 15) + 4 * v + (l >> 4);
    P[(int64_t)y * ldp + i] = sum;
}

// E_i of an interior row i (1 <= i <= N-2) from the partials and the two
// surface-interval corrections
__device__ inline double space_sum(const double* __restrict__ P, int64_t ldp, const double* __restrict__ corr,
                                   const double* __restrict__ F, int i) {
    const int ny = space_chunks(i / kSpTile);
    double e = 0.0;
    for (int y = 0; y < ny; y++) e += P[(int64_t)y * ldp + i];
    e = fma(corr[2 * (i - 1)], F[1], e);
    return fma(corr[2 * (i - 1) + 1], F[2], e);
}

// Kernel launchers (weights.hip, history.hip, space.hip, step1d.hip, lod3d.hip)
fpde_status launch_weights(const fpde_spec& s, DeviceWeights& w, hipStream_t st);
fpde_status launch_space_weights(const fpde_spec& s, DeviceWeights& w, double h, hipStream_t st);
fpde_status launch_space_flux(const double* U, int N, double h, double ksat, int mode, double* F, int64_t ldp,
                              hipStream_t st);
fpde_status launch_space_apply(const double* F, int N, const double* sp_omega, double* P, int64_t ldp,
                               hipStream_t st);
fpde_status launch_space_reduce(double* V, const double* P, const double* F, const double* sp_corr, int N,
                                int64_t ldp, hipStream_t st);
// plain FP64 stream over t < t_hi: partial rows part[ks*ld + i], ks < the
// returned row count of hist_plain_plan (at most kPlainMaxSplits)
constexpr int kPlainMaxSplits = 64;
int hist_plain_plan(int nodes, int t_hi, int* chunk_out);
fpde_status launch_hist_plain(const double* D, int64_t ld, int nodes, int n, int t_hi, const double* omega,
                              double* part, int* splits_out, hipStream_t st);
fpde_status launch_ref_sum(double* S, const double* old, int64_t old_size, int64_t os, int64_t nodes,
                           const double* tau, const double* kb, int sum_alg, double sum_param, int bs,
                           hipStream_t st);
fpde_status launch_ref_series(double* S, const double* old, int64_t old_size, int64_t os, int64_t nodes,
                              const double* tau, int tstep, double* Htmp, const double* coef, int P, double gtj,
                              double gt, double alpha, hipStream_t st);
// The GEMM launchers write K-split partial rows P[ks][r][i], ks < *splits_out
// (at most max_splits when that is > 0); the readers add them in split order:
// the 3D history combine itself, the 1D step kernels after launch_part_reduce
// has folded them into P[0].
fpde_status launch_gemm_f64(const double* D, int64_t ld, int nodes, int n0, int tc, int block,
                            const double* omega, double* P, int* splits_out, hipStream_t st,
                            int max_splits = 0);
fpde_status launch_part_reduce(double* P, int splits, int rows, int64_t ld, int nodes, hipStream_t st);
fpde_status launch_pack_split(int mode, const double* D, int64_t ld, int nodes, int g_lo, int g_hi, void* c0,
                              void* c1, void* c2, float* scale, hipStream_t st);
fpde_status launch_gemm_f64_dense(const double* D, int64_t ld, int nodes, int tc, int block, const double* W,
                                  int64_t ldw, double* P, int* splits_out, hipStream_t st,
                                  int max_splits = 0);
// W != nullptr: dense weight rows W[r][t] (leading dimension ldw >= tc) of a
// prescribed grid in place of the Toeplitz vector omega.  afrag / ainv:
// workspace of split_afrag_bytes(mode, block, tc / 32) bytes and
// (tc / 32) * (block / 16) doubles for the block's A fragments.
size_t split_afrag_bytes(int mode, int block, int groups);
fpde_status launch_gemm_split(int mode, const void* c0, const void* c1, const void* c2, const float* scale,
                              int64_t ld, int nodes, int n0, int tc, int block, const double* omega,
                              const double* W, int64_t ldw, void* afrag, double* ainv, double* P,
                              int* splits_out, hipStream_t st, int max_splits = 0);
fpde_status launch_step1d(const Step1DArgs& a, hipStream_t st);
fpde_status launch_step1d_block(const Step1DArgs& a, hipStream_t st);
bool step1d_block_supported(int C);
bool step1d_batch_supported(int64_t wgs);
// Batched 3D LOD step: one step of up to kMax3DBatch 3D solves (n <= 128,
// partitioned sweeps) in one launch per kernel of the step (history combine,
// coefficients, y sweep, coefficients, x sweep, z sweep).  Workgroups
// [first_*[j], first_*[j+1]) of each launch belong to solve j; each solve's
// row holds what its single-solve launches take.
constexpr int kMax3DBatch = 8;
struct Lod3DRow {
    double *U, *S, *Uprev, *Dlev;
    const double *part, *D, *omega, *corr, *utop;
    int64_t part_stride, part_off, ld, nodes;
    int n, nsplit, step, near_lo, coef_mode, split;
    // history block: first step, rows, weight-row stride, folded near groups
    int n0, blk;
    int64_t omega_stride;
    double* mid;
    double h, h2inv, ksat, L, hscale;
    // arguments the solo sweeps take that a batched solve leaves at their
    // neutral values: kap = 1 and no line sums E (no space-fractional term),
    // no coefficient arrays for the depth (Kz, iCz) and lateral (Kl, iCl)
    // sweeps, which then evaluate K and 1/C themselves.  They are run-time
    // values so that the batched kernels compile to the solo kernels'
    // arithmetic: a literal 1.0 or null would let the compiler fold them and
    // fuse products into FMAs differently, changing the rounding
    double kap;
    const double *E, *Kz, *iCz, *Kl, *iCl;
};
struct Lod3DBatch {
    int count;
    int first_node[kMax3DBatch + 1];   // 256-node blocks (history combine, coefficients)
    int first_lat[kMax3DBatch + 1];    // 32-line blocks (lateral sweeps)
    int first_z[kMax3DBatch + 1];      // 16-line blocks (depth sweep)
    Lod3DRow row[kMax3DBatch];
};
fpde_status launch_lod3d_batch(const Lod3DBatch& b, hipStream_t st);
constexpr int kZLinesHost = 16;      // depth lines per workgroup of the partitioned z sweep
bool solver_3d_batchable(const fpde_solver* s);
fpde_status advance_3d_batch(fpde_solver* const* v, int count, int nsteps);
// 3D solves in batched LOD steps (groups of kMax3DBatch), the rest on their own streams
fpde_status advance_riders(fpde_solver* const* v, int count, int nsteps);

// One launch of fr_do.cpp's program kernels (frdo.hip), decoded by the facade.
enum { FRDO_FILL = 0, FRDO_DOC, FRDO_SUM, FRDO_EQ_F, FRDO_EQ_SOLVE };
struct FrdoLaunch {
    int kernel = 0;
    bool is_double = true;      // the program's element type T
    int64_t n = 0;              // Fill: levels
    int nf = 1, ntau = 0, ntau2 = 0, na = 0, t0 = 0, j0 = 0, maxt = -1, use_f = 0, append = 0;
    int64_t tot = 0;            // DOC: levels per grid point of F (total_ntau)
    int nx = 0, Nt = 0, j = 0;  // eq_F / eq_Solve (j: the level, t of eq_Solve)
    double tau = 0, D = 0, b0 = 0, h = 0, A = 0, ex = 0;
    void* b[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
};
fpde_status launch_frdo(const FrdoLaunch& L, hipStream_t st);

int64_t barrier_wait_ticks();
bool solver_batchable(const fpde_solver* s);
int solver_chunks(const fpde_solver* s);

// Device memory arena of the request server: segments of device memory
// allocated when no free block fits (at least seg_bytes each, the sum capped
// at limit), handed out first-fit in 256-byte units and returned with
// coalescing, so that a steady stream of admissions and retirements reuses
// the same memory without hipMalloc/hipFree.  Not thread-safe (the server's
// scheduler owns it).
class DeviceArena {
public:
    DeviceArena(size_t seg_bytes, size_t limit) : seg_bytes_(seg_bytes), limit_(limit) {}
    ~DeviceArena();
    void* alloc(size_t bytes);          // nullptr when nothing fits within the limit
    void release(void* p);
    size_t reserved() const { return reserved_; }   // device memory held in segments
    size_t in_use() const { return in_use_; }
    int64_t segment_allocations() const { return (int64_t)segs_.size(); }
private:
    struct Block { size_t size; int seg; };
    size_t seg_bytes_, limit_, reserved_ = 0, in_use_ = 0;
    std::vector<std::pair<char*, size_t>> segs_;
    std::map<char*, Block> free_, live_;
};

// A solver whose device buffers come from `arena` and whose launches go to
// `stream` (both owned by the caller, the request server): destroying it
// synchronises the stream and returns the buffers to the arena.
fpde_status solver_create_in(const fpde_spec* spec, DeviceArena* arena, hipStream_t stream, fpde_solver** out);
fpde_status launch_step1d_batch(const Step1DArgs* base, const BatchTick& t, hipStream_t st);
fpde_status launch_coef3d(const double* U, int n, double h, double ksat, int mode, double* K, double* iC,
                          hipStream_t st);
// K, iC null: the partitioned sweep (lateral3d_partitioned) evaluates the
// coefficients at U itself; the one-thread-per-line sweep needs them
fpde_status launch_lateral3d(double* U, const double* K, const double* iC, const double* H, int n, int axis,
                             double h2inv, double L, double* Al, double* Bt, double kap, const double* E,
                             double h, double ksat, int mode, hipStream_t st);
inline bool lateral3d_partitioned(int n, double L) { return n <= 128 && L > 0.0; }
// K, iC: the coefficients at U when a coefficient pass already formed them
// (else nullptr: the sweep evaluates them itself)
fpde_status launch_zline3d(double* U, double* Uprev, const double* utop, int n, double h, double h2inv,
                           double ksat, int mode, double L, double* Dlev, double kap, const double* E,
                           const double* K, const double* iC, const double* H, hipStream_t st);
fpde_status launch_space3d(const double* U, const double* K, int n, int axis, double h, const double* sp_omega,
                           const double* sp_corr, double* E, hipStream_t st);
fpde_status launch_series3d(double* S, double* H, int P, const double* dnew, const double* ser_c,
                            const double* ser_v, int64_t nodes, int64_t ld, double hscale, hipStream_t st);
// 3D history term of step n (far partials, near levels [near_lo, n),
// corrections); omega is the weight row of the block's first step n0 (rows
// omega_stride apart on a prescribed grid, else one Toeplitz vector); with
// mid (a block of blk rows) the near levels go in 16-level groups folded
// ahead, each at the step that completes it.
fpde_status launch_hist_combine(double* S, const double* part, int nsplit, int64_t part_stride, int64_t part_off,
                                const double* D, int64_t ld, int64_t nodes, int n, int near_lo,
                                const double* omega, int64_t omega_stride, int n0, int blk, double* mid,
                                const double* corr, double hscale, hipStream_t st);
int gemm_max_splits(int64_t nodes);

}  // namespace fpde
