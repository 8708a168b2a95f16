// 3D locally-one-dimensional step (gfx950), fpde_hpc_1d.cpp:5144-5626.
//
// Field layout idx = z + n*(x + n*y): depth z fastest (the reference's N axis,
// idx() at :1181).  One step = three sweeps in the reference order:
//   y (m3d_c=2, carries the history term), x (m3d_c=1), z (m3d_c=0); with
//   split3d odd each of the three carries a third of it.
// Every sweep solves, line by line, L*(U - U_in) = div_axis(K grad U) + H
// with coefficients evaluated at the field the sweep starts from.
//  - y and x lines: one thread per line (adjacent threads = adjacent z, so
//    every load is coalesced) running the reference Thomas recurrences
//    al1/bt1/U1 (:4002-4027, :4585-4590, :4672-4725) with zero flux at both
//    ends (Al[1]=1, Bt[1]=0; bcN()=0), coefficients precomputed per node;
//  - z lines: one workgroup per line, one node per thread, PCR in LDS,
//    Dirichlet top / zero-flux bottom; it also writes D[n] = U^{n+1}-U^n.
// With the space-fractional scheme every sweep also takes kap and the line
// sums E of its own axis (space.hip, k_space3d).
#include "fpde_internal.h"

namespace fpde {

// K and 1/C of every node at the current field.
__device__ __forceinline__ void coef3d_body(const double* __restrict__ U, int n, double h, double ksat, int mode,
                                            double* __restrict__ K, double* __restrict__ iC, int bid) {
    int64_t id = (int64_t)bid * 256 + threadIdx.x;
    int64_t nn = (int64_t)n * n * n;
    if (id >= nn) return;
    int z = (int)(id % n);
    double k, ic;
    vgm_coefs(U[id], z * h, ksat, mode, k, ic);
    K[id] = k;
    iC[id] = ic;
}
__global__ __launch_bounds__(256) void k_coef3d(const double* __restrict__ U, int n, double h,
                                                double ksat, int mode, double* __restrict__ K,
                                                double* __restrict__ iC) {
    coef3d_body(U, n, h, ksat, mode, K, iC, blockIdx.x);
}

// Lateral sweep, thread per line.  axis 1 = x (stride n), 2 = y (stride n*n).
// Scratch Al/Bt: [n+1][nlines], coalesced.  The row data of a line are
// fetched 16 rows at a time (independent loads) ahead of the recurrence, and
// the back substitution fetches Al/Bt the same way, so a line costs ~n/16
// memory round trips instead of n.
// Space-fractional scheme (E != nullptr): the diffusion part of the rows is
// scaled by kap = Dg kb(gamma,i-1,i,i) and Om carries -inv_dw_dh E (E: the
// line sums of k_space3d), as in the 1D rows.
__global__ __launch_bounds__(256) void k_lateral3d(double* __restrict__ U, const double* __restrict__ K,
                                                   const double* __restrict__ iC,
                                                   const double* __restrict__ H, int n, int axis,
                                                   double h2inv, double L, double* __restrict__ Al,
                                                   double* __restrict__ Bt, double kap,
                                                   const double* __restrict__ E) {
    constexpr int R = 16;
    int line = blockIdx.x * 256 + threadIdx.x;
    int nlines = n * n;
    if (line >= nlines) return;
    int z = line % n, o = line / n;
    int64_t base, stride;
    if (axis == 2) { base = z + (int64_t)n * o; stride = (int64_t)n * n; }   // (z,x=o), along y
    else { base = z + (int64_t)n * n * o; stride = n; }                        // (z,y=o), along x
    const int N = n - 1;
    // al1 with Al[1]=1 (lateral, :4021), bt1 with Bt[1]=bc0()=0 (:4511)
    double al = 1.0, bt = 0.0;
    Al[(int64_t)1 * nlines + line] = al;
    Bt[(int64_t)1 * nlines + line] = bt;
    double Kprev = K[base];
    for (int i0 = 1; i0 < N; i0 += R) {
        double kk[R], ic[R], uu[R], hh[R], ee[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            int i = i0 + j;
            bool ok = i < N;
            int64_t id = base + stride * (ok ? i : 0);
            kk[j] = ok ? K[id] : 0.0;
            ic[j] = ok ? iC[id] : 0.0;
            uu[j] = ok ? U[id] : 0.0;
            hh[j] = (ok && H) ? H[id] : 0.0;
            ee[j] = (ok && E) ? E[id] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < R; j++) {
            int i = i0 + j;
            if (i < N) {
                double Km = Kprev, Kp = kk[j];
                double A_ = Km * ic[j] * h2inv * kap;
                double B_ = Kp * ic[j] * h2inv * kap;
                double R_ = ic[j] * ((Kp * h2inv) + (Km * h2inv)) * kap + L;
                double Om = -L * uu[j] + hh[j];
                if (E) Om -= ic[j] * ee[j];
                double aln = B_ / (R_ - A_ * al);
                bt = (aln / B_) * (A_ * bt - Om);
                al = aln;
                Al[(int64_t)(i + 1) * nlines + line] = al;
                Bt[(int64_t)(i + 1) * nlines + line] = bt;
                Kprev = Kp;
            }
        }
    }
    // zero-flux end, U1 :4704-4720 with bcN()=0
    double uN;
    if (fabs(bt) > 1e-15) uN = (fabs(1.0 - al) > 1e-15) ? bt / (1.0 - al) : U[base + stride * (N - 1)];
    else uN = (fabs(1.0 - al) > 1e-15) ? 0.0 : U[base + stride * (N - 1)];
    U[base + stride * N] = uN;
    double u = uN;
    for (int i1 = N - 1; i1 >= 0; i1 -= R) {
        double aa[R], bb[R];
#pragma unroll
        for (int j = 0; j < R; j++) {
            int i = i1 - j;
            bool ok = i >= 0;
            aa[j] = ok ? Al[(int64_t)(i + 1) * nlines + line] : 0.0;
            bb[j] = ok ? Bt[(int64_t)(i + 1) * nlines + line] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < R; j++) {
            int i = i1 - j;
            if (i >= 0) {
                u = aa[j] * u + bb[j];
                U[base + stride * i] = u;
            }
        }
    }
}


// Depth sweep: block = one z line, one node per thread (blockDim = n rounded
// up to 64, n <= 256).  PCR on rows normalised to b = 1 with ping-pong LDS
// buffers (one barrier and one reciprocal per level).  Writes U^{n+1}, the
// history level D[t] = U^{n+1} - Uprev and refreshes Uprev.
__global__ __launch_bounds__(256) void k_zline3d(double* __restrict__ U, double* __restrict__ Uprev,
                                                 const double* __restrict__ utop, int n, double h,
                                                 double h2inv, double ksat, int mode, double L,
                                                 double* __restrict__ Dlev, double kap,
                                                 const double* __restrict__ E,
                                                 const double* __restrict__ H) {
    __shared__ double a[2][256], c[2][256], d[2][256], sK[256];
    const int tid = threadIdx.x;
    const int line = blockIdx.x;                   // x + n*y
    const int64_t base = (int64_t)n * line;
    const bool act = tid < n;
    double u = 0.0, k = 0.0, ic = 0.0, uprev = 0.0;
    if (act) {
        u = U[base + tid];
        uprev = Uprev[base + tid];
        vgm_coefs(u, tid * h, ksat, mode, k, ic);
        sK[tid] = k;
    }
    double top = utop[line];
    __syncthreads();
    if (act) {
        double ar, br, cr, dr;
        if (tid == 0) { ar = 0.0; br = 1.0; cr = 0.0; dr = top; }
        else if (tid == n - 1) { ar = -1.0; br = 1.0; cr = 0.0; dr = 0.0; }
        else {
            double Km = sK[tid - 1], Kp = k;
            ar = Km * ic * h2inv * kap;
            cr = Kp * ic * h2inv * kap;
            br = -(ic * ((Kp * h2inv) + (Km * h2inv)) * kap + L);
            dr = -L * u + (H ? H[base + tid] : 0.0);
            if (E) dr -= ic * E[base + tid];
        }
        double ib = 1.0 / br;
        a[0][tid] = ar * ib; c[0][tid] = cr * ib; d[0][tid] = dr * ib;
    }
    __syncthreads();
    int p = 0;
    for (int delta = 1; delta < n; delta <<= 1) {
        if (act) {
            double aq = a[p][tid], cq = c[p][tid], dq = d[p][tid];
            double am = 0, cm = 0, dm = 0, ap = 0, cp = 0, dp = 0;
            if (tid - delta >= 0) { am = a[p][tid - delta]; cm = c[p][tid - delta]; dm = d[p][tid - delta]; }
            if (tid + delta < n) { ap = a[p][tid + delta]; cp = c[p][tid + delta]; dp = d[p][tid + delta]; }
            double inv = 1.0 / (1.0 - aq * cm - cq * ap);
            a[1 - p][tid] = -aq * am * inv;
            c[1 - p][tid] = -cq * cp * inv;
            d[1 - p][tid] = (dq - aq * dm - cq * dp) * inv;
        }
        __syncthreads();
        p = 1 - p;
    }
    if (act) {
        double un = d[p][tid];
        int64_t id = base + tid;
        U[id] = un;
        Dlev[id] = un - uprev;
        Uprev[id] = un;
    }
}

// Lateral (y or x) sweep as a partitioned solve: each line's n nodes are cut
// into kLatSeg segments of at most S rows, one thread each (32 lines x 8
// segments per workgroup; the 32 lines are consecutive z, so every row load
// is coalesced).  The line's system is the reference one: zero flux at both
// ends (U_0 = U_1: Al_1 = 1, Bt_1 = 0, fpde_hpc_1d.cpp:4021, :4511; U_N =
// U_{N-1}: the bottom of U1, :4704-4720) and the rows of A_/B/R/Om between.
//  1. A segment [lo, hi] moves its couplings to U_{lo-1} and U_{hi+1} to two
//     right-hand sides (a_lo e_1, c_hi e_last) and runs the Thomas
//     recurrences for the three: U_i = x_i + y_i U_{lo-1} + z_i U_{hi+1}.
//  2. One barrier publishes the segments' end rows; every thread runs the
//     recursion over the segments (f_s = U_lo, l_s = U_hi: l_{s-1} = P + Q f_s
//     forward, f_s = F_s + G_s f_{s+1} back) and back-substitutes its rows.
// Same system as the one-thread-per-line sweeps (k_lateral3d*).  The
// reference's 1e-15 guards on the bottom row act on singular lines (|1 -
// Al_N| <= 1e-15, which L > 0 rules out: every row is diagonally dominant)
// and on a bottom value within 1e-15 of zero, where the two differ by less
// than that absolute amount.
// K, iC: the coefficients at U from a coefficient pass, or nullptr: then the
// sweep evaluates them itself, at the rows it solves and the node before its
// first row (every node of a lateral line has the same depth z0 h), the same
// values the pass would store.
constexpr int kLatSeg = 8;
static_assert(16 * kLatSeg == 128, "lateral3d_partitioned: 8 segments of at most 16 rows");
template <int S>
__device__ __forceinline__ void lateral3d_seg_body(double* __restrict__ U, const double* __restrict__ K,
                                                   const double* __restrict__ iC, const double* __restrict__ H,
                                                   int n, int axis, double h2inv, double L, double kap,
                                                   const double* __restrict__ E, double h, double ksat, int mode,
                                                   int bid) {
    __shared__ double sEnd[kLatSeg][32][7];    // x, y, z of the first and the last row
    const int tid = threadIdx.x;
    const int l = tid & 31, sg = tid >> 5;
    const int line = bid * 32 + l;
    const bool ok_line = line < n * n;
    const int ln = ok_line ? line : 0;
    const int z0 = ln % n, o = ln / n;
    int64_t base, stride;
    if (axis == 2) { base = z0 + (int64_t)n * o; stride = (int64_t)n * n; }   // (z,x=o), along y
    else { base = z0 + (int64_t)n * n * o; stride = n; }                        // (z,y=o), along x
    const int N = n - 1;
    const int Sg = (n + kLatSeg - 1) / kLatSeg;      // rows per segment (<= S)
    const int W = (n + Sg - 1) / Sg;                 // segments present
    const int lo = sg * Sg, hi = min(n, lo + Sg) - 1;
    const bool act = ok_line && lo <= hi;
    // 1. Thomas on the segment: cp = c', x/y the modified right-hand sides
    // (straight-line code: rows past hi and the two end rows by selects)
    double cp[S], xs[S], ys[S];
    double zl = 0.0;                                 // z of the last row
    double cprev = 0.0, xprev = 0.0, yprev = 0.0;
    const double zd = z0 * h;
    double Km = 0.0;
    if (act && lo >= 1) {
        if (K) {
            Km = K[base + stride * (lo - 1)];
        } else {
            double icm;
            vgm_coefs(U[base + stride * (lo - 1)], zd, ksat, mode, Km, icm);
        }
    }
    // the rows' coefficients first (independent evaluations that the
    // scheduler interleaves), then the recurrence
    double Kr[S], ir[S], ur[S];
#pragma unroll
    for (int j = 0; j < S; j++) {
        const int i = lo + j;
        const bool inner = act && i <= hi && i >= 1 && i <= N - 1;
        const int64_t id = base + stride * (inner ? i : 0);
        ur[j] = inner ? U[id] : 0.0;
        if (K) {
            Kr[j] = inner ? K[id] : 0.0;
            ir[j] = inner ? iC[id] : 0.0;
        } else {
            double kv = 0.0, iv = 0.0;
            if (inner) vgm_coefs(ur[j], zd, ksat, mode, kv, iv);
            Kr[j] = kv;
            ir[j] = iv;
        }
    }
#pragma unroll
    for (int j = 0; j < S; j++) {
        const int i = lo + j;
        const bool valid = act && i <= hi;
        const bool inner = valid && i >= 1 && i <= N - 1;
        const int64_t id = base + stride * (inner ? i : 0);
        const double Kp = Kr[j];
        const double icv = ir[j];
        const double uv = ur[j];
        const double hv = (inner && H) ? H[id] : 0.0;
        const double ev = (inner && E) ? E[id] : 0.0;
        const double a = inner ? Km * icv * h2inv * kap : (i == N ? -1.0 : 0.0);
        const double c = inner ? Kp * icv * h2inv * kap : (i == 0 ? -1.0 : 0.0);
        const double b = inner ? -(icv * ((Kp * h2inv) + (Km * h2inv)) * kap + L) : 1.0;
        const double d = inner ? -L * uv + hv - icv * ev : 0.0;
        Km = Kp;
        const double inv = rcp_nr(j == 0 ? b : b - a * cprev);
        cp[j] = (valid && i < hi) ? c * inv : 0.0;
        xs[j] = valid ? (j == 0 ? d : d - a * xprev) * inv : 0.0;
        ys[j] = valid ? (j == 0 ? -a : -a * yprev) * inv : 0.0;
        if (valid && i == hi) zl = -c * inv;
        cprev = cp[j]; xprev = xs[j]; yprev = ys[j];
    }
    // back: x, y in place, z_i = -c'_i z_{i+1} into cp (rows past hi stay 0)
    double xn = 0.0, yn = 0.0, zn = 0.0;
#pragma unroll
    for (int j = S - 1; j >= 0; j--) {
        const int i = lo + j;
        const bool valid = act && i <= hi;
        const bool last = valid && i == hi;
        xn = last ? xs[j] : xs[j] - cp[j] * xn;
        yn = last ? ys[j] : ys[j] - cp[j] * yn;
        zn = last ? zl : -cp[j] * zn;
        xs[j] = xn; ys[j] = yn; cp[j] = zn;
        if (valid && j == 0) { sEnd[sg][l][0] = xn; sEnd[sg][l][1] = yn; sEnd[sg][l][2] = zn; }
        if (last) { sEnd[sg][l][3] = xn; sEnd[sg][l][4] = yn; sEnd[sg][l][5] = zn; }
    }
    __syncthreads();
    if (!act) return;
    // 2. the segments' end values: forward l_{s-1} = P + Q f_s, back f_s = F_s + G_s f_{s+1}
    double Fv[kLatSeg], Gv[kLatSeg], Pv[kLatSeg], Qv[kLatSeg];
    double pv = 0.0, qv = 0.0;
#pragma unroll
    for (int k = 0; k < kLatSeg; k++) {
        if (k < W) {
            const double Xf = sEnd[k][l][0], Yf = sEnd[k][l][1], Zf = sEnd[k][l][2];
            const double Xl = sEnd[k][l][3], Yl = sEnd[k][l][4], Zl = sEnd[k][l][5];
            const double inv = rcp_nr(1.0 - Yf * qv);
            Fv[k] = (Xf + Yf * pv) * inv;
            Gv[k] = Zf * inv;
            const double yq = Yl * qv;
            pv = Xl + Yl * pv + yq * Fv[k];
            qv = yq * Gv[k] + Zl;
            Pv[k] = pv; Qv[k] = qv;
        }
    }
    double fr = 0.0, vr = 0.0, vl = 0.0;             // f_{k+1}; this segment's right and left values
#pragma unroll
    for (int k = kLatSeg - 1; k >= 0; k--) {
        if (k < W) {
            if (k == sg) vr = fr;
            if (k == sg - 1) vl = Pv[k] + Qv[k] * fr;    // l_{sg-1}
            fr = Fv[k] + Gv[k] * fr;
        }
    }
#pragma unroll
    for (int j = 0; j < S; j++) {
        const int i = lo + j;
        if (i <= hi) U[base + stride * i] = xs[j] + ys[j] * vl + cp[j] * vr;
    }
}
template <int S>
__global__ __launch_bounds__(256, 2) void k_lateral3d_seg(double* __restrict__ U, const double* __restrict__ K,
                                                      const double* __restrict__ iC,
                                                      const double* __restrict__ H, int n, int axis,
                                                      double h2inv, double L, double kap,
                                                      const double* __restrict__ E, double h, double ksat,
                                                      int mode) {
    lateral3d_seg_body<S>(U, K, iC, H, n, axis, h2inv, L, kap, E, h, ksat, mode, blockIdx.x);
}


// Depth (z) sweep as a partitioned solve over 16 contiguous lines per
// workgroup: the lines' U, their VGM coefficients at U and the right-hand
// sides are staged in LDS by coalesced loads, then each line's n nodes are
// cut into kZSeg segments of at most 8 rows (one thread each) and solved as
// in k_lateral3d_seg (Thomas per segment with the two coupling right-hand
// sides, one barrier, the recursion over the segments).  Rows: Dirichlet top
// U_0 = u_top, the rows of A_/B/R/Om, zero-flux bottom U_{n-1} = U_{n-2};
// the new U, D[n] = U - Uprev and Uprev go out through LDS, coalesced.
constexpr int kZSeg = 16, kZLines = 16, kZThreads = kZSeg * kZLines;
static_assert(kZLines == kZLinesHost, "the host's block count of the depth sweep");
__device__ __forceinline__ void zline3d_seg_body(double* __restrict__ U, double* __restrict__ Uprev,
                                                 const double* __restrict__ utop, int n, double h, double h2inv,
                                                 double ksat, int mode, double L, double* __restrict__ Dlev,
                                                 double kap, const double* __restrict__ E,
                                                 const double* __restrict__ Kin, const double* __restrict__ iCin,
                                                 const double* __restrict__ H, int bid) {
    constexpr int LD = 129;                      // odd row stride: the 16 lines hit distinct banks
    __shared__ double sK[kZLines][LD], sIC[kZLines][LD], sD[kZLines][LD];
    __shared__ double sEnd[kZSeg][kZLines][7];
    const int tid = threadIdx.x;
    const int nlines = n * n;
    const int line0 = bid * kZLines;
    const int nl = min(kZLines, nlines - line0);        // lines of this group
    const int64_t base0 = (int64_t)n * line0;
    // stage: K, 1/C and the right-hand side -L U - E/C of every node (four
    // nodes in flight per thread: the VGM chains are latency bound); K and
    // 1/C come from memory when a coefficient pass already formed them at U
#pragma unroll 4
    for (int q = tid; q < nl * n; q += kZThreads) {
        const int l = q / n, i = q % n;
        const double u = U[base0 + q];
        double k, ic;
        if (Kin) { k = Kin[base0 + q]; ic = iCin[base0 + q]; }
        else vgm_coefs(u, i * h, ksat, mode, k, ic);
        sK[l][i] = k;
        sIC[l][i] = ic;
        sD[l][i] = (-L * u + (H ? H[base0 + q] : 0.0)) - (E ? ic * E[base0 + q] : 0.0);
    }
    __syncthreads();
    const int l = tid & (kZLines - 1), sg = tid / kZLines;
    const int N = n - 1;
    const int Sg = (n + kZSeg - 1) / kZSeg;             // rows per segment (<= 8)
    const int W = (n + Sg - 1) / Sg;
    const int lo = sg * Sg, hi = min(n, lo + Sg) - 1;
    const bool act = l < nl && lo <= hi;
    const double top = (act && lo == 0) ? utop[line0 + l] : 0.0;
    constexpr int S = 8;
    double cp[S], xs[S], ys[S];
    double zl = 0.0, cprev = 0.0, xprev = 0.0, yprev = 0.0;
#pragma unroll
    for (int j = 0; j < S; j++) {
        const int i = lo + j;
        const bool valid = act && i <= hi;
        const bool inner = valid && i >= 1 && i <= N - 1;
        const int ii = inner ? i : 1;
        const double Km = inner ? sK[l][ii - 1] : 0.0, Kp = inner ? sK[l][ii] : 0.0;
        const double icv = inner ? sIC[l][ii] : 0.0;
        const double a = inner ? Km * icv * h2inv * kap : (i == N ? -1.0 : 0.0);
        const double c = inner ? Kp * icv * h2inv * kap : 0.0;
        const double b = inner ? -(icv * ((Kp * h2inv) + (Km * h2inv)) * kap + L) : 1.0;
        const double d = inner ? sD[l][ii] : (i == 0 ? top : 0.0);
        const double inv = rcp_nr(j == 0 ? b : b - a * cprev);
        cp[j] = (valid && i < hi) ? c * inv : 0.0;
        xs[j] = valid ? (j == 0 ? d : d - a * xprev) * inv : 0.0;
        ys[j] = valid ? (j == 0 ? -a : -a * yprev) * inv : 0.0;
        if (valid && i == hi) zl = -c * inv;
        cprev = cp[j]; xprev = xs[j]; yprev = ys[j];
    }
    double xn = 0.0, yn = 0.0, zn = 0.0;
#pragma unroll
    for (int j = S - 1; j >= 0; j--) {
        const int i = lo + j;
        const bool valid = act && i <= hi;
        const bool last = valid && i == hi;
        xn = last ? xs[j] : xs[j] - cp[j] * xn;
        yn = last ? ys[j] : ys[j] - cp[j] * yn;
        zn = last ? zl : -cp[j] * zn;
        xs[j] = xn; ys[j] = yn; cp[j] = zn;
        if (valid && j == 0) { sEnd[sg][l][0] = xn; sEnd[sg][l][1] = yn; sEnd[sg][l][2] = zn; }
        if (last) { sEnd[sg][l][3] = xn; sEnd[sg][l][4] = yn; sEnd[sg][l][5] = zn; }
    }
    __syncthreads();
    if (act) {
        double Fv[kZSeg], Gv[kZSeg], Pv[kZSeg], Qv[kZSeg];
        double pv = 0.0, qv = 0.0;
#pragma unroll
        for (int k = 0; k < kZSeg; k++) {
            if (k < W) {
                const double Xf = sEnd[k][l][0], Yf = sEnd[k][l][1], Zf = sEnd[k][l][2];
                const double Xl = sEnd[k][l][3], Yl = sEnd[k][l][4], Zl = sEnd[k][l][5];
                const double inv = rcp_nr(1.0 - Yf * qv);
                Fv[k] = (Xf + Yf * pv) * inv;
                Gv[k] = Zf * inv;
                const double yq = Yl * qv;
                pv = Xl + Yl * pv + yq * Fv[k];
                qv = yq * Gv[k] + Zl;
                Pv[k] = pv; Qv[k] = qv;
            }
        }
        double fr = 0.0, vr = 0.0, vl = 0.0;
#pragma unroll
        for (int k = kZSeg - 1; k >= 0; k--) {
            if (k < W) {
                if (k == sg) vr = fr;
                if (k == sg - 1) vl = Pv[k] + Qv[k] * fr;
                fr = Fv[k] + Gv[k] * fr;
            }
        }
        // this thread's rows of sD are read by no other thread: overwrite
        // them with the solution
#pragma unroll
        for (int j = 0; j < S; j++) {
            const int i = lo + j;
            if (i <= hi) sD[l][i] = xs[j] + ys[j] * vl + cp[j] * vr;
        }
    }
    __syncthreads();
    for (int q = tid; q < nl * n; q += kZThreads) {
        const int ll = q / n, i = q % n;
        const double un = sD[ll][i];
        const int64_t id = base0 + q;
        const double up = Uprev[id];
        U[id] = un;
        Dlev[id] = un - up;
        Uprev[id] = un;
    }
}
__global__ __launch_bounds__(kZThreads, 2) void k_zline3d_seg(double* __restrict__ U, double* __restrict__ Uprev,
                                                       const double* __restrict__ utop, int n, double h,
                                                       double h2inv, double ksat, int mode, double L,
                                                       double* __restrict__ Dlev, double kap,
                                                       const double* __restrict__ E,
                                                       const double* __restrict__ Kin,
                                                       const double* __restrict__ iCin,
                                                       const double* __restrict__ H) {
    zline3d_seg_body(U, Uprev, utop, n, h, h2inv, ksat, mode, L, Dlev, kap, E, Kin, iCin, H, blockIdx.x);
}


// S[i] = sum of far partials + near levels [near_lo, n) + two-level corrections,
// times hscale (1, or 1/3 when every sweep carries a third of the term).
// omega is the weight row of the block's first step n0 (rows omega_stride
// apart).  With mid, the 16-level groups [near_lo + 16 j, +16) of the near
// field are folded into mid for the steps [s, n0 + blk) at the step s that
// completes them (the block's first step folds those already complete), and
// step n takes the complete groups as one mid value plus the levels after
// them.
__device__ __forceinline__ void hist_combine_body(double* __restrict__ S, const double* __restrict__ part,
                                                  int nsplit, int64_t part_stride, int64_t part_off,
                                                  const double* __restrict__ D, int64_t ld, int64_t nodes, int n,
                                                  int near_lo, const double* __restrict__ omega_blk,
                                                  int64_t omega_stride, int n0, int blk, double* __restrict__ mid,
                                                  const double* __restrict__ corr, double hscale, int bid) {
    const double* __restrict__ omega = omega_blk + (int64_t)(n - n0) * omega_stride;
    int64_t i = (int64_t)bid * 256 + threadIdx.x;
    if (i >= nodes) return;
    constexpr int W = 16;      // independent loads per batch
    double far = 0.0;
    for (int sp0 = 0; sp0 < nsplit; sp0 += W) {
        double v[W];
#pragma unroll
        for (int j = 0; j < W; j++) v[j] = (sp0 + j < nsplit) ? part[(sp0 + j) * part_stride + part_off + i] : 0.0;
#pragma unroll
        for (int j = 0; j < W; j++) far += v[j];
    }
    double nr0 = 0.0, nr1 = 0.0;
    int lo = near_lo;
    if (mid) {
        const int done = n > near_lo ? (n - near_lo) / kNearGroup : 0;       // groups complete by now
        const int before = (n > n0 && n - 1 > near_lo) ? (n - 1 - near_lo) / kNearGroup : 0;
        for (int j = before; j < done; j++)
            fold_near_group(D, ld, omega_blk, omega_stride, n0, mid, i, near_lo + kNearGroup * j, n, n0 + blk,
                            j == 0);
        if (done > 0) {
            nr0 = mid[(int64_t)(n - n0) * ld + i];
            lo = near_lo + kNearGroup * done;
        }
    }
    for (int t0 = lo; t0 < n; t0 += W) {
        double v[W], w[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            int t = t0 + j;
            bool ok = t < n;
            v[j] = ok ? D[(int64_t)t * ld + i] : 0.0;
            w[j] = ok ? omega[n - t - 1] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < W; j += 2) {
            nr0 = fma(w[j], v[j], nr0);
            nr1 = fma(w[j + 1], v[j + 1], nr1);
        }
    }
    double cr = 0.0;
    if (n >= 1) cr = corr[2 * n] * D[i];
    if (n >= 2) cr = fma(corr[2 * n + 1], D[ld + i], cr);
    S[i] = (far + (nr0 + nr1) + cr) * hscale;
}
__global__ __launch_bounds__(256) void k_hist_combine(double* __restrict__ S, const double* __restrict__ part,
                                                      int nsplit, int64_t part_stride, int64_t part_off,
                                                      const double* __restrict__ D, int64_t ld, int64_t nodes,
                                                      int n, int near_lo, const double* __restrict__ omega,
                                                      int64_t omega_stride, int n0, int blk,
                                                      double* __restrict__ mid,
                                                      const double* __restrict__ corr, double hscale) {
    hist_combine_body(S, part, nsplit, part_stride, part_off, D, ld, nodes, n, near_lo, omega, omega_stride, n0,
                      blk, mid, corr, hscale, blockIdx.x);
}

// Series-compressed history (sum_alg 2): add the newest interval dnew to the
// moments and sum the series, S[i] = sum_j v_j (H[j][i] += c_j dnew[i]);
// dnew == nullptr (step 0) gives S = 0.
__global__ __launch_bounds__(256) void k_series3d(double* __restrict__ S, double* __restrict__ H, int P,
                                                  const double* __restrict__ dnew,
                                                  const double* __restrict__ ser_c,
                                                  const double* __restrict__ ser_v, int64_t nodes, int64_t ld,
                                                  double hscale) {
    int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nodes) return;
    if (!dnew) { S[i] = 0.0; return; }
    S[i] = series_update(H, ld, i, P, dnew[i], ser_c, ser_v) * hscale;
}

// ---------------------------------------------------------------------------
// Batched step of several 3D solves: the same bodies, each workgroup mapped to
// (solve j, its local block) by the prefix counts of the table.
__device__ inline int batch_slot(const int* first, int count, int bid) {
    int j = 0;
    while (j + 1 < count && bid >= first[j + 1]) j++;
    return j;
}

__global__ __launch_bounds__(256) void k_hist_combine_batch(Lod3DBatch b) {
    const int j = batch_slot(b.first_node, b.count, blockIdx.x);
    const Lod3DRow& r = b.row[j];
    hist_combine_body(r.S, r.part, r.nsplit, r.part_stride, r.part_off, r.D, r.ld, r.nodes, r.step, r.near_lo,
                      r.omega, r.omega_stride, r.n0, r.blk, r.mid, r.corr, r.hscale,
                      (int)blockIdx.x - b.first_node[j]);
}

// y sweep (axis 2) carries the history term; the x sweep (axis 1) a third of
// it with split3d odd
__global__ __launch_bounds__(256, 2) void k_lateral3d_batch(Lod3DBatch b, int axis) {
    const int j = batch_slot(b.first_lat, b.count, blockIdx.x);
    const Lod3DRow& r = b.row[j];
    lateral3d_seg_body<16>(r.U, r.Kl, r.iCl, (axis == 2 || r.split) ? r.S : nullptr, r.n, axis, r.h2inv, r.L,
                           r.kap, r.E, r.h, r.ksat, r.coef_mode, (int)blockIdx.x - b.first_lat[j]);
}

__global__ __launch_bounds__(kZThreads, 2) void k_zline3d_batch(Lod3DBatch b) {
    const int j = batch_slot(b.first_z, b.count, blockIdx.x);
    const Lod3DRow& r = b.row[j];
    zline3d_seg_body(r.U, r.Uprev, r.utop, r.n, r.h, r.h2inv, r.ksat, r.coef_mode, r.L, r.Dlev, r.kap, r.E,
                     r.Kz, r.iCz, r.split ? r.S : nullptr, (int)blockIdx.x - b.first_z[j]);
}

fpde_status launch_lod3d_batch(const Lod3DBatch& b, hipStream_t st) {
    if (b.count < 1 || b.count > kMax3DBatch) { set_error("batched 3D step: 1 to 8 solves"); return FPDE_ERR_ARG; }
    for (int j = 0; j < b.count; j++)
        if (b.row[j].n > 16 * kLatSeg || !(b.row[j].L > 0.0)) {
            set_error("batched 3D step: partitioned sweeps only (n <= 128)");
            return FPDE_ERR_ARG;
        }
    const int gn = b.first_node[b.count], gl = b.first_lat[b.count], gz = b.first_z[b.count];
    // the lateral sweeps evaluate their coefficients themselves (rows' Kl, iCl null)
    hipLaunchKernelGGL(k_hist_combine_batch, dim3(gn), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_lateral3d_batch, dim3(gl), dim3(256), 0, st, b, 2);
    hipLaunchKernelGGL(k_lateral3d_batch, dim3(gl), dim3(256), 0, st, b, 1);
    hipLaunchKernelGGL(k_zline3d_batch, dim3(gz), dim3(kZThreads), 0, st, b);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_series3d(double* S, double* H, int P, const double* dnew, const double* ser_c,
                            const double* ser_v, int64_t nodes, int64_t ld, double hscale, hipStream_t st) {
    hipLaunchKernelGGL(k_series3d, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, st, S, H, P, dnew, ser_c,
                       ser_v, nodes, ld, hscale);
    return FPDE_OK;
}

fpde_status launch_coef3d(const double* U, int n, double h, double ksat, int mode, double* K,
                          double* iC, hipStream_t st) {
    int64_t nn = (int64_t)n * n * n;
    hipLaunchKernelGGL(k_coef3d, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st, U, n, h, ksat, mode, K, iC);
    return FPDE_OK;
}

fpde_status launch_lateral3d(double* U, const double* K, const double* iC, const double* H, int n,
                             int axis, double h2inv, double L, double* Al, double* Bt, double kap,
                             const double* E, double h, double ksat, int mode, hipStream_t st) {
    int nl = n * n;
    if (lateral3d_partitioned(n, L)) {   // 8 segments of <= 16 rows per line
        hipLaunchKernelGGL(k_lateral3d_seg<16>, dim3((nl + 31) / 32), dim3(256), 0, st, U, K, iC, H, n, axis, h2inv,
                           L, kap, E, h, ksat, mode);
        return FPDE_OK;
    }
    if (!K || !iC) { set_error("3D lateral sweep over 128 nodes needs the coefficient pass"); return FPDE_ERR_ARG; }
    hipLaunchKernelGGL(k_lateral3d, dim3((nl + 255) / 256), dim3(256), 0, st, U, K, iC, H, n, axis, h2inv, L, Al, Bt,
                       kap, E);
    return FPDE_OK;
}

fpde_status launch_zline3d(double* U, double* Uprev, const double* utop, int n, double h, double h2inv,
                           double ksat, int mode, double L, double* Dlev, double kap, const double* E,
                           const double* K, const double* iC, const double* H, hipStream_t st) {
    if (n > 256) { set_error("3D grid: at most 256 nodes per axis"); return FPDE_ERR_ARG; }
    if (n <= 8 * kZSeg && L > 0.0) {        // partitioned: 16 lines per workgroup, 16 segments of <= 8 rows
        hipLaunchKernelGGL(k_zline3d_seg, dim3((n * n + kZLines - 1) / kZLines), dim3(kZThreads), 0, st, U, Uprev, utop, n,
                           h, h2inv, ksat, mode, L, Dlev, kap, E, K, iC, H);
        return FPDE_OK;
    }
    int bs = ((n + 63) / 64) * 64;
    hipLaunchKernelGGL(k_zline3d, dim3(n * n), dim3(bs), 0, st, U, Uprev, utop, n, h, h2inv, ksat, mode, L, Dlev,
                       kap, E, H);
    return FPDE_OK;
}

fpde_status launch_hist_combine(double* S, const double* part, int nsplit, int64_t part_stride,
                                int64_t part_off, const double* D, int64_t ld, int64_t nodes, int n,
                                int near_lo, const double* omega, int64_t omega_stride, int n0, int blk,
                                double* mid, const double* corr, double hscale, hipStream_t st) {
    hipLaunchKernelGGL(k_hist_combine, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, st, S, part,
                       nsplit, part_stride, part_off, D, ld, nodes, n, near_lo, omega, omega_stride, n0, blk, mid,
                       corr, hscale);
    return FPDE_OK;
}

}  // namespace fpde
