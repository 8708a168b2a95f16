// One launch of the 1D implicit FP64 step (gfx950).
//
// The reference step (fpde_hpc_1d.cpp:4672-4725: al1, Om1, bt1, back
// substitution) is a Thomas sweep over the whole line; on the GPU a single
// sequential sweep of 1e4 rows is latency bound, so the line is cut into C
// chunks of <= 256 nodes, one workgroup each (partitioned / SPIKE-style
// solve):
//  - inside a chunk the interior rows are solved by parallel cyclic reduction
//    in LDS for three right-hand sides, giving U_i = x_i + y_i U_s + z_i U_e
//    in terms of the chunk's end nodes s, e;
//  - the end-node rows then form a tridiagonal "reduced" system of 2C rows,
//    written to global memory;
//  - the NEXT launch solves the reduced system redundantly in every
//    workgroup (PCR in LDS) and back-substitutes its chunk.
// So one launch = [finish the previous solve] + [history term, coefficients,
// rows, chunk elimination of this solve]: one launch per Picard iteration and
// no grid-wide barrier.  Rows (fpde_hpc_1d.cpp:3095-3142 with Dg=1, gamma=1):
//   A_i U_{i-1} - R_i U_i + B_i U_{i+1} = Om_i,
//   A_i = K_{i-1}/C_i/h^2, B_i = K_i/C_i/h^2, R_i = (K_i+K_{i-1})/C_i/h^2 + L,
//   Om_i = -L U^n_i + S_n[i]      (Om1, :4165-4268; S_n = Da*history sum)
// row 0: Dirichlet U_0 = u_top; row N-1: zero flux U_{N-1} = U_{N-2}
// (al1/bt1 with the bottom formula of U1, :4684-4703, bcN()=0).
#include "fpde_internal.h"

namespace fpde {

constexpr int kMaxRed = 512;   // reduced rows (2C) per launch: N <= 65536

// Parallel cyclic reduction on rows normalised to b = 1 (row r: a_r x_{r-1}
// + x_r + c_r x_{r+1} = d_r), 256 threads, two rows per thread, ping-pong LDS
// buffers so each level costs one barrier and one reciprocal per row.
// Rows outside [0,n) are identity rows.  Returns the buffer holding x in D.
__device__ inline int pcr_norm(double (*A)[kMaxRed], double (*Cc)[kMaxRed], double (*Dd)[kMaxRed], int n) {
    const int tid = threadIdx.x;
    int p = 0;
    for (int delta = 1; delta < n; delta <<= 1) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = tid + 256 * j;
            if (r < n) {
                double ar = A[p][r], cr = Cc[p][r], dr = Dd[p][r];
                double am = 0.0, cm = 0.0, dm = 0.0, ap = 0.0, cp = 0.0, dp = 0.0;
                if (r - delta >= 0) { am = A[p][r - delta]; cm = Cc[p][r - delta]; dm = Dd[p][r - delta]; }
                if (r + delta < n) { ap = A[p][r + delta]; cp = Cc[p][r + delta]; dp = Dd[p][r + delta]; }
                double inv = 1.0 / (1.0 - ar * cm - cr * ap);
                A[1 - p][r] = -ar * am * inv;
                Cc[1 - p][r] = -cr * cp * inv;
                Dd[1 - p][r] = (dr - ar * dm - cr * dp) * inv;
            }
        }
        __syncthreads();
        p = 1 - p;
    }
    return p;
}

// Far partial sums plus the near levels t in [near_lo, n-2] already in
// memory, gathered in 32-wide batches of independent loads (a sum over m
// terms costs ~m/32 memory round trips).  Every load is guarded: omega is
// read only at n-t-1 >= 1.
__device__ inline double hist_mem(const Step1DArgs& g, int i, int n) {
    constexpr int W = 32;
    double far = 0.0;
    for (int sp0 = 0; sp0 < g.nsplit; sp0 += W) {
        double v[W];
#pragma unroll
        for (int j = 0; j < W; j++)
            v[j] = (sp0 + j < g.nsplit) ? g.part[(int64_t)(sp0 + j) * g.part_stride + g.part_off + i] : 0.0;
#pragma unroll
        for (int j = 0; j < W; j++) far += v[j];
    }
    double nr0 = 0.0, nr1 = 0.0;
    const int t_end = n - 1;   // exclusive: D[n-1] is the newest level, added by the caller
    for (int t0 = g.near_lo; t0 < t_end; t0 += W) {
        double v[W], w[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            int t = t0 + j;
            bool ok = t < t_end;
            v[j] = ok ? g.D[(int64_t)t * g.ld + i] : 0.0;
            w[j] = ok ? g.omega[n - t - 1] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < W; j += 2) {
            nr0 = fma(w[j], v[j], nr0);
            nr1 = fma(w[j + 1], v[j + 1], nr1);
        }
    }
    return far + (nr0 + nr1);
}

// max over the 256 threads of a block (wave shuffles + LDS across the 4 waves)
__device__ inline double block_max(double v, double* sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    v = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    __syncthreads();
    return v;
}

__global__ __launch_bounds__(256) void k_step1d(Step1DArgs g) {
    __shared__ double rA[2][kMaxRed], rC[2][kMaxRed], rD[2][kMaxRed];
    __shared__ double pA[2][256], pC[2][256], pX[2][256], pY[2][256], pZ[2][256];
    __shared__ double sK[258];
    __shared__ double sred[4];
    const int tid = threadIdx.x;
    const int c = blockIdx.x;
    const int N = g.N, C = g.C;
    const int s = (int)(((int64_t)c * N) / C);
    const int e = (int)(((int64_t)(c + 1) * N) / C) - 1;
    const int nc = e - s + 1;
    const int i = s + tid;
    const bool act = tid < nc;
    const int R2 = 2 * C;
    const int n = g.n;
    const bool first = g.k <= 0;          // first Picard iteration of step n (or finalize)

    // ---- 0. loads that do not depend on the pending solve (one round trip) ---
    double rrow[2][4];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int r = tid + 256 * j;
        bool ok = g.recover && r < R2;
        const double* red = g.red + (size_t)(1 - g.parity) * R2 * 4 + 4 * r;
#pragma unroll
        for (int q = 0; q < 4; q++) rrow[j][q] = ok ? red[q] : 0.0;
    }
    const bool inner_node = act && i != s && i != e;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (g.recover && inner_node) { fx = g.xyz[i]; fy = g.xyz[g.ld + i]; fz = g.xyz[2 * g.ld + i]; }
    double uold = act ? g.U[i] : 0.0;                    // U^{n-1} (recovering, k=0) or U^n
    double Sk = (!first && act) ? g.Sbuf[i] : 0.0;
    const bool series = g.ser_P > 0;
    const int64_t drow = series ? ((n - 1) & 1) : (n - 1);    // row of D[n-1]
    double dmem = (first && !g.recover && act && n >= 1) ? g.D[drow * g.ld + i] : 0.0;
    double d0m = (first && !series && act && n >= 2) ? g.D[i] : 0.0;
    double d1m = (first && !series && act && n >= 3) ? g.D[g.ld + i] : 0.0;
    // Picard convergence of the previous launch (decided one launch late: the
    // test needs the max over every chunk): dmax of the previous launch and the
    // iterate it stored
    const bool check_conv = g.recover && g.prev_k >= 1;
    double dprev = (check_conv && tid < C) ? g.dmax[(size_t)(1 - g.parity) * C + tid] : 0.0;
    double uitp = ((check_conv || g.k >= 2) && act) ? g.Uit[i] : 0.0;
    double hm = (first && !series && g.k == 0 && act && n >= 1) ? hist_mem(g, i, n) : 0.0;

    // ---- 1. finish the previous solve: reduced system, then back-substitute --
    double val = 0.0;
    int pr = 0;
    if (g.recover) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = tid + 256 * j;
            if (r < R2) {
                double ib = 1.0 / rrow[j][1];
                rA[0][r] = rrow[j][0] * ib;
                rC[0][r] = rrow[j][2] * ib;
                rD[0][r] = rrow[j][3] * ib;
            }
        }
        __syncthreads();
        pr = pcr_norm(rA, rC, rD, R2);
        if (act) {
            if (i == s) val = rD[pr][2 * c];
            else if (i == e) val = rD[pr][2 * c + 1];
            else val = fx + fy * rD[pr][2 * c] + fz * rD[pr][2 * c + 1];
        }
    }
    // the previous Picard iterate met the tolerance: it is the step's result
    bool conv = false;
    if (check_conv) conv = block_max(dprev, sred) <= g.tol;
    if (conv) val = uitp;

    // ---- 2. advance the state -----------------------------------------------
    double un = 0.0, uit = 0.0, S = 0.0;
    if (first) {
        if (act) {
            un = uold;
            double dnew = dmem;
            if (g.recover) {                     // val = U^n, uold = U^{n-1}
                dnew = val - uold;
                g.D[drow * g.ld + i] = dnew;
                g.U[i] = val;
                un = val;
            }
            if (g.k >= 0 && n >= 1 && series) {
                // moments += newest interval, then the series (sum_alg 2)
                S = series_update(g.H, g.ld, i, g.ser_P, dnew, g.ser_c, g.ser_v);
            } else if (g.k >= 0 && n >= 1) {
                // S_n = far + near (in memory) + newest level + two-level corrections
                double newest = (g.near_lo <= n - 1) ? g.omega[0] * dnew : 0.0;
                double cr = (n == 1) ? g.corr[2] * dnew : g.corr[2 * n] * d0m;
                if (n == 2) cr = fma(g.corr[2 * n + 1], dnew, cr);
                else if (n >= 3) cr = fma(g.corr[2 * n + 1], d1m, cr);
                S = hm + newest + cr;
            }
            if (g.k >= 0) g.Sbuf[i] = S;
        }
        uit = un;
    } else if (act) {
        un = uold;
        S = Sk;
        uit = val;
    }
    if (g.k < 0) return;                         // finalize only (uniform)

    if (!first) {
        double* dm = g.dmax + (size_t)g.parity * C;
        if (conv) {
            // converged: hand the iterate on unchanged through an identity
            // reduced system and x = U, y = z = 0 (dmax 0 keeps later launches
            // of this step passing it on)
            double* redw = g.red + (size_t)g.parity * R2 * 4;
            if (act && i != s && i != e) { g.xyz[i] = uit; g.xyz[g.ld + i] = 0.0; g.xyz[2 * g.ld + i] = 0.0; }
            if (tid == 0) {
                redw[4 * (2 * c) + 0] = 0.0; redw[4 * (2 * c) + 1] = 1.0;
                redw[4 * (2 * c) + 2] = 0.0; redw[4 * (2 * c) + 3] = uit;
                dm[c] = 0.0;
            }
            if (tid == nc - 1) {
                redw[4 * (2 * c + 1) + 0] = 0.0; redw[4 * (2 * c + 1) + 1] = 1.0;
                redw[4 * (2 * c + 1) + 2] = 0.0; redw[4 * (2 * c + 1) + 3] = uit;
            }
            return;                              // uniform: conv is block-wide
        }
        // |U^{k} - U^{k-1}| over the chunk; U^0 of the step is U^n
        double prev_it = (g.k == 1) ? un : uitp;
        double d = act ? fabs(uit - prev_it) : 0.0;
        d = block_max(d, sred);
        if (tid == 0) dm[c] = d;
        if (act) g.Uit[i] = uit;
    }

    // ---- 3. coefficients at the iterate ------------------------------------
    double K = 0.0, invC = 0.0;
    // K of the node left of the chunk, on a thread without a node of its own
    // when there is one (else thread 0 does both evaluations)
    const int halo_tid = (nc < 256) ? 255 : 0;
    if (tid == halo_tid && c > 0) {
        double ulo = g.recover ? rD[pr][2 * c - 1] : g.U[s - 1];
        double Kl, iCl;
        vgm_coefs(ulo, (s - 1) * g.h, g.ksat, g.coef_mode, Kl, iCl);
        sK[0] = Kl;
    }
    if (act) {
        vgm_coefs(uit, i * g.h, g.ksat, g.coef_mode, K, invC);
        sK[tid + 1] = K;
    }
    __syncthreads();

    // ---- 4. rows -------------------------------------------------------------
    double ar = 0.0, br = 1.0, cr = 0.0, dr = 0.0;
    if (act) {
        if (i == 0) { ar = 0.0; br = 1.0; cr = 0.0; dr = g.utop; }
        else if (i == N - 1) { ar = -1.0; br = 1.0; cr = 0.0; dr = 0.0; }
        else {
            double Km = sK[tid], Kp = sK[tid + 1];
            double A = Km * invC * g.h2inv;
            double B = Kp * invC * g.h2inv;
            double R = invC * (Kp * g.h2inv + Km * g.h2inv) + g.L;
            ar = A; br = -R; cr = B; dr = -g.L * un + S;
        }
    }

    // ---- 5. chunk elimination: interior rows 1..nc-2, three right-hand sides
    // (x: the row's own RHS; y, z: unit values at the chunk ends s and e) ---
    const int ni = nc - 2;
    const int q = tid - 1;
    const bool inner = (tid >= 1) && (tid <= nc - 2);
    if (inner) {
        double ib = 1.0 / br;
        pA[0][q] = (tid == 1) ? 0.0 : ar * ib;
        pC[0][q] = (tid == nc - 2) ? 0.0 : cr * ib;
        pX[0][q] = dr * ib;
        pY[0][q] = (tid == 1) ? -ar * ib : 0.0;
        pZ[0][q] = (tid == nc - 2) ? -cr * ib : 0.0;
    }
    __syncthreads();
    int pp = 0;
    for (int delta = 1; delta < ni; delta <<= 1) {
        if (inner) {
            double aq = pA[pp][q], cq = pC[pp][q];
            double am = 0, cm = 0, xm = 0, ym = 0, zm = 0, ap = 0, cp = 0, xp = 0, yp = 0, zp = 0;
            if (q - delta >= 0) {
                am = pA[pp][q - delta]; cm = pC[pp][q - delta];
                xm = pX[pp][q - delta]; ym = pY[pp][q - delta]; zm = pZ[pp][q - delta];
            }
            if (q + delta < ni) {
                ap = pA[pp][q + delta]; cp = pC[pp][q + delta];
                xp = pX[pp][q + delta]; yp = pY[pp][q + delta]; zp = pZ[pp][q + delta];
            }
            double inv = 1.0 / (1.0 - aq * cm - cq * ap);
            pA[1 - pp][q] = -aq * am * inv;
            pC[1 - pp][q] = -cq * cp * inv;
            pX[1 - pp][q] = (pX[pp][q] - aq * xm - cq * xp) * inv;
            pY[1 - pp][q] = (pY[pp][q] - aq * ym - cq * yp) * inv;
            pZ[1 - pp][q] = (pZ[pp][q] - aq * zm - cq * zp) * inv;
        }
        __syncthreads();
        pp = 1 - pp;
    }
    if (inner) {
        g.xyz[i] = pX[pp][q];
        g.xyz[g.ld + i] = pY[pp][q];
        g.xyz[2 * g.ld + i] = pZ[pp][q];
    }

    // ---- 6. reduced rows of the chunk ends ----------------------------------
    double* redw = g.red + (size_t)g.parity * R2 * 4;
    if (tid == 0) {
        double x, y, z;
        if (nc == 2) { x = 0.0; y = 0.0; z = 1.0; }
        else { x = pX[pp][0]; y = pY[pp][0]; z = pZ[pp][0]; }
        redw[4 * (2 * c) + 0] = ar;
        redw[4 * (2 * c) + 1] = br + cr * y;
        redw[4 * (2 * c) + 2] = cr * z;
        redw[4 * (2 * c) + 3] = dr - cr * x;
    }
    if (tid == nc - 1) {
        double x, y, z;
        if (nc == 2) { x = 0.0; y = 1.0; z = 0.0; }
        else { x = pX[pp][ni - 1]; y = pY[pp][ni - 1]; z = pZ[pp][ni - 1]; }
        redw[4 * (2 * c + 1) + 0] = ar * y;
        redw[4 * (2 * c + 1) + 1] = br + ar * z;
        redw[4 * (2 * c + 1) + 2] = cr;
        redw[4 * (2 * c + 1) + 3] = dr - ar * x;
    }
}

fpde_status launch_step1d(const Step1DArgs& a, hipStream_t st) {
    if (2 * a.C > kMaxRed) { set_error("1D line too long for the reduced system (N <= 65536)"); return FPDE_ERR_ARG; }
    hipLaunchKernelGGL(k_step1d, dim3(a.C), dim3(256), 0, st, a);
    return FPDE_OK;
}

}  // namespace fpde
