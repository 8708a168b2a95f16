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
// + x_r + c_r x_{r+1} = d_r), 256 threads, two rows per thread kept in
// registers, ping-pong LDS buffers for the partners, so each level costs one
// barrier and one reciprocal per row.  Rows outside [0,n) are identity rows.
// Returns the buffer holding x in D.
__device__ inline int pcr_norm(double (*A)[kMaxRed], double (*Cc)[kMaxRed], double (*Dd)[kMaxRed], int n) {
    const int tid = threadIdx.x;
    double ar[2], cr[2], dr[2];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int r = tid + 256 * j;
        ar[j] = r < n ? A[0][r] : 0.0;
        cr[j] = r < n ? Cc[0][r] : 0.0;
        dr[j] = r < n ? Dd[0][r] : 0.0;
    }
    int p = 0;
    for (int delta = 1; delta < n; delta <<= 1) {
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = tid + 256 * j;
            if (r < n) {
                double am = 0.0, cm = 0.0, dm = 0.0, ap = 0.0, cp = 0.0, dp = 0.0;
                if (r - delta >= 0) { am = A[p][r - delta]; cm = Cc[p][r - delta]; dm = Dd[p][r - delta]; }
                if (r + delta < n) { ap = A[p][r + delta]; cp = Cc[p][r + delta]; dp = Dd[p][r + delta]; }
                double inv = rcp_nr(1.0 - ar[j] * cm - cr[j] * ap);
                dr[j] = (dr[j] - ar[j] * dm - cr[j] * dp) * inv;
                ar[j] = -ar[j] * am * inv;
                cr[j] = -cr[j] * cp * inv;
                A[1 - p][r] = ar[j];
                Cc[1 - p][r] = cr[j];
                Dd[1 - p][r] = dr[j];
            }
        }
        __syncthreads();
        p = 1 - p;
    }
    return p;
}

// The same reduction by one wave on its own LDS slice (A/Cc/Dd: [2][128]),
// for n <= 128 rows (rows r = lane, lane + 64).  Every wave of the block
// solves the reduced system redundantly, so no workgroup barrier is needed:
// LDS operations of one wave complete in order, and the wavefront fences
// keep the compiler from moving them across a level.
__device__ inline int pcr_wave(double* A, double* Cc, double* Dd, int n) {
    const int lane = threadIdx.x & 63;
    int p = 0;
    for (int delta = 1; delta < n; delta <<= 1) {
        double na[2], nc[2], nd[2];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = lane + 64 * j;
            na[j] = 0.0; nc[j] = 0.0; nd[j] = 0.0;
            if (r < n) {
                const double* a = A + p * 128;
                const double* cc = Cc + p * 128;
                const double* d = Dd + p * 128;
                double ar = a[r], cr = cc[r], dr = d[r];
                double am = 0.0, cm = 0.0, dm = 0.0, ap = 0.0, cp = 0.0, dp = 0.0;
                if (r - delta >= 0) { am = a[r - delta]; cm = cc[r - delta]; dm = d[r - delta]; }
                if (r + delta < n) { ap = a[r + delta]; cp = cc[r + delta]; dp = d[r + delta]; }
                double inv = rcp_nr(1.0 - ar * cm - cr * ap);
                na[j] = -ar * am * inv;
                nc[j] = -cr * cp * inv;
                nd[j] = (dr - ar * dm - cr * dp) * inv;
            }
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = lane + 64 * j;
            if (r < n) {
                A[(1 - p) * 128 + r] = na[j];
                Cc[(1 - p) * 128 + r] = nc[j];
                Dd[(1 - p) * 128 + r] = nd[j];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        p = 1 - p;
    }
    return p;
}

// Chunk elimination: the interior rows q in [0, ni) of a chunk, normalised
// to b = 1 (a_q x_{q-1} + x_q + c_q x_{q+1} = d_q) and held by thread q + 1
// in registers, for three right-hand sides (x: the row's own; y, z: unit
// values at the chunk ends), by parallel cyclic reduction: each thread keeps
// its own row in registers and reads its two partners of the level from the
// LDS ping-pong buffers.  Returns the buffer holding the solution of every
// row.
__device__ inline int chunk_solve(double (*pA)[256], double (*pC)[256], double (*pX)[256], double (*pY)[256],
                                  double (*pZ)[256], int ni, int q, bool inner, double aq, double cq, double xq,
                                  double yq, double zq) {
    if (inner) { pA[0][q] = aq; pC[0][q] = cq; pX[0][q] = xq; pY[0][q] = yq; pZ[0][q] = zq; }
    __syncthreads();
    int pp = 0;
    for (int delta = 1; delta < ni; delta <<= 1) {
        if (inner) {
            double am = 0, cm = 0, xm = 0, ym = 0, zm = 0, ap = 0, cp = 0, xp = 0, yp = 0, zp = 0;
            if (q - delta >= 0) {
                am = pA[pp][q - delta]; cm = pC[pp][q - delta];
                xm = pX[pp][q - delta]; ym = pY[pp][q - delta]; zm = pZ[pp][q - delta];
            }
            if (q + delta < ni) {
                ap = pA[pp][q + delta]; cp = pC[pp][q + delta];
                xp = pX[pp][q + delta]; yp = pY[pp][q + delta]; zp = pZ[pp][q + delta];
            }
            const double inv = rcp_nr(1.0 - aq * cm - cq * ap);
            const double na = -aq * am * inv, nc = -cq * cp * inv;
            xq = (xq - aq * xm - cq * xp) * inv;
            yq = (yq - aq * ym - cq * yp) * inv;
            zq = (zq - aq * zm - cq * zp) * inv;
            aq = na;
            cq = nc;
            pA[1 - pp][q] = aq; pC[1 - pp][q] = cq; pX[1 - pp][q] = xq; pY[1 - pp][q] = yq; pZ[1 - pp][q] = zq;
        }
        __syncthreads();
        pp = 1 - pp;
    }
    return pp;
}

// Far partial sums plus the near levels t in [near_lo, n-2] already in
// memory, gathered in 32-wide batches of independent loads (a sum over m
// terms costs ~m/32 memory round trips).  Every load is guarded: omega is
// read only at n-t-1 >= 1.  With `folded` > 0 the first folded * 16 near
// levels come as one value of g.mid (fold_near_group) and the loads start after
// them.
__device__ inline double hist_mem(const Step1DArgs& g, int i, int n, int64_t part_off,
                                  const double* __restrict__ omega, int folded = 0) {
    constexpr int W = 32;
    double far = 0.0;
    for (int sp0 = 0; sp0 < g.nsplit; sp0 += W) {
        double v[W];
#pragma unroll
        for (int j = 0; j < W; j++)
            v[j] = (sp0 + j < g.nsplit) ? g.part[(int64_t)(sp0 + j) * g.part_stride + part_off + i] : 0.0;
#pragma unroll
        for (int j = 0; j < W; j++) far += v[j];
    }
    double nr0 = folded > 0 ? g.mid[(int64_t)(n - g.n0) * g.ld + i] : 0.0, nr1 = 0.0;
    const int t_end = n - 1;   // exclusive: D[n-1] is the newest level, added by the caller
    for (int t0 = g.near_lo + kNearGroup * folded; t0 < t_end; t0 += W) {
        double v[W], w[W];
#pragma unroll
        for (int j = 0; j < W; j++) {
            int t = t0 + j;
            bool ok = t < t_end;
            v[j] = ok ? g.D[(int64_t)t * g.ld + i] : 0.0;
            w[j] = ok ? omega[n - t - 1] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < W; j += 2) {
            nr0 = fma(w[j], v[j], nr0);
            nr1 = fma(w[j + 1], v[j + 1], nr1);
        }
    }
    return far + (nr0 + nr1);
}

// The newest level and the two-level corrections of the oldest intervals,
// written with explicit fma so that every kernel rounds them alike:
// om0 dnew + corr[2n] d0 + corr[2n+1] d1, with d0 = dnew at n = 1 and
// d1 = dnew at n = 2.
// c0 = corr[2n], c1 = corr[2n+1].
__device__ inline double hist_tail(double om0, double c0, double c1, int n, int near_lo, double dnew, double d0m,
                                   double d1m) {
    const double w0 = (near_lo <= n - 1) ? om0 : 0.0;
    double cr = (n == 1) ? c0 * dnew : c0 * (n >= 2 ? d0m : 0.0);
    if (n == 2) cr = fma(c1, dnew, cr);
    else if (n >= 3) cr = fma(c1, d1m, cr);
    return fma(w0, dnew, cr);
}

// max over the 256 threads of a block (wave maxima + LDS across the 4 waves)
__device__ inline double block_max(double v, double* sh) {
    v = wave_max(v);
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

    // reduced systems of up to 128 rows are solved by every wave on its own
    // LDS slice (no workgroup barrier), larger ones by the whole block
    const bool wave_red = R2 <= 128;
    const int wave = tid >> 6, lane = tid & 63;

    // ---- 0. loads that do not depend on the pending solve (one round trip) ---
    double rrow[2][4];
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int r = wave_red ? lane + 64 * j : tid + 256 * j;
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
    // previous iterate: the initial one for iteration 1, else the stored one
    double uitp = 0.0;
    if ((check_conv || g.k >= 1) && act) uitp = (g.k == 1) ? g.Uit0[i] : g.Uit[i];
    // halo node s-1 (owned by chunk c-1, whose workgroup rewrites U and D in
    // this launch): the converged iterate and U^{n-1} come from buffers that
    // no workgroup writes in this launch
    const int halo_tid = (nc < 256) ? 255 : 0;
    const bool halo = tid == halo_tid && c > 0;
    const bool extrap0 = first && g.k == 0 && g.extrap && n >= 1;
    double uit_halo = (halo && check_conv) ? g.Uit[s - 1] : 0.0;
    double uprev_halo = (halo && extrap0) ? g.uend[(size_t)((n - 1) & 1) * R2 + 2 * c - 1] : 0.0;
    double hm = (first && !series && g.k == 0 && act && n >= 1) ? hist_mem(g, i, n, g.part_off, g.omega) : 0.0;

    // ---- 1. finish the previous solve: reduced system, then back-substitute --
    double val = 0.0;
    const double* xs = nullptr;          // solution of the reduced system
    if (g.recover) {
        if (wave_red) {
            double* A = &rA[0][0] + wave * 256;
            double* Cc = &rC[0][0] + wave * 256;
            double* Dd = &rD[0][0] + wave * 256;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                int r = lane + 64 * j;
                if (r < R2) {
                    double ib = rcp_nr(rrow[j][1]);
                    A[r] = rrow[j][0] * ib;
                    Cc[r] = rrow[j][2] * ib;
                    Dd[r] = rrow[j][3] * ib;
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            int pr = pcr_wave(A, Cc, Dd, R2);
            xs = Dd + pr * 128;
        } else {
#pragma unroll
            for (int j = 0; j < 2; j++) {
                int r = tid + 256 * j;
                if (r < R2) {
                    double ib = rcp_nr(rrow[j][1]);
                    rA[0][r] = rrow[j][0] * ib;
                    rC[0][r] = rrow[j][2] * ib;
                    rD[0][r] = rrow[j][3] * ib;
                }
            }
            __syncthreads();
            int pr = pcr_norm(rA, rC, rD, R2);
            xs = &rD[pr][0];
        }
        if (act) {
            if (i == s) val = xs[2 * c];
            else if (i == e) val = xs[2 * c + 1];
            else val = fx + fy * xs[2 * c] + fz * xs[2 * c + 1];
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
                S = hm + hist_tail(g.omega[0], g.corr[2 * n], g.corr[2 * n + 1], n, g.near_lo, dnew, d0m, d1m);
            }
            if (g.k >= 0) g.Sbuf[i] = S;
            // Picard starts from the extrapolation 2 U^n - U^{n-1}; the
            // iterate is kept for the convergence test of the next launch and
            // the chunk ends of U^n for the next step's halo
            uit = un;
            if (g.extrap && g.k == 0) {
                if (n >= 1) uit = un + dnew;
                g.Uit0[i] = uit;
                double* ue = g.uend + (size_t)(n & 1) * R2;
                if (i == s) ue[2 * c] = un;
                if (i == e) ue[2 * c + 1] = un;
            }
        }
    } else if (act) {
        un = uold;
        S = Sk;
        uit = val;
    }
    if (g.k < 0) {                               // finalize only (uniform)
        // space-fractional scheme: F of U^n for this chunk's rows (Om1
        // equation 0, fpde_hpc_1d.cpp:4176-4179), the arithmetic of
        // k_space_flux.  The neighbours s-1 and e+1 are chunk ends of the
        // reduced solution (or of the converged iterate), so the next step
        // needs no flux launch.
        if (g.sp_on && g.recover) {
            double* sU = &pA[0][0];              // U^n of nodes s-1 .. e+1
            double* sKf = &pC[0][0];             // K at U^n of nodes s-1 .. e
            if (act) {
                double k, icv;
                vgm_coefs(un, i * g.h, g.ksat, g.coef_mode, k, icv);
                sU[tid + 1] = un;
                sKf[tid + 1] = k;
            }
            if (tid == halo_tid) {
                if (c > 0) {
                    const double ul = conv ? uit_halo : xs[2 * c - 1];
                    double k, icv;
                    vgm_coefs(ul, (s - 1) * g.h, g.ksat, g.coef_mode, k, icv);
                    sU[0] = ul;
                    sKf[0] = k;
                }
                if (c < C - 1) sU[nc + 1] = conv ? g.Uit[e + 1] : xs[2 * c + 2];
            }
            __syncthreads();
            if (act) {
                double f = 0.0;
                if (i >= 1 && i <= N - 2) {
                    const double um = sU[tid], u0 = sU[tid + 1], up = sU[tid + 2];
                    const double Km = sKf[tid], Kj = sKf[tid + 1];
                    const double hh = g.h * g.h;
                    f = Kj * up / hh - (Kj / hh + Km / hh) * u0 + Km * um / hh;
                }
                g.sp_F[i] = f;
            }
        }
        return;
    }

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
        // |U^{k} - U^{k-1}| over the chunk; U^0 of the step is the initial
        // iterate stored by the k = 0 launch
        double d = act ? fabs(uit - uitp) : 0.0;
        d = block_max(d, sred);
        if (tid == 0) dm[c] = d;
        if (act) g.Uit[i] = uit;
    }

    // ---- 3. coefficients at the iterate ------------------------------------
    double K = 0.0, invC = 0.0;
    // K of the node left of the chunk, on a thread without a node of its own
    // when there is one (else thread 0 does both evaluations), at the same
    // iterate as the chunk's own nodes
    if (halo) {
        double ulo = g.recover ? (conv ? uit_halo : xs[2 * c - 1]) : g.U[s - 1];
        if (extrap0) ulo += ulo - uprev_halo;
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
    // (space-fractional scheme: the diffusion part of A_/B/R scaled by
    // Dg kb(gamma,i-1,i,i) and Om1's explicit sum -inv_dw_dh E_i, :4182-4189;
    // E of step n is formed by its first launch and reused by the others)
    double ar = 0.0, br = 1.0, cr = 0.0, dr = 0.0;
    if (act) {
        if (i == 0) { ar = 0.0; br = 1.0; cr = 0.0; dr = g.utop; }
        else if (i == N - 1) { ar = -1.0; br = 1.0; cr = 0.0; dr = 0.0; }
        else {
            const double kap = g.sp_on ? g.sp_kap : 1.0;
            double Km = sK[tid], Kp = sK[tid + 1];
            double A = Km * invC * g.h2inv * kap;
            double B = Kp * invC * g.h2inv * kap;
            double R = invC * (Kp * g.h2inv + Km * g.h2inv) * kap + g.L;
            ar = A; br = -R; cr = B; dr = -g.L * un + S;
            if (g.sp_on) {
                double E;
                if (first) {
                    E = space_sum(g.sp_P, g.sp_ldp, g.sp_corr, g.sp_F, i);
                    g.sp_E[i] = E;
                } else {
                    E = g.sp_E[i];
                }
                dr -= invC * E;
            }
        }
    }

    // ---- 5. chunk elimination: interior rows 1..nc-2, three right-hand sides
    // (x: the row's own RHS; y, z: unit values at the chunk ends s and e) ---
    const int ni = nc - 2;
    const int q = tid - 1;
    const bool inner = (tid >= 1) && (tid <= nc - 2);
    double ib = inner ? rcp_nr(br) : 0.0;
    const int pp = chunk_solve(pA, pC, pX, pY, pZ, ni, q, inner, (tid == 1) ? 0.0 : ar * ib,
                               (tid == nc - 2) ? 0.0 : cr * ib, dr * ib, (tid == 1) ? -ar * ib : 0.0,
                               (tid == nc - 2) ? -cr * ib : 0.0);
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

// ---------------------------------------------------------------------------
// Cooperative form: one launch runs the steps [n, n_end) of a history block
// with every Picard iteration inside, all C workgroups co-resident.  Per
// iteration each workgroup builds its rows at the iterate (VGM, chunk PCR)
// and publishes its two reduced rows with the convergence measure of the
// previous iterate; one grid barrier later every workgroup has the whole
// reduced system, decides convergence the reference way (stop once
// max|U^k - U^{k-1}| <= tol, the iterate U^k is the result) and otherwise
// solves the reduced system and back-substitutes its chunk.  A step's
// history needs only the workgroup's own nodes (its own earlier writes of D)
// and the precomputed far partials, so steps follow each other without a
// barrier; U and D of every step are written as in the per-launch form.

// A grid barrier waits at most `wait_ticks` ticks of the device's constant
// wall clock (wall_clock64; the host sizes it to 2 s from the clock rate).
// Workgroups of a cooperative launch can become resident late when other
// streams' kernels hold CUs at dispatch (the 3D and far-part kernels that
// advance beside a batch): those finish within micro- to milliseconds, far
// inside the bound, which exists only so that a launch whose workgroups can
// never all be resident ends with an error instead of hanging.

// Grid exchange of the chunk workgroups.  Every value a workgroup publishes
// for the others (its two reduced rows, its convergence measure) travels as
// two 64-bit words, each holding one 32-bit half of the double and the 32-bit
// tag of the barrier it belongs to (xtag | (barrier index + 1) mod 2^16; xtag
// carries the launch count in its upper 16 bits, never 0).  Words are written
// and read with relaxed agent-scope atomics: each is single-copy atomic and
// coherent across the XCDs, and a reader takes a word once it carries the
// barrier's tag, so it holds that barrier's value.  No fence orders one word
// against another, so a barrier needs no L2 write-back or invalidate (the
// release/acquire pair of a flag barrier costs more than the exchange).
// Readers poll the measure words first (thread j < C polls chunk j's: few
// pollers, little traffic while the last chunks are still computing), then
// load the rows, which have landed or are about to; polling every word of
// every row from the start was measured slower (its traffic delays the
// writers' stores).
//  - Words alternate by barrier parity.  A workgroup writes barrier b+1's
//    words only after it has read every workgroup's barrier-b words, and each
//    workgroup wrote those after reading its barrier-(b-1) words, so no word
//    is overwritten while a reader still needs it.
//  - At launch start each workgroup zeroes its own words (no tag is 0), so no
//    word an earlier launch left can carry one of this launch's tags.
//  - A reader waits at most `wait_ticks` ticks of the device's constant wall
//    clock per wait (the host sizes it to 2 s): workgroups of a cooperative
//    launch can become resident late when other streams' kernels hold CUs at
//    dispatch, which ends within micro- to milliseconds; the bound exists so
//    that a launch whose workgroups can never all be resident ends with an
//    error instead of hanging.
// Slots (two words each) per parity: value q of reduced row r at 4 r + q
// (r < 2C), the measure of chunk j at kXchMeasure + j.
constexpr int kXchMeasure = 8 * kMaxChunks;
constexpr int kXchWords = 2 * (kXchMeasure + kMaxChunks);       // per parity
static_assert(2 * kXchWords * 8 <= (int)kXchBytes, "exchange buffer size");

__device__ inline void xput(unsigned long long* x, int slot, unsigned tag, double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long t = (unsigned long long)tag << 32;
    __hip_atomic_store(x + 2 * slot, t | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(x + 2 * slot + 1, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// History term of step n at node i (levels t < n-1 in memory, the newest
// difference dnew in a register, the two-level corrections of the oldest
// intervals from d0m, d1m; Toeplitz weights or the step's own row; the
// first `folded` near-field groups from g.mid).
__device__ inline double hist_step(const Step1DArgs& g, int i, int n, double dnew, double d0m, double d1m,
                                   int folded = 0) {
    const double* om = g.omega + (int64_t)(n - g.n0) * g.omega_stride;
    double hm = hist_mem(g, i, n, (int64_t)(n - g.n0) * g.ld, om, folded);
    return hm + hist_tail(om[0], g.corr[2 * n], g.corr[2 * n + 1], n, g.near_lo, dnew, d0m, d1m);
}

// Steps [g.n, g.n_end) of one solve by its chunk workgroup c (0 <= c < g.C):
// the body of the single-solve launch and of the batched one.
__device__ __forceinline__ void step_block(const Step1DArgs& g, const int c) {
    __shared__ double rA[2][kMaxRed], rC[2][kMaxRed], rD[2][kMaxRed];
    __shared__ double pA[2][256], pC[2][256], pX[2][256], pY[2][256], pZ[2][256];
    __shared__ double sK[258];
    __shared__ double sdp[4];                  // wave maxima of the last solve's |U^{k+1} - U^k|
    __shared__ double sMeas[kMaxChunks];       // every chunk's published measure of the last barrier
    __shared__ int sh_fail;                    // a barrier wait ran out of time
    const int tid = threadIdx.x;
    const int N = g.N, C = g.C;
    const int s = (int)(((int64_t)c * N) / C);
    const int e = (int)(((int64_t)(c + 1) * N) / C) - 1;
    const int nc = e - s + 1;
    const int i = s + tid;
    const bool act = tid < nc;
    const int R2 = 2 * C;
    const bool series = g.ser_P > 0;
    const int halo_tid = (nc < 256) ? 255 : 0;
    const bool halo = tid == halo_tid && c > 0;
    const int wave = tid >> 6, lane = tid & 63;
    const int ni = nc - 2;
    const int q = tid - 1;
    const bool inner = (tid >= 1) && (tid <= nc - 2);

    // state carried from step to step: U^n and D[n-1] of the own node and of
    // the halo node s-1 (from memory at the first step of the launch)
    const int n_first = g.n;
    double un = act ? g.U[i] : 0.0;
    double dlast = 0.0;
    if (act && n_first >= 1) dlast = g.D[(series ? ((n_first - 1) & 1) : (n_first - 1)) * g.ld + i];
    double uh = halo ? g.U[s - 1] : 0.0;
    double dh = 0.0;
    if (halo && n_first >= 1) dh = g.D[(series ? ((n_first - 1) & 1) : (n_first - 1)) * g.ld + s - 1];
    // D[0], D[1] (corrections of the two oldest intervals): from memory when
    // written by an earlier launch, else taken when this launch makes them
    double d0m = (!series && act && n_first >= 1) ? g.D[i] : 0.0;
    double d1m = (!series && act && n_first >= 2) ? g.D[g.ld + i] : 0.0;
    int bk = 0;                                                    // Picard barriers passed in this launch
    // Near-field groups of this history block folded into g.mid so far.
    // The first build of a step s of the block that has a later step in the
    // block folds every group j (levels near_lo + 16 j ...+15) with
    // near_lo + 16 j + 17 <= s, i.e. whose levels are final (t <= s - 2)
    // whatever the previous step's decision; after the steps [n0, s] of the
    // block that makes (s - near_lo - 1) / 16 groups.  Those of earlier
    // launches of the block are counted here.
    int folded = 0;
    if (g.mid) {
        const int s_last = min(n_first - 1, g.n0 + g.blk - 2);
        folded = s_last >= g.n0 ? (s_last - g.near_lo - 1) / kNearGroup : 0;
    }
    // the exchange words of this workgroup, both parities, zeroed before its
    // first publication (its rows 2c and 2c+1, its measure)
    auto xbase = [&](int b) { return g.xch + (size_t)(b & 1) * kXchWords; };
    auto xtag = [&](int b) { return g.xtag | ((unsigned)(b + 1) & 0xffffu); };
    if (tid == 0 || tid == nc - 1) {
        for (int par = 0; par < 2; par++) {
            unsigned long long* x = g.xch + (size_t)par * kXchWords;
            const int r = (tid == 0) ? 2 * c : 2 * c + 1;
            for (int w = 0; w < 8; w++)
                __hip_atomic_store(x + 8 * r + w, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (tid == 0)
                for (int w = 0; w < 2; w++)
                    __hip_atomic_store(x + 2 * (kXchMeasure + c) + w, 0ull, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (tid == 0) sh_fail = 0;
    // the words of barrier b: thread j < C the measure of chunk j (into
    // sMeas), then (rows = true) thread r reduced rows r and r + 256; then a
    // workgroup barrier.  False when a wait ran out of time (the launch then
    // ends; the host reports it).
    auto read_barrier = [&](int b, bool rows, double (*rrow)[4]) -> bool {
        const unsigned long long* xr = xbase(b);
        const unsigned tg = xtag(b);
        const unsigned long long none = (unsigned long long)tg << 32;
        const int64_t tmax = g.wait_ticks;
        // load the words the predicate leaves to this thread until all carry
        // the tag (missing ones reloaded together); false after tmax ticks
        auto take = [&](unsigned long long* w, int cnt, const unsigned long long* src) -> bool {
            int64_t t0 = -1;
            for (;;) {
                bool all = true;
                for (int k = 0; k < cnt; k++) all = all && (unsigned)(w[k] >> 32) == tg;
                if (all) return true;
                const int64_t now = (int64_t)wall_clock64();
                if (t0 < 0) t0 = now;
                else if (now - t0 > tmax) return false;
                __builtin_amdgcn_s_sleep(1);
                for (int k = 0; k < cnt; k++)
                    if ((unsigned)(w[k] >> 32) != tg)
                        w[k] = __hip_atomic_load(src + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        };
        bool ok = true;
        // first the measures (one pair of words per chunk, the last words a
        // chunk publishes from its thread 0): few pollers while the others
        // are still on their way
        unsigned long long mw[2] = {none, none};
        if (tid < C) {
            const unsigned long long* src = xr + 2 * (kXchMeasure + tid);
            mw[0] = __hip_atomic_load(src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            mw[1] = __hip_atomic_load(src + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = take(mw, 2, src);
        }
        // then the rows, which have landed or are about to
        unsigned long long rw[2][8];
        if (rows) {
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 2; j++) {
                const int r = tid + 256 * j;
#pragma unroll
                for (int w = 0; w < 8; w++)
                    rw[j][w] = (r < R2) ? __hip_atomic_load(xr + 8 * r + w, __ATOMIC_RELAXED,
                                                            __HIP_MEMORY_SCOPE_AGENT)
                                        : none;
            }
#pragma unroll
            for (int j = 0; j < 2; j++)
                if (tid + 256 * j < R2) ok = ok && take(rw[j], 8, xr + 8 * (tid + 256 * j));
        }
        if (!ok) {
            __hip_atomic_store(g.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            sh_fail = 1;
        }
        auto dec = [](unsigned long long lo, unsigned long long hi) {
            return __longlong_as_double((long long)((hi << 32) | (lo & 0xffffffffull)));
        };
        if (tid < C) sMeas[tid] = dec(mw[0], mw[1]);
        if (rows)
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int qq = 0; qq < 4; qq++) rrow[j][qq] = dec(rw[j][2 * qq], rw[j][2 * qq + 1]);
        __syncthreads();
        return sh_fail == 0;
    };
    // the maximum of the staged measures, formed by every wave itself
    auto measures_max = [&]() {
        double dv = 0.0;
        for (int j = lane; j < C; j += 64) dv = fmax(dv, sMeas[j]);
        return wave_max(dv);
    };

    if (tid < 4) sdp[tid] = 0.0;
    // the chunk's max |U^{k+1} - U^k| of the last solve (thread 0, after a
    // workgroup barrier that follows that solve)
    auto last_measure = [&]() { return fmax(fmax(sdp[0], sdp[1]), fmax(sdp[2], sdp[3])); };

    // ---- one Picard build: coefficients at the iterate (uit, uith), rows,
    // chunk elimination (chunk factors into fx, fy, fz), the two reduced rows
    // published with the last solve's measure in the convergence slot, then
    // the grid barrier
    double fx = 0.0, fy = 0.0, fz = 0.0;
    // the part of the next step's history sum that is already in memory
    // (far partials and near levels t < n), formed while a barrier is pending
    double hm_next = 0.0;
    double tl_next[3] = {0.0, 0.0, 0.0};       // and its tail weights: omega[0], corr[2n], corr[2n+1]
    auto build = [&](double uit, double uith, double S, double Ln, double un_, int pre_n, int fold_n) -> bool {
        double K = 0.0, invC = 0.0;
        if (halo) {
            double Kl, iCl;
            vgm_coefs(uith, (s - 1) * g.h, g.ksat, g.coef_mode, Kl, iCl);
            sK[0] = Kl;
        }
        if (act) {
            vgm_coefs(uit, i * g.h, g.ksat, g.coef_mode, K, invC);
            sK[tid + 1] = K;
        }
        __syncthreads();
        double ar = 0.0, br = 1.0, cr = 0.0, dr = 0.0;
        if (act) {
            if (i == 0) { ar = 0.0; br = 1.0; cr = 0.0; dr = g.utop; }
            else if (i == N - 1) { ar = -1.0; br = 1.0; cr = 0.0; dr = 0.0; }
            else {
                double Km = sK[tid], Kp = sK[tid + 1];
                double A = Km * invC * g.h2inv;
                double B = Kp * invC * g.h2inv;
                double R = invC * (Kp * g.h2inv + Km * g.h2inv) + Ln;
                ar = A; br = -R; cr = B; dr = -Ln * un_ + S;
            }
        }
        double ib = inner ? rcp_nr(br) : 0.0;
        const int pp = chunk_solve(pA, pC, pX, pY, pZ, ni, q, inner, (tid == 1) ? 0.0 : ar * ib,
                                   (tid == nc - 2) ? 0.0 : cr * ib, dr * ib, (tid == 1) ? -ar * ib : 0.0,
                                   (tid == nc - 2) ? -cr * ib : 0.0);
        fx = 0.0; fy = 0.0; fz = 0.0;
        if (inner) { fx = pX[pp][q]; fy = pY[pp][q]; fz = pZ[pp][q]; }
        unsigned long long* xw = xbase(bk);
        const unsigned tg = xtag(bk);
        if (tid == 0) {
            double x, y, z;
            if (nc == 2) { x = 0.0; y = 0.0; z = 1.0; }
            else { x = pX[pp][0]; y = pY[pp][0]; z = pZ[pp][0]; }
            xput(xw, 4 * (2 * c) + 0, tg, ar);
            xput(xw, 4 * (2 * c) + 1, tg, br + cr * y);
            xput(xw, 4 * (2 * c) + 2, tg, cr * z);
            xput(xw, 4 * (2 * c) + 3, tg, dr - cr * x);
            xput(xw, kXchMeasure + c, tg, last_measure());
        }
        if (tid == nc - 1) {
            double x, y, z;
            if (nc == 2) { x = 0.0; y = 1.0; z = 0.0; }
            else { x = pX[pp][ni - 1]; y = pY[pp][ni - 1]; z = pZ[pp][ni - 1]; }
            xput(xw, 4 * (2 * c + 1) + 0, tg, ar * y);
            xput(xw, 4 * (2 * c + 1) + 1, tg, br + ar * z);
            xput(xw, 4 * (2 * c + 1) + 2, tg, cr);
            xput(xw, 4 * (2 * c + 1) + 3, tg, dr - ar * x);
        }
        // the next near-field group whose levels are final (t <= step - 2)
        // folded for the later steps of the block, while the exchange is
        // pending (first build of a step only)
        if (fold_n >= 0 && g.mid) {
            while (g.near_lo + kNearGroup * folded + kNearGroup + 1 <= fold_n && fold_n + 1 < g.n0 + g.blk) {
                if (act)
                    fold_near_group(g.D, g.ld, g.omega, g.omega_stride, g.n0, g.mid, i,
                                    g.near_lo + kNearGroup * folded, fold_n + 1, g.n0 + g.blk, folded == 0);
                folded++;
            }
        }
        if (pre_n >= 0 && act) {
            const double* om = g.omega + (int64_t)(pre_n - g.n0) * g.omega_stride;
            tl_next[0] = om[0];
            tl_next[1] = g.corr[2 * pre_n];
            tl_next[2] = g.corr[2 * pre_n + 1];
            hm_next = hist_mem(g, i, pre_n, (int64_t)(pre_n - g.n0) * g.ld, om, folded);
        }
        bk++;                                  // the solve reads this barrier's words
        return true;
    };
    // ---- after a build's barrier: the reduced rows; with check != 0 first
    // the decision on the published convergence slots (check 1: U^k met the
    // tolerance -> return 1, nothing solved; check 2: the tentatively ended
    // previous step did not -> return 2, nothing solved); then the reduced
    // solve and back-substitution (uit, uith updated), return 0.  Every wave
    // forms the decision's maximum itself, so it is uniform without a
    // workgroup barrier.  The chunk's max |U^{k+1} - U^k| is left as four
    // wave maxima in sdp for thread 0, which publishes it with the next build.
    auto solve = [&](int check, double& uit, double& uith) -> int {
        double rrow[2][4];
        if (!read_barrier(bk - 1, true, rrow)) return -1;
        if (check) {
            const bool met = measures_max() <= g.tol;
            if (check == 1 && met) return 1;
            if (check == 2 && !met) return 2;
        }
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = tid + 256 * j;
            if (r < R2) {
                double ib = rcp_nr(rrow[j][1]);
                rA[0][r] = rrow[j][0] * ib;
                rC[0][r] = rrow[j][2] * ib;
                rD[0][r] = rrow[j][3] * ib;
            }
        }
        __syncthreads();
        const double* xs = &rD[pcr_norm(rA, rC, rD, R2)][0];
        double unew = 0.0;
        if (act) {
            if (i == s) unew = xs[2 * c];
            else if (i == e) unew = xs[2 * c + 1];
            else unew = fx + fy * xs[2 * c] + fz * xs[2 * c + 1];
        }
        double unewh = halo ? xs[2 * c - 1] : 0.0;
        const double d = wave_max(act ? fabs(unew - uit) : 0.0);
        if (lane == 0) sdp[wave] = d;
        uit = unew;
        uith = unewh;
        return 0;
    };
    // ---- end of step m with U^{m+1} = (res, resh): D[m], U, carried state
    auto finish = [&](int m, double res, double resh) {
        double dn = res - un;
        if (act) {
            g.D[(series ? (m & 1) : m) * g.ld + i] = dn;
            g.U[i] = res;
        }
        dh = resh - uh;
        uh = resh;
        dlast = dn;
        un = res;
        if (m == 0) d0m = dn;
        if (m == 1) d1m = dn;
    };

    // A step whose last Picard iteration is due (k = picard_max - 1) ends
    // tentatively with U^k when the launch has a next step: its convergence
    // measure goes out in the convergence slots of the next step's first
    // build, so the barrier that decides it is the one that exchanges those
    // rows.  If it did not converge, every workgroup (the decision is
    // uniform) restores the step, runs its last iteration, rewrites D[n] and
    // U, and starts the next step again.  Series moments are updated in
    // place, so sum_alg 2 decides convergence before the step ends.
    bool pending = false;          // the previous step ended tentatively
    int sv_n = 0;                  // that step's state, for a rollback
    double sv_un = 0.0, sv_uh = 0.0, sv_dh = 0.0, sv_dlast = 0.0, sv_d0m = 0.0, sv_d1m = 0.0;
    double sv_S = 0.0, sv_Ln = 0.0, sv_uit = 0.0, sv_uith = 0.0;
    double hm_cur = 0.0;           // that part for the current step
    double tl_cur[3] = {0.0, 0.0, 0.0};
    bool have_hm = false;
    int nsolve = 0;                // Picard iterations solved (the decisions are uniform)
    int n = n_first;
    while (n < g.n_end) {
        // ---- history term of step n -------------------------------------
        // weights of step n: one Toeplitz vector, or the step's own row
        // (prescribed grid); local coefficient likewise
        const double Ln = g.Lv ? g.Lv[n] : g.L;
        double S = 0.0;
        if (n >= 1 && act) {
            if (series) {
                S = series_update(g.H, g.ld, i, g.ser_P, dlast, g.ser_c + (size_t)n * g.ser_P,
                                  g.ser_v + (size_t)n * g.ser_P);
            } else if (have_hm) {
                S = hm_cur + hist_tail(tl_cur[0], tl_cur[1], tl_cur[2], n, g.near_lo, dlast, d0m, d1m);
            } else {
                S = hist_step(g, i, n, dlast, d0m, d1m, folded);
            }
        }
        // the next step's in-memory history part goes with this step's first
        // barrier (levels t < n are written; D[n] enters through the tail)
        const int pre_n = (!series && n + 1 < g.n_end) ? n + 1 : -1;
        // Picard starts from 2 U^n - U^{n-1} (a single pass: lagged at U^n)
        double uit = un, uith = uh;
        if (g.extrap && n >= 1) { uit = un + dlast; uith = uh + dh; }
        bool tentative = false, rollback = false;
        for (int k = 0; k < g.picard_max; k++) {
            const bool last_due = k >= 1 && k == g.picard_max - 1;
            if (last_due && !series && n + 1 < g.n_end) {
                tentative = true;            // U^k ends the step, decided with the next build
                break;
            }
            if (last_due) {
                // the last build is skipped when U^k already met the
                // tolerance: decide that first (one barrier on the measure)
                __syncthreads();
                if (tid == 0) xput(xbase(bk), kXchMeasure + c, xtag(bk), last_measure());
                bk++;
                if (!read_barrier(bk - 1, false, nullptr)) return;
                if (measures_max() <= g.tol) break;           // U^k is the result (uniform decision)
            }
            const bool pend_check = k == 0 && pending;
            if (!build(uit, uith, S, Ln, un, k == 0 ? pre_n : -1, k == 0 ? n : -1)) return;
            const int r = solve(pend_check ? 2 : ((k >= 1 && !last_due) ? 1 : 0), uit, uith);
            if (r < 0) return;                                 // a barrier wait ran out of time
            if (r == 1) break;                                 // U^k is the result
            if (r == 2) { rollback = true; break; }
            nsolve++;
            pending = false;
        }
        if (rollback) {
            // the tentatively ended step sv_n needs its last iteration
            un = sv_un; uh = sv_uh; dh = sv_dh; dlast = sv_dlast; d0m = sv_d0m; d1m = sv_d1m;
            double ruit = sv_uit, ruith = sv_uith;
            if (!build(ruit, ruith, sv_S, sv_Ln, un, -1, -1)) return;
            if (solve(0, ruit, ruith) < 0) return;
            nsolve++;
            finish(sv_n, ruit, ruith);
            pending = false;
            n = sv_n + 1;
            continue;
        }
        if (tentative) {
            sv_n = n; sv_un = un; sv_uh = uh; sv_dh = dh; sv_dlast = dlast; sv_d0m = d0m; sv_d1m = d1m;
            sv_S = S; sv_Ln = Ln; sv_uit = uit; sv_uith = uith;
            pending = true;
        }
        // ---- step n is done: the last iterate is U^{n+1} -------------------
        finish(n, uit, uith);
        hm_cur = hm_next;
        tl_cur[0] = tl_next[0]; tl_cur[1] = tl_next[1]; tl_cur[2] = tl_next[2];
        have_hm = pre_n >= 0;
        n++;
    }
    // the launch's Picard iterations, counted once per solve (err[1])
    if (c == 0 && tid == 0) __hip_atomic_fetch_add(g.err + 1, nsolve, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(256) void k_step1d_block(Step1DArgs g) { step_block(g, (int)blockIdx.x); }

// Many independent solves in one cooperative launch: workgroups
// [t.first[j], t.first[j+1]) are the chunks of solve j.  Each solve keeps its
// own exchange words, convergence slots and Picard decisions, so its arithmetic
// is that of its single-solve launch; the fields that change from block to
// block come with the launch, the rest from the solve's row of `base`.
// Registers are held to 256 per lane (a few spills) so that two workgroups
// share a CU: a batch holds up to 512 chunk workgroups, and the latency-bound
// waves of two workgroups interleave on each SIMD.  (The single-solve kernel
// keeps 272 registers and one workgroup per CU, its fastest form.)
__global__ __launch_bounds__(256, 2) void k_step1d_batch(const Step1DArgs* __restrict__ base, BatchTick t) {
    const int wg = (int)blockIdx.x;
    int j = 0;
    while (j + 1 < t.count && wg >= t.first[j + 1]) j++;
    Step1DArgs g = base[j];
    g.n = t.n[j];
    g.n_end = t.n_end[j];
    g.n0 = t.n0[j];
    g.nsplit = t.nsplit[j];
    g.near_lo = t.near_lo[j];
    g.xtag = t.xtag[j];
    g.part_stride = t.part_stride[j];
    g.part = t.part[j];
    step_block(g, wg - t.first[j]);
}

// Barrier bound in wall-clock ticks: 2 s at the device's wall clock rate
// (hipDeviceAttributeWallClockRate, kHz).
int64_t barrier_wait_ticks() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
        khz = 100000;                                  // the 100 MHz constant clock of gfx9 parts
    return (int64_t)khz * 2000;
}

static bool coop_fits(const void* kern, int64_t wgs) {
    int dev = 0, coop = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev) != hipSuccess || !coop) return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return false;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 256, 0) != hipSuccess) return false;
    return (int64_t)per_cu * cus >= wgs;
}

// Whether the C workgroups of a cooperative block launch can be co-resident.
bool step1d_block_supported(int C) { return coop_fits((const void*)k_step1d_block, C); }
bool step1d_batch_supported(int64_t wgs) { return coop_fits((const void*)k_step1d_batch, wgs); }

fpde_status launch_step1d_batch(const Step1DArgs* base, const BatchTick& t, hipStream_t st) {
    if (t.count < 1 || t.count > kMaxBatch) { set_error("batched step: 1 to 32 solves"); return FPDE_ERR_ARG; }
    for (int j = 0; j < t.count; j++)
        if (t.first[j + 1] < t.first[j]) { set_error("batched step: bad workgroup table"); return FPDE_ERR_ARG; }
    const int wgs = t.first[t.count];
    if (wgs == 0) return FPDE_OK;
    const Step1DArgs* b = base;
    BatchTick tick = t;
    void* params[] = {&b, &tick};
    FPDE_HIP_TRY(hipLaunchCooperativeKernel((const void*)k_step1d_batch, dim3(wgs), dim3(256), params, 0, st));
    return FPDE_OK;
}

fpde_status launch_step1d_block(const Step1DArgs& a, hipStream_t st) {
    if (2 * a.C > kMaxRed || a.C > kMaxChunks) { set_error("1D line too long for the reduced system (N <= 65536)"); return FPDE_ERR_ARG; }
    if (a.sp_on) { set_error("cooperative step: the space-fractional sum runs with per-iteration launches"); return FPDE_ERR_ARG; }
    if (a.n_end <= a.n) return FPDE_OK;
    Step1DArgs arg = a;
    void* params[] = {&arg};
    FPDE_HIP_TRY(hipLaunchCooperativeKernel((const void*)k_step1d_block, dim3(a.C), dim3(256), params, 0, st));
    return FPDE_OK;
}

fpde_status launch_step1d(const Step1DArgs& a, hipStream_t st) {
    if (2 * a.C > kMaxRed) { set_error("1D line too long for the reduced system (N <= 65536)"); return FPDE_ERR_ARG; }
    hipLaunchKernelGGL(k_step1d, dim3(a.C), dim3(256), 0, st, a);
    return FPDE_OK;
}

}  // namespace fpde
