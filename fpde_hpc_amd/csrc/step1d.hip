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

// In-place parallel cyclic reduction on LDS rows [0,n), 256 threads, one RHS.
// On return d[r] holds the solution.  Rows outside [0,n) act as identity.
__device__ inline void pcr_block(double* a, double* b, double* c, double* d, int n) {
    const int tid = threadIdx.x;
    for (int delta = 1; delta < n; delta <<= 1) {
        double na[2], nb[2], nc[2], nd[2];
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = tid + 256 * j;
            if (r < n) {
                double ar = a[r], br = b[r], cr = c[r], dr = d[r];
                double k1 = 0.0, k2 = 0.0, am = 0.0, cm = 0.0, dm = 0.0, ap = 0.0, cp = 0.0, dp = 0.0;
                if (r - delta >= 0) { k1 = ar / b[r - delta]; am = a[r - delta]; cm = c[r - delta]; dm = d[r - delta]; }
                if (r + delta < n) { k2 = cr / b[r + delta]; ap = a[r + delta]; cp = c[r + delta]; dp = d[r + delta]; }
                na[j] = -k1 * am;
                nc[j] = -k2 * cp;
                nb[j] = br - k1 * cm - k2 * ap;
                nd[j] = dr - k1 * dm - k2 * dp;
            }
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 2; j++) {
            int r = tid + 256 * j;
            if (r < n) { a[r] = na[j]; b[r] = nb[j]; c[r] = nc[j]; d[r] = nd[j]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int r = tid + 256 * j;
        if (r < n) d[r] = d[r] / b[r];
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_step1d(Step1DArgs g) {
    __shared__ double ra[kMaxRed], rb[kMaxRed], rc[kMaxRed], rd[kMaxRed];
    __shared__ double pa[256], pb[256], pc[256], px[256], py[256], pz[256];
    __shared__ double sK[258];
    const int tid = threadIdx.x;
    const int c = blockIdx.x;
    const int N = g.N, C = g.C;
    const int s = (int)(((int64_t)c * N) / C);
    const int e = (int)(((int64_t)(c + 1) * N) / C) - 1;
    const int nc = e - s + 1;
    const int i = s + tid;
    const bool act = tid < nc;
    const int R2 = 2 * C;

    // ---- 1. finish the previous solve -------------------------------------
    double val = 0.0;
    if (g.recover) {
        const double* red = g.red + (size_t)(1 - g.parity) * R2 * 4;
        for (int r = tid; r < R2; r += 256) {
            ra[r] = red[4 * r + 0];
            rb[r] = red[4 * r + 1];
            rc[r] = red[4 * r + 2];
            rd[r] = red[4 * r + 3];
        }
        __syncthreads();
        pcr_block(ra, rb, rc, rd, R2);
        if (act) {
            if (i == s) val = rd[2 * c];
            else if (i == e) val = rd[2 * c + 1];
            else {
                val = g.xyz[i] + g.xyz[g.ld + i] * rd[2 * c] + g.xyz[2 * g.ld + i] * rd[2 * c + 1];
            }
        }
    }

    // ---- 2. advance the state -----------------------------------------------
    double un = 0.0, uit = 0.0, S = 0.0;
    const int n = g.n;
    if (g.k <= 0) {
        if (act) {
            un = g.U[i];
            double dnew = 0.0;
            if (g.recover) {                     // val = U^n, un = U^{n-1}
                dnew = val - un;
                g.D[(int64_t)(n - 1) * g.ld + i] = dnew;
                g.U[i] = val;
                un = val;
            } else if (n >= 1) {
                dnew = g.D[(int64_t)(n - 1) * g.ld + i];
            }
            if (g.k < 0) return;
            if (n >= 1) {
                // history term S_n: far partials + near levels + corrections
                double far = 0.0;
                for (int sp = 0; sp < g.nsplit; sp++) far += g.part[sp * g.part_stride + g.part_off + i];
                double nr0 = 0.0, nr1 = 0.0;
                int t = g.near_lo;
                for (; t + 2 <= n - 1; t += 2) {
                    nr0 = fma(g.omega[n - t - 1], g.D[(int64_t)t * g.ld + i], nr0);
                    nr1 = fma(g.omega[n - t - 2], g.D[(int64_t)(t + 1) * g.ld + i], nr1);
                }
                for (; t < n - 1; t++) nr0 = fma(g.omega[n - t - 1], g.D[(int64_t)t * g.ld + i], nr0);
                double newest = (g.near_lo <= n - 1) ? g.omega[0] * dnew : 0.0;
                double d0 = (n - 1 == 0) ? dnew : g.D[i];
                double cr = g.corr[2 * n] * d0;
                if (n >= 2) {
                    double d1 = (n - 1 == 1) ? dnew : g.D[g.ld + i];
                    cr = fma(g.corr[2 * n + 1], d1, cr);
                }
                S = far + (nr0 + nr1) + newest + cr;
            }
            g.Sbuf[i] = S;
        } else if (g.k < 0) {
            return;
        }
        uit = un;
    } else {
        if (act) {
            un = g.U[i];
            S = g.Sbuf[i];
            uit = val;
        }
    }
    if (g.k < 0) return;

    // ---- 3. coefficients at the iterate ------------------------------------
    double K = 0.0, invC = 0.0;
    if (act) {
        vgm_coefs(uit, i * g.h, g.ksat, g.coef_mode, K, invC);
        sK[tid + 1] = K;
    }
    if (tid == 0 && c > 0) {
        double ulo = g.recover ? rd[2 * c - 1] : g.U[s - 1];
        double Kl, iCl;
        vgm_coefs(ulo, (s - 1) * g.h, g.ksat, g.coef_mode, Kl, iCl);
        sK[0] = Kl;
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

    // ---- 5. chunk elimination: interior rows 1..nc-2 -------------------------
    const int ni = nc - 2;
    const int q = tid - 1;
    const bool inner = (tid >= 1) && (tid <= nc - 2);
    if (inner) {
        pa[q] = (tid == 1) ? 0.0 : ar;
        pc[q] = (tid == nc - 2) ? 0.0 : cr;
        pb[q] = br;
        px[q] = dr;
        py[q] = (tid == 1) ? -ar : 0.0;
        pz[q] = (tid == nc - 2) ? -cr : 0.0;
    }
    __syncthreads();
    for (int delta = 1; delta < ni; delta <<= 1) {
        double na = 0, nb = 1, ncc = 0, nx = 0, ny = 0, nz = 0;
        if (inner) {
            double k1 = 0, k2 = 0, am = 0, cm = 0, xm = 0, ym = 0, zm = 0, ap = 0, cp = 0, xp = 0, yp = 0, zp = 0;
            double aq = pa[q], cq = pc[q];
            if (q - delta >= 0) {
                k1 = aq / pb[q - delta];
                am = pa[q - delta]; cm = pc[q - delta]; xm = px[q - delta]; ym = py[q - delta]; zm = pz[q - delta];
            }
            if (q + delta < ni) {
                k2 = cq / pb[q + delta];
                ap = pa[q + delta]; cp = pc[q + delta]; xp = px[q + delta]; yp = py[q + delta]; zp = pz[q + delta];
            }
            na = -k1 * am;
            ncc = -k2 * cp;
            nb = pb[q] - k1 * cm - k2 * ap;
            nx = px[q] - k1 * xm - k2 * xp;
            ny = py[q] - k1 * ym - k2 * yp;
            nz = pz[q] - k1 * zm - k2 * zp;
        }
        __syncthreads();
        if (inner) { pa[q] = na; pb[q] = nb; pc[q] = ncc; px[q] = nx; py[q] = ny; pz[q] = nz; }
        __syncthreads();
    }
    if (inner) {
        double ib = 1.0 / pb[q];
        double x = px[q] * ib, y = py[q] * ib, z = pz[q] * ib;
        g.xyz[i] = x;
        g.xyz[g.ld + i] = y;
        g.xyz[2 * g.ld + i] = z;
        px[q] = x; py[q] = y; pz[q] = z;
    }
    __syncthreads();

    // ---- 6. reduced rows of the chunk ends ----------------------------------
    double* redw = g.red + (size_t)g.parity * R2 * 4;
    if (tid == 0) {
        double x, y, z;
        if (nc == 2) { x = 0.0; y = 0.0; z = 1.0; }
        else { x = px[0]; y = py[0]; z = pz[0]; }
        redw[4 * (2 * c) + 0] = ar;
        redw[4 * (2 * c) + 1] = br + cr * y;
        redw[4 * (2 * c) + 2] = cr * z;
        redw[4 * (2 * c) + 3] = dr - cr * x;
    }
    if (tid == nc - 1) {
        double x, y, z;
        if (nc == 2) { x = 0.0; y = 1.0; z = 0.0; }
        else { x = px[ni - 1]; y = py[ni - 1]; z = pz[ni - 1]; }
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
