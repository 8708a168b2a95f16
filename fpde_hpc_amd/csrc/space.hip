// Space-fractional explicit scheme of the reference (implicit=0, space_der=0,
// equation 0): per step, the flux differences of U^n
//   F[j] = (K_j (U_{j+1}-U_j) - K_{j-1} (U_j-U_{j-1})) / h^2   (Om1, :4176-4179)
// and the lower-triangular Toeplitz sum of calc_v (:2303-2335)
//   E_i = sum_{j=1}^{i-1} w(i-j) F[j] + corrections of j = 1, 2,
//   w(d) = Dg kb(gamma, j-1, j, i) (closed form, a function of d only).
// The sum costs N^2/2 FP64 multiply-adds per step.  It runs on the FP64
// matrix cores as a block-Toeplitz product: with i = 16 I + p and
// j = 16 J + q, the 16 x 16 block (I, J) is T_d[p][q] = w(16 d + p - q),
// d = I - J, so a tile of 256 rows (I = I0 .. I0+15) accumulates, per block
// diagonal d, T_d (A operand, 16 x 16) times the 16 x 16 matrix of F blocks
// F_{I-d} (B operand) in four 4-deep v_mfma_f64_16x16x4f64 steps.  Row tiles
// x chunks of kSpDiag block diagonals form the grid; the chunks' partial
// sums are added by the consumer (space_sum).
#include "fpde_internal.h"

namespace fpde {

// F of U (in memory) with K at U; zero outside the interior [1, N-2] up to ldp.
__global__ __launch_bounds__(256) void k_space_flux(const double* __restrict__ U, int N, double h, double ksat,
                                                    int mode, double* __restrict__ F, int64_t ldp) {
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= ldp) return;
    double f = 0.0;
    if (j >= 1 && j <= N - 2) {
        const double um = U[j - 1], u0 = U[j], up = U[j + 1];
        double Km, Kj, iC;
        vgm_coefs(um, (double)(j - 1) * h, ksat, mode, Km, iC);
        vgm_coefs(u0, (double)j * h, ksat, mode, Kj, iC);
        const double hh = h * h;
        f = Kj * up / hh - (Kj / hh + Km / hh) * u0 + Km * um / hh;
    }
    F[j] = f;
}

// Partial sums of the spatial sum on the matrix cores: one work unit (row
// tile t, block-diagonal chunk y, space_unit) per workgroup.
__global__ __launch_bounds__(256) void k_space_apply(const double* __restrict__ F, int N,
                                                     const double* __restrict__ om, double* __restrict__ P,
                                                     int64_t ldp) {
    __shared__ SpaceUnitLds lds;
    const int t = blockIdx.x, y = blockIdx.y;
    if (y >= space_chunks(t)) return;                  // uniform over the block
    space_unit(F, N, om, P, ldp, t, y, lds);
}

// V[i] = E_i on the interior rows, 0 on rows 0 and N-1
__global__ __launch_bounds__(256) void k_space_reduce(double* __restrict__ V, const double* __restrict__ P,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ corr, int N, int64_t ldp) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= N) return;
    V[i] = (i >= 1 && i <= N - 2) ? space_sum(P, ldp, corr, F, i) : 0.0;
}

// 3D: the space-fractional sums of every line of one LOD sweep.  The
// reference's sweeps switch Z, dL and N to the sweep axis (y sweep:
// fpde_hpc_1d.cpp:5262-5265), so each line carries the 1D operator along its
// own axis: E[line][i] = sum_{j<i} w(i-j) F[line][j] + the two surface
// corrections, F from the field U the sweep starts from and its conductivity
// K.  A workgroup takes 16 lines (16 consecutive z for the lateral axes, 16
// consecutive x for the depth axis, so loads and stores are coalesced) and
// forms E = F T^T on v_mfma_f64_16x16x4f64 with M = line, N = the row i
// (16 per tile), K = the column j; wave w takes the row tiles w, w+4, ...
// Field layout idx = z + n (x + n y); axis 0 = z (depth), 1 = x, 2 = y.
constexpr int kSp3Max = 256;
__global__ __launch_bounds__(256) void k_space3d(const double* __restrict__ U, const double* __restrict__ K,
                                                 int n, int axis, double h, const double* __restrict__ om,
                                                 const double* __restrict__ corr, double* __restrict__ E) {
    constexpr int LD = kSp3Max + 1;            // odd row stride: the 16 lines of an A read hit distinct banks
    __shared__ double sA[16][LD];              // U, then F, then E
    __shared__ double Ws[kSp3Max + 48];        // w(m) at m + 16, m in [-16, n + 31]
    __shared__ double sF12[16][2];
    const int tid = threadIdx.x;
    const int g = blockIdx.x, o = blockIdx.y;  // line group of 16, outer index
    int64_t base0, lstride, stride;            // line l of the group starts at base0 + l*lstride
    if (axis == 0) { base0 = (int64_t)n * (16 * g + (int64_t)n * o); lstride = n; stride = 1; }   // x = 16g+l, y = o
    else if (axis == 1) { base0 = 16 * g + (int64_t)n * n * o; lstride = 1; stride = n; }         // z = 16g+l, y = o
    else { base0 = 16 * g + (int64_t)n * o; lstride = 1; stride = (int64_t)n * n; }              // z = 16g+l, x = o
    const int nl = min(16, n - 16 * g);        // lines present in this group
    const int np = (n + 15) & ~15;
    const int total = 16 * np;
    // element q -> (line, i): consecutive threads walk the contiguous index
    auto line_of = [&](int q) { return axis == 0 ? q / np : (q & 15); };
    auto row_of = [&](int q) { return axis == 0 ? q % np : (q >> 4); };
    for (int q = tid; q < total; q += 256) {
        const int l = line_of(q), i = row_of(q);
        sA[l][i] = (l < nl && i < n) ? U[base0 + lstride * l + stride * i] : 0.0;
    }
    for (int q = tid; q < np + 48; q += 256) {
        const int m = q - 16;
        Ws[q] = (m >= 1 && m <= n) ? om[m - 1] : 0.0;
    }
    __syncthreads();
    // F (Om1 equation 0, :4176-4179) of the thread's elements, then into sA
    constexpr int PER = (16 * kSp3Max) / 256;
    double f[PER];
    const double hh = h * h;
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int q = tid + 256 * k;
        f[k] = 0.0;
        if (q < total) {
            const int l = line_of(q), i = row_of(q);
            if (l < nl && i >= 1 && i <= n - 2) {
                const int64_t id = base0 + lstride * l + stride * i;
                const double Kj = K[id], Km = K[id - stride];
                f[k] = Kj * sA[l][i + 1] / hh - (Kj / hh + Km / hh) * sA[l][i] + Km * sA[l][i - 1] / hh;
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int q = tid + 256 * k;
        if (q < total) sA[line_of(q)][row_of(q)] = f[k];
    }
    __syncthreads();
    if (tid < 32) sF12[tid >> 1][tid & 1] = sA[tid >> 1][1 + (tid & 1)];
    const int lane = tid & 63, wave = tid >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;
    const int ntile = np / 16;
    sd4_t acc[kSp3Max / 64];
#pragma unroll
    for (int c = 0; c < kSp3Max / 64; c++) acc[c] = (sd4_t){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int c = 0; c < kSp3Max / 64; c++) {
        const int it = wave + 4 * c;
        if (it < ntile) {
            const int i0 = 16 * it;
            for (int j0 = 0; j0 < i0 + 16; j0 += 4) {
                const double a = sA[r16][j0 + k4];                    // F[line][j]
                const double b = Ws[i0 + r16 - j0 - k4 + 16];         // w(i - j)
                acc[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[c], 0, 0, 0);
            }
        }
    }
    __syncthreads();                           // every wave is done with F
#pragma unroll
    for (int c = 0; c < kSp3Max / 64; c++) {
        const int it = wave + 4 * c;
        if (it < ntile) {
#pragma unroll
            for (int v = 0; v < 4; v++) sA[4 * v + k4][16 * it + r16] = acc[c][v];
        }
    }
    __syncthreads();
    for (int q = tid; q < total; q += 256) {
        const int l = line_of(q), i = row_of(q);
        if (l < nl && i < n) {
            double e = 0.0;
            if (i >= 1 && i <= n - 2)
                e = fma(corr[2 * (i - 1) + 1], sF12[l][1], fma(corr[2 * (i - 1)], sF12[l][0], sA[l][i]));
            E[base0 + lstride * l + stride * i] = e;
        }
    }
}

fpde_status launch_space3d(const double* U, const double* K, int n, int axis, double h, const double* sp_omega,
                           const double* sp_corr, double* E, hipStream_t st) {
    if (n > kSp3Max) { set_error("3D space-fractional sweep: at most 256 nodes per axis"); return FPDE_ERR_ARG; }
    hipLaunchKernelGGL(k_space3d, dim3((n + 15) / 16, n), dim3(256), 0, st, U, K, n, axis, h, sp_omega, sp_corr, E);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_space_flux(const double* U, int N, double h, double ksat, int mode, double* F, int64_t ldp,
                              hipStream_t st) {
    if (ldp < N || ldp % kSpTile) { set_error("space flux: bad padded length"); return FPDE_ERR_ARG; }
    hipLaunchKernelGGL(k_space_flux, dim3((unsigned)(ldp / 256)), dim3(256), 0, st, U, N, h, ksat, mode, F, ldp);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_space_apply(const double* F, int N, const double* sp_omega, double* P, int64_t ldp,
                               hipStream_t st) {
    if (ldp < N || ldp % kSpTile) { set_error("space sum: bad padded length"); return FPDE_ERR_ARG; }
    const int T = (int)(ldp / kSpTile);
    hipLaunchKernelGGL(k_space_apply, dim3(T, space_chunks(T - 1)), dim3(256), 0, st, F, N, sp_omega, P, ldp);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_space_reduce(double* V, const double* P, const double* F, const double* sp_corr, int N,
                                int64_t ldp, hipStream_t st) {
    hipLaunchKernelGGL(k_space_reduce, dim3((N + 255) / 256), dim3(256), 0, st, V, P, F, sp_corr, N, ldp);
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

}  // namespace fpde
