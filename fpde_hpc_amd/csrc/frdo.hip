// The GPU kernels of fr_do.cpp's program (Fill, DOC, Sum, eq_F, eq_Solve:
// fr_do.cpp:136-297) behind the OpenCL_* facade, for the reference's host
// code of the distributed-order derivative and equation (DO_CUDA_init,
// DO_Cder_CUDA, eq_one_step_cuda, eq_solve_cuda: :310-430, :781-863).  The
// element type T is the program's (float or double, -DT); the facade checks
// every index range against the buffer sizes before a launch.  The
// distributed-order derivative is computed with one thread per (grid point,
// time level, order) output instead of one per (grid point, order) walking
// all levels, in the same summation order; DOC's tensor-core mode 1 gives the
// same sums in T (it multiplied half-precision operands).
#include "fpde_internal.h"

namespace fpde {

// bk of the program text (:129-134): d1^(1-b) - d2^(1-b), a zero distance
// leaving the other's power alone; f(x) = x^2 is the function the text
// differentiates (:125-128).
template <typename T> __device__ inline T frdo_bk(T d1, T d2, T beta) {
    if (d1 == (T)0) return pow(d2, (T)1 - beta);
    if (d2 == (T)0) return pow(d1, (T)1 - beta);
    return pow(d1, (T)1 - beta) - pow(d2, (T)1 - beta);
}

// Fill: df[id] = (f((id+1) tau) - f(id tau)) / tau, bkk[id na + j] = bk((id+1) tau, id tau, a_j).
template <typename T>
__global__ void k_frdo_fill(int64_t n, T tau, int na, T* __restrict__ df, T* __restrict__ bkk) {
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n) return;
    const T x1 = tau * (T)(id + 1), x0 = tau * (T)id;
    df[id] = (x1 * x1 - x0 * x0) / tau;
    for (int j = 0; j < na; j++)
        bkk[id * na + j] = frdo_bk<T>(x1, x0, (na == 1) ? (T)0 : (T)j / (T)(na - 1));
}

// DOC: do1[pt na ntau + na t + a] = sum_{j in [j0, min(t, maxt))} f_j bkk[(t-j-1) na + a] / gs[a]
// for t in [t0, ntau), f_j = df[j] or (F[pt tot + j + 1] - F[pt tot + j]) / tau.
template <typename T>
__global__ void k_frdo_doc(int nf, int ntau, T tau, int na, int t0, int j0, int maxt, int use_f,
                           const T* __restrict__ F, const T* __restrict__ df, const T* __restrict__ bkk,
                           T* __restrict__ do1, const T* __restrict__ gs, int64_t tot) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nt = ntau - t0;
    if (nt <= 0 || idx >= (int64_t)nf * nt * na) return;
    const int a = (int)(idx % na);
    const int64_t r = idx / na;
    const int t = t0 + (int)(r % nt), pt = (int)(r / nt);
    const int jhi = (maxt == -1) ? t : ((t > maxt) ? maxt : t);
    T sum = (T)0;
    for (int j = (j0 > 0 ? j0 : 0); j < jhi; j++) {
        const T f = use_f ? (F[pt * tot + j + 1] - F[pt * tot + j]) / tau : df[j];
        sum += f * bkk[(int64_t)(t - j - 1) * na + a];
    }
    sum /= gs[a];
    do1[(int64_t)pt * na * ntau + (int64_t)na * t + a] = sum;
}

// Sum: do2[pt ntau2 + t] (+)= sum_{i < na-1} (wa_i do1_i + wa_{i+1} do1_{i+1}) / 2.
template <typename T>
__global__ void k_frdo_sum(int nf, int ntau, int ntau2, int t0, int na, int append, const T* __restrict__ do1,
                           T* __restrict__ do2, const T* __restrict__ wa) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nt = ntau - t0;
    if (nt <= 0 || idx >= (int64_t)nf * nt) return;
    const int pt = (int)(idx / nt), t = t0 + (int)(idx % nt);
    const T* d = do1 + (int64_t)pt * na * ntau + (int64_t)na * t;
    T sum = (T)0;
    for (int i = 0; i < na - 1; i++) sum += (T)0.5 * (wa[i] * d[i] + wa[i + 1] * d[i + 1]);
    T* o = do2 + (int64_t)pt * ntau2 + t;
    if (append == 0) *o = sum;
    else *o += sum;
}

// eq_ff of the program text (:256-263): the source of example 0.  The text
// multiplies a float constant for pi: 2 pi and 800 pi pi are formed in float
// before they meet t and x, so the constants here are too.
template <typename T> __device__ inline T frdo_ff(T t, T x, double ex) {
    if (ex == 0.0) {
        const float pf = (float)M_PI;
        const T two_pi = (T)(2.0f * pf), k800 = (T)(800.0f * pf * pf);
        return sin(two_pi * x) * (k800 * t * t * t * t + (2400 * (t * t * t * t - t * t * t) / log(t)));
    }
    return (T)0;
}

// eq_F: the Crank-Nicolson right-hand side of level j+1 at node id.
template <typename T>
__global__ void k_frdo_eq_f(int nx, int Nt, T D, T b0, T tau, T h, int j, const T* __restrict__ C,
                            const T* __restrict__ DF, T* __restrict__ F, double ex) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx) return;
    if (id == 0 || id == nx - 1) { F[id] = (T)0; return; }
    const int64_t c = (int64_t)id * Nt + j;
    const T rp = (T)(-0.5) * (D / (h * h)) * (C[c - Nt] - (T)2 * C[c] + C[c + Nt]);
    T lp = -(b0 * C[c] / tau) - (T)0.5 * (frdo_ff<T>((T)j * tau, (T)id * h, ex) + frdo_ff<T>((T)(j + 1) * tau, (T)id * h, ex));
    const T nlp = DF[c + 1];
    if (!isfinite(lp)) lp = -(b0 * C[c] / tau) - frdo_ff<T>((T)(j + 1) * tau, (T)id * h, ex);
    F[id] = rp + lp + nlp;
}

// eq_Solve: C[id nt + t] = sum_i inv[id nx + i] F[i] / A (the tridiagonal
// system through its known inverse).
template <typename T>
__global__ void k_frdo_eq_solve(int t, int nt, int nx, T A, T* __restrict__ C, const T* __restrict__ inv,
                                const T* __restrict__ F) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= nx) return;
    T sum = (T)0;
    const T* row = inv + (int64_t)id * nx;
    for (int i = 0; i < nx; i++) sum += row[i] * F[i];
    C[(int64_t)id * nt + t] = sum / A;
}

static dim3 grid_for(int64_t n) { return dim3((unsigned)((n + 255) / 256)); }

template <typename T>
static fpde_status frdo_launch_t(const FrdoLaunch& L, hipStream_t st) {
    switch (L.kernel) {
        case FRDO_FILL:
            hipLaunchKernelGGL(k_frdo_fill<T>, grid_for(L.n), dim3(256), 0, st, L.n, (T)L.tau, L.na, (T*)L.b[0],
                               (T*)L.b[1]);
            break;
        case FRDO_DOC: {
            const int64_t work = (int64_t)L.nf * (L.ntau - L.t0) * L.na;
            if (work > 0)
                hipLaunchKernelGGL(k_frdo_doc<T>, grid_for(work), dim3(256), 0, st, L.nf, L.ntau, (T)L.tau, L.na,
                                   L.t0, L.j0, L.maxt, L.use_f, (const T*)L.b[0], (const T*)L.b[1],
                                   (const T*)L.b[2], (T*)L.b[3], (const T*)L.b[4], L.tot);
            break;
        }
        case FRDO_SUM: {
            const int64_t work = (int64_t)L.nf * (L.ntau - L.t0);
            if (work > 0)
                hipLaunchKernelGGL(k_frdo_sum<T>, grid_for(work), dim3(256), 0, st, L.nf, L.ntau, L.ntau2, L.t0,
                                   L.na, L.append, (const T*)L.b[0], (T*)L.b[1], (const T*)L.b[2]);
            break;
        }
        case FRDO_EQ_F:
            hipLaunchKernelGGL(k_frdo_eq_f<T>, grid_for(L.nx), dim3(256), 0, st, L.nx, L.Nt, (T)L.D, (T)L.b0,
                               (T)L.tau, (T)L.h, L.j, (const T*)L.b[0], (const T*)L.b[1], (T*)L.b[2], L.ex);
            break;
        case FRDO_EQ_SOLVE:
            hipLaunchKernelGGL(k_frdo_eq_solve<T>, grid_for(L.nx), dim3(256), 0, st, L.j, L.Nt, L.nx, (T)L.A,
                               (T*)L.b[0], (const T*)L.b[1], (const T*)L.b[2]);
            break;
        default: set_error("unknown fr_do kernel"); return FPDE_ERR_ARG;
    }
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_frdo(const FrdoLaunch& L, hipStream_t st) {
    return L.is_double ? frdo_launch_t<double>(L, st) : frdo_launch_t<float>(L, st);
}

}  // namespace fpde
