// History-sum kernels of the Caputo L1 derivative (gfx950).
//
// Storage: D[t][i] = U^{t+1}_i - U^t_i, level-major with node index fastest
// (leading dimension ld, a multiple of 64), so one level is a contiguous row
// and every wavefront reads 64 consecutive nodes.  The step-n history term
// is S_n[i] = sum_{t<n} w(n,t) D[t][i] with the Toeplitz weight
// w(n,t) = omega[n-t-1] (plus the two-level correction added by the step
// kernel).  Kernels here:
//   k_hist_plain   : FP64 streaming sum over a t-range, split over t-chunks
//                    (one partial row per chunk), HBM bound, 8 B per term;
//   k_ref_sum      : the reference "Sum" kernel semantics on level arrays
//                    (fpde_hpc_1d.cpp:1247-1284) behind the OpenCL_* facade;
//   k_gemm_f64     : time-block GEMM  P[ks][r][i] = sum_t omega[n0+r-t-1] D[t][i]
//                    for the B future steps r of a block, on v_mfma_f64_16x16x4;
//   k_pack_split / k_gemm_split : the same GEMM on split bf16/fp16 operands
//                    (hi+lo or hi+mid+lo) with fp32 MFMA accumulation flushed
//                    to fp64 after every 32-level group.
#include "fpde_internal.h"

namespace fpde {

// ---------------------------------------------------------------------------
// Plain FP64 streaming kernel: part[ks*ld + i] = sum_{t in chunk ks, t < t_hi}
// omega[n-t-1] * D[t*ld + i].  256 threads, 2 nodes per thread (16-B loads);
// the weights of the chunk are staged in LDS windows of kPlainWin levels
// (wave-uniform reads, one barrier pair per window).
constexpr int kPlainWin = 1024;
__global__ __launch_bounds__(256) void k_hist_plain(const double* __restrict__ D, int64_t ld,
                                                    int nodes, int n, int t_hi, int chunk,
                                                    const double* __restrict__ omega,
                                                    double* __restrict__ part) {
    __shared__ double wwin[kPlainWin];                // wwin[t - w0] = omega[n-t-1]
    const int i2 = (blockIdx.x * 256 + threadIdx.x) * 2;
    const bool ok = i2 < nodes;                       // threads past the nodes still stage weights
    const int ks = blockIdx.y;
    const int t0 = ks * chunk;
    const int t1 = min(t_hi, t0 + chunk);
    double a0 = 0.0, a1 = 0.0, b0 = 0.0, b1 = 0.0;
    for (int w0 = t0; w0 < t1; w0 += kPlainWin) {
        const int w1 = min(t1, w0 + kPlainWin);
        __syncthreads();
        for (int q = threadIdx.x; q < w1 - w0; q += 256) wwin[q] = omega[n - (w0 + q) - 1];
        __syncthreads();
        if (!ok) continue;
        int t = w0;
        for (; t + 4 <= w1; t += 4) {
            double2 d0 = *reinterpret_cast<const double2*>(D + (int64_t)(t + 0) * ld + i2);
            double2 d1 = *reinterpret_cast<const double2*>(D + (int64_t)(t + 1) * ld + i2);
            double2 d2 = *reinterpret_cast<const double2*>(D + (int64_t)(t + 2) * ld + i2);
            double2 d3 = *reinterpret_cast<const double2*>(D + (int64_t)(t + 3) * ld + i2);
            const double* w = wwin + (t - w0);
            double x0 = w[0], x1 = w[1], x2 = w[2], x3 = w[3];
            a0 = fma(x0, d0.x, a0); a1 = fma(x0, d0.y, a1);
            b0 = fma(x1, d1.x, b0); b1 = fma(x1, d1.y, b1);
            a0 = fma(x2, d2.x, a0); a1 = fma(x2, d2.y, a1);
            b0 = fma(x3, d3.x, b0); b1 = fma(x3, d3.y, b1);
        }
        for (; t < w1; t++) {
            double2 d0 = *reinterpret_cast<const double2*>(D + (int64_t)t * ld + i2);
            double x0 = wwin[t - w0];
            a0 = fma(x0, d0.x, a0); a1 = fma(x0, d0.y, a1);
        }
    }
    if (!ok) return;
    double2 r;
    r.x = a0 + b0;
    r.y = a1 + b1;
    *reinterpret_cast<double2*>(part + (int64_t)ks * ld + i2) = r;
}

// ---------------------------------------------------------------------------
// The reference Sum kernel, sum_alg 0/1 (fpde_hpc_1d.cpp:1247-1284): newest
// interval first, chunks of BS levels; sum_alg 1 stops after the first chunk whose
// weights are all below sum_param in magnitude.  One thread per node.
__global__ __launch_bounds__(256) void k_ref_sum(double* __restrict__ S, const double* __restrict__ old,
                                                 int64_t old_size, int64_t os, int64_t nodes,
                                                 const double* __restrict__ tau,
                                                 const double* __restrict__ kb, int sum_alg,
                                                 double sum_param, int bs) {
    int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= nodes) return;
    double time_sum = 0.0;
    if (old_size >= 2) {
        for (int64_t t = old_size - 2; t >= 0; t -= bs) {
            int br = 0, cnt = 0;
            for (int tt = 0; tt < bs && t - tt >= 0; tt++) {
                int64_t l = t - tt;
                double kv = kb[l];
                double diff = old[(l + 1) * os + id] - old[l * os + id];
                time_sum += kv * diff / tau[l];
                if (fabs(kv) < sum_param) br++;
                cnt++;
            }
            if (sum_alg == 1 && br == bs) break;
            (void)cnt;
        }
    }
    S[id] = time_sum;
}

// ---------------------------------------------------------------------------
// The reference Sum kernel, sum_alg 2 (fpde_hpc_1d.cpp:1285-1302): per node,
// the moments Htmp[id*P + j] += gt^{j+1} coef[j] (old[1][id] - old[0][id])
// (non-finite factors read as 0, as the reference kernel does), then
// S = sum_j v1_j Htmp[id*P + j] / tau[tstep] with the coefficients of
// _kb_row2_fixed_coefs (:1159-1180): v1_0 = gtj^-alpha,
// v1_j = v1_{j-1} (alpha+j-1)/(j gtj).
__global__ __launch_bounds__(256) void k_ref_series(double* __restrict__ S, const double* __restrict__ old,
                                                    int64_t old_size, int64_t os, int64_t nodes,
                                                    const double* __restrict__ tau, int tstep,
                                                    double* __restrict__ Htmp, const double* __restrict__ coef,
                                                    int P, double gtj, double gt, double alpha) {
    int64_t id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (id >= nodes) return;
    double ts = 0.0;
    if (old_size >= 2) {
        double cur = old[os + id] - old[id];
        double* h = Htmp + id * P;
        double m = gt;
        for (int j = 0; j < P; j++) {
            double cc = coef[j];
            if (!isfinite(cc)) cc = 0.0;
            double mm = isfinite(m) ? m : 0.0;
            h[j] += mm * cc * cur;
            m *= gt;
        }
        double v1 = pow(gtj, -alpha);
        for (int j = 0; j < P; j++) {
            ts += (isfinite(v1) ? v1 : 0.0) * h[j];
            double i = (double)(j + 1);
            v1 *= -(-alpha - i + 1.0) / (i * gtj);
        }
        ts /= tau[tstep];
    }
    S[id] = ts;
}

fpde_status launch_ref_series(double* S, const double* old, int64_t old_size, int64_t os, int64_t nodes,
                              const double* tau, int tstep, double* Htmp, const double* coef, int P, double gtj,
                              double gt, double alpha, hipStream_t st) {
    if (nodes <= 0) return FPDE_OK;
    hipLaunchKernelGGL(k_ref_series, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, st, S, old, old_size, os,
                       nodes, tau, tstep, Htmp, coef, P, gtj, gt, alpha);
    return FPDE_OK;
}

// ---------------------------------------------------------------------------
// MFMA operand types
typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf8_t __attribute__((ext_vector_type(8)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));

// FP64 time-block GEMM.  Grid: x = 64-node tiles, y = K splits (t-chunks).
// Block 256 = 4 waves; wave w owns nodes [16w,16w+16) of the tile and all BR
// rows (BR/16 MFMA tiles).  v_mfma_f64_16x16x4f64 operand layout (checked
// on gfx950): A lane l = A[l%16][l/16], B lane l = B[l/16][l%16], and the
// four results of lane l are D[4v + l/16][l%16], v=0..3 (rows interleaved,
// unlike the f32-output MFMAs where lane l holds rows 4(l/16)+v).
template <int BR>
__global__ __launch_bounds__(256) void k_gemm_f64(const double* __restrict__ D, int64_t ld, int nodes,
                                                  int n0, int tc, int kchunk,
                                                  const double* __restrict__ omega,
                                                  double* __restrict__ P) {
    constexpr int RT = BR / 16;
    constexpr int SLAB = 1024;                  // t-values per LDS weight window
    __shared__ double win[BR + SLAB];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ks = blockIdx.y;
    const int tlo = ks * kchunk;
    const int thi = min(tc, tlo + kchunk);
    const int node = blockIdx.x * 64 + wave * 16 + (lane & 15);
    const bool node_ok = node < nodes;
    const int kk = lane >> 4;          // k index within the 4-deep MFMA
    d4_t acc[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) acc[rt] = (d4_t){0.0, 0.0, 0.0, 0.0};
    for (int s0 = tlo; s0 < thi; s0 += SLAB) {
        const int s1 = min(thi, s0 + SLAB);
        // win[q] = omega[n0 - s1 + q], so omega[n0+r-t-1] = win[r - t - 1 + s1]
        const int W = BR + (s1 - s0) - 1;
        __syncthreads();
        for (int q = threadIdx.x; q < W; q += 256) win[q] = omega[n0 - s1 + q];
        __syncthreads();
        // four 4-deep K steps per iteration; the next iteration's four B
        // loads are in flight while this one's MFMAs run
        double bn[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int tt = s0 + 4 * q + kk;
            bn[q] = (tt < s1 && node_ok) ? D[(int64_t)tt * ld + node] : 0.0;
        }
        for (int t = s0; t < s1; t += 16) {
            double b[4];
#pragma unroll
            for (int q = 0; q < 4; q++) b[q] = bn[q];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int tt = t + 16 + 4 * q + kk;
                bn[q] = (tt < s1 && node_ok) ? D[(int64_t)tt * ld + node] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int tt = t + 4 * q + kk;
                bool ok = tt < s1;
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    int r = rt * 16 + (lane & 15);
                    double a = ok ? win[r - tt - 1 + s1] : 0.0;
                    acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[q], acc[rt], 0, 0, 0);
                }
            }
        }
    }
    if (node_ok) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int r = rt * 16 + 4 * v + (lane >> 4);
                P[((int64_t)ks * BR + r) * ld + node] = acc[rt][v];
            }
        }
    }
}

// FP64 time-block GEMM with dense weights (a prescribed non-uniform grid,
// where the weights are not Toeplitz): P[ks][r][i] = sum_t W[r][t] D[t][i],
// W row-major with leading dimension ldw.  Same tiling and MFMA layout as
// k_gemm_f64; the A operand comes from an LDS tile of BR x 64 weights.
template <int BR>
__global__ __launch_bounds__(256) void k_gemm_f64_dense(const double* __restrict__ D, int64_t ld, int nodes,
                                                        int tc, int kchunk, const double* __restrict__ W,
                                                        int64_t ldw, double* __restrict__ P) {
    constexpr int RT = BR / 16;
    constexpr int SL = 64;
    __shared__ double win[BR][SL + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ks = blockIdx.y;
    const int tlo = ks * kchunk;
    const int thi = min(tc, tlo + kchunk);
    const int node = blockIdx.x * 64 + wave * 16 + (lane & 15);
    const bool node_ok = node < nodes;
    const int kk = lane >> 4;
    d4_t acc[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) acc[rt] = (d4_t){0.0, 0.0, 0.0, 0.0};
    for (int s0 = tlo; s0 < thi; s0 += SL) {
        const int s1 = min(thi, s0 + SL);
        __syncthreads();
        for (int q = threadIdx.x; q < BR * SL; q += 256) {
            int r = q / SL, c = q % SL;
            win[r][c] = (s0 + c < s1) ? W[(int64_t)r * ldw + s0 + c] : 0.0;
        }
        __syncthreads();
        for (int t = s0; t < s1; t += 16) {
            double b[4];
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int tt = t + 4 * q + kk;
                b[q] = (tt < s1 && node_ok) ? D[(int64_t)tt * ld + node] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 4; q++) {
                int tt = t + 4 * q + kk;
                bool ok = tt < s1;
#pragma unroll
                for (int rt = 0; rt < RT; rt++) {
                    int r = rt * 16 + (lane & 15);
                    double a = ok ? win[r][tt - s0] : 0.0;
                    acc[rt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[q], acc[rt], 0, 0, 0);
                }
            }
        }
    }
    if (node_ok) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int r = rt * 16 + 4 * v + (lane >> 4);
                P[((int64_t)ks * BR + r) * ld + node] = acc[rt][v];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Split storage.  Group g holds levels [32g, 32g+32); for node i the 32
// values of each component are contiguous: comp[(g*ld + i)*32 + e].  That is
// exactly the B operand of v_mfma_f32_16x16x32_{bf16,f16}: lane l needs
// B[8(l/16)+e][l%16], e=0..7, i.e. 16 contiguous bytes.
// fp16 groups carry a power-of-two scale per (group,node): value = scale*(hi+lo).
template <typename T> struct SplitTraits;
template <> struct SplitTraits<__bf16> {
    typedef bf8_t vec;
    static constexpr bool scaled = false;
};
template <> struct SplitTraits<_Float16> {
    typedef h8_t vec;
    static constexpr bool scaled = true;
};

// Pack levels of complete groups [g_lo, g_hi) from the FP64 history.
// One thread per (group, node).
template <typename T, int NS>
__global__ __launch_bounds__(256) void k_pack_split(const double* __restrict__ D, int64_t ld, int nodes,
                                                    int g_lo, int g_hi, T* __restrict__ c0,
                                                    T* __restrict__ c1, T* __restrict__ c2,
                                                    float* __restrict__ scale) {
    int node = blockIdx.x * 256 + threadIdx.x;
    int g = g_lo + blockIdx.y;
    if (node >= ld || g >= g_hi) return;
    double v[kGroup];
    double mx = 0.0;
#pragma unroll
    for (int e = 0; e < kGroup; e++) {
        v[e] = (node < nodes) ? D[(int64_t)(g * kGroup + e) * ld + node] : 0.0;
        mx = fmax(mx, fabs(v[e]));
    }
    double sc = 1.0;
    if (SplitTraits<T>::scaled) {
        // scaled magnitudes in [2^14, 2^15): hi and lo stay normal fp16 for
        // entries within 2^-10 of the group maximum
        if (mx > 0.0) sc = ldexp(1.0, ilogb(mx) - 14);
        scale[(int64_t)g * ld + node] = (float)sc;
    }
    double inv = 1.0 / sc;
    int64_t base = ((int64_t)g * ld + node) * kGroup;
#pragma unroll
    for (int e = 0; e < kGroup; e++) {
        double x = v[e] * inv;
        T h = (T)(float)x;
        double r = x - (double)(float)h;
        T m = (T)(float)r;
        c0[base + e] = h;
        c1[base + e] = m;
        if (NS == 3) {
            double r2 = r - (double)(float)m;
            c2[base + e] = (T)(float)r2;
        }
    }
}

// A fragments of the split GEMM for the groups [0, tc/32) of one block: the
// weights of rows r < BR against the levels of a group (Toeplitz
// omega[n0+r-t-1], or with DENSE the rows W[r][t] of a prescribed grid),
// split hi/lo(/lo2) into the MFMA A-operand layout (lane l of fragment rt:
// row 16 rt + l%16, levels 8(l/16)..8(l/16)+7 of the group).  They depend on
// the block only, not on the node tile or the K split, so they are formed
// once here instead of in every GEMM workgroup.  Layout: afrag[((g NS + c)
// RT + rt) 64 + lane], ainv[g RT + rt] = 1 / the fragment's power-of-two
// scale (fp16; 1 for bf16).  Grid: one workgroup per group; each wave builds
// whole fragments, so the dense form's fragment maximum is a wave maximum.
template <typename T, int NS, int BR, bool DENSE>
__global__ __launch_bounds__(256) void k_afrag_split(int n0, const double* __restrict__ omega,
                                                     const double* __restrict__ W, int64_t ldw,
                                                     typename SplitTraits<T>::vec* __restrict__ afrag,
                                                     double* __restrict__ ainv) {
    typedef typename SplitTraits<T>::vec vec_t;
    constexpr int RT = BR / 16;
    const int g = blockIdx.x;
    for (int f = threadIdx.x; f < RT * 64; f += 256) {
        const int fl = f & 63, frt = f >> 6;
        const int r = frt * 16 + (fl & 15);
        const int tbase = g * kGroup + 8 * (fl >> 4);
        double wd[8];
#pragma unroll
        for (int e = 0; e < 8; e++) wd[e] = DENSE ? W[(int64_t)r * ldw + tbase + e] : omega[n0 + r - (tbase + e) - 1];
        // fp16: power-of-two scale per fragment so that its largest weight
        // sits in [2^14, 2^15) and the lo parts stay normal.  Toeplitz
        // weights decrease with the lag n0+r-t, so the extremes are at the
        // corners; dense rows take the maximum over the fragment.
        double sa = 1.0;
        if constexpr (SplitTraits<T>::scaled && DENSE) {
            double mx = 0.0;
#pragma unroll
            for (int e = 0; e < 8; e++) mx = fmax(mx, fabs(wd[e]));
            mx = wave_max(mx);
            if (mx > 0.0) sa = ldexp(1.0, 14 - ilogb(mx));
        } else if (SplitTraits<T>::scaled) {
            const int lag_lo = n0 + frt * 16 - (g * kGroup + kGroup - 1);
            const int lag_hi = n0 + frt * 16 + 15 - g * kGroup;
            const double wmax = fmax(fabs(omega[lag_lo - 1]), fabs(omega[lag_hi - 1]));
            if (wmax > 0.0) sa = ldexp(1.0, 14 - ilogb(wmax));
        }
        if (fl == 0) ainv[g * RT + frt] = 1.0 / sa;
        vec_t h, m, l;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const double wv = wd[e] * sa;
            T hh = (T)(float)wv;
            double rr = wv - (double)(float)hh;
            T mm = (T)(float)rr;
            h[e] = hh;
            m[e] = mm;
            if (NS == 3) l[e] = (T)(float)(rr - (double)(float)mm);
        }
        const size_t base = ((size_t)g * NS * RT + frt) * 64 + fl;
        afrag[base] = h;
        afrag[base + RT * 64] = m;
        if (NS == 3) afrag[base + 2 * RT * 64] = l;
    }
}

// Split GEMM.  Grid: x = 64-node tiles, y = K splits.  Per 32-level group:
// wave w (nodes 16w..16w+15) loads its B fragments (one 16-B load per
// component), the A fragments (k_afrag_split) come from LDS, where the block
// stages them for 4 groups at a time, and the products are accumulated in
// fp32 by magnitude class, then flushed to fp64.
template <typename T, int NS, int BR>
__global__ __launch_bounds__(256) void k_gemm_split(const T* __restrict__ c0, const T* __restrict__ c1,
                                                    const T* __restrict__ c2,
                                                    const float* __restrict__ scale, int64_t ld,
                                                    int nodes, int tc, int kchunk,
                                                    const typename SplitTraits<T>::vec* __restrict__ afrag_g,
                                                    const double* __restrict__ ainv_g,
                                                    double* __restrict__ P) {
    typedef typename SplitTraits<T>::vec vec_t;
    constexpr int RT = BR / 16;
    constexpr int GB = 4;                                  // groups per LDS batch
    // afrag[comp][gb][rt][lane] : 8 elements each
    __shared__ vec_t afrag[NS][GB][RT][64];
    __shared__ double ainv[GB][RT];                        // 1 / weight scale of a fragment
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ks = blockIdx.y;
    const int glo = (ks * kchunk) / kGroup;
    const int ghi = min(tc, ks * kchunk + kchunk) / kGroup;
    const int node = blockIdx.x * 64 + wave * 16 + (lane & 15);
    const bool node_ok = node < nodes;
    d4_t acc[RT];
#pragma unroll
    for (int rt = 0; rt < RT; rt++) acc[rt] = (d4_t){0.0, 0.0, 0.0, 0.0};

    // software pipeline: the next batch's B fragments (history) and A
    // fragments (weights) load into registers while this batch computes
    constexpr int AE = GB * RT * 64 / 256;                // A entries per thread per batch
    vec_t bn[GB][NS], an[AE][NS];
    double sn[GB], ain[AE];
    auto load_batch = [&](int gb0) {
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
            const int g = min(gb0 + gi, ghi - 1);
            const int64_t off = ((int64_t)g * ld + (node_ok ? node : 0)) * kGroup + 8 * (lane >> 4);
            bn[gi][0] = *reinterpret_cast<const vec_t*>(c0 + off);
            bn[gi][1] = *reinterpret_cast<const vec_t*>(c1 + off);
            if (NS == 3) bn[gi][NS - 1] = *reinterpret_cast<const vec_t*>(c2 + off);
            sn[gi] = 1.0;
            if (SplitTraits<T>::scaled) sn[gi] = node_ok ? (double)scale[(int64_t)g * ld + node] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < AE; k++) {
            const int f = threadIdx.x + 256 * k;
            const int fl = f & 63, frt = (f >> 6) % RT, fgb = f / (64 * RT);
            const int g = min(gb0 + fgb, ghi - 1);          // groups past ghi are never used
            const size_t base = ((size_t)g * NS * RT + frt) * 64 + fl;
            an[k][0] = afrag_g[base];
            an[k][1] = afrag_g[base + RT * 64];
            if (NS == 3) an[k][NS - 1] = afrag_g[base + 2 * RT * 64];
            ain[k] = ainv_g[g * RT + frt];
        }
    };
    if (glo < ghi) load_batch(glo);
    for (int gb0 = glo; gb0 < ghi; gb0 += GB) {
        vec_t bfr[GB][NS];
        double bsc[GB];
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
#pragma unroll
            for (int c = 0; c < NS; c++) bfr[gi][c] = bn[gi][c];
            bsc[gi] = sn[gi];
        }
        __syncthreads();                       // the previous batch is done with the LDS fragments
#pragma unroll
        for (int k = 0; k < AE; k++) {
            const int f = threadIdx.x + 256 * k;
            const int fl = f & 63, frt = (f >> 6) % RT, fgb = f / (64 * RT);
#pragma unroll
            for (int c = 0; c < NS; c++) afrag[c][fgb][frt][fl] = an[k][c];
            if (fl == 0) ainv[fgb][frt] = ain[k];
        }
        if (gb0 + GB < ghi) load_batch(gb0 + GB);
        __syncthreads();
#pragma unroll
        for (int gi = 0; gi < GB; gi++) {
            int g = gb0 + gi;
            if (g >= ghi) break;
            const vec_t bh = bfr[gi][0], bm = bfr[gi][1];
#pragma unroll
            for (int rt = 0; rt < RT; rt++) {
                vec_t ah = afrag[0][gi][rt][lane];
                vec_t am = afrag[1][gi][rt][lane];
                // one fp32 chain per group: hi*hi first, then the smaller
                // cross products (their rounding is 2^-24 of the group sum)
                f4_t z = {0.f, 0.f, 0.f, 0.f};
                f4_t p;
                if constexpr (SplitTraits<T>::scaled) {
                    p = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, z, 0, 0, 0);
                    p = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bm, p, 0, 0, 0);
                    p = __builtin_amdgcn_mfma_f32_16x16x32_f16(am, bh, p, 0, 0, 0);
                    if (NS == 3) {
                        vec_t al = afrag[NS - 1][gi][rt][lane];
                        vec_t bl = bfr[gi][NS - 1];
                        p = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, p, 0, 0, 0);
                        p = __builtin_amdgcn_mfma_f32_16x16x32_f16(am, bm, p, 0, 0, 0);
                        p = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, p, 0, 0, 0);
                    }
                } else {
                    p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, z, 0, 0, 0);
                    p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, p, 0, 0, 0);
                    p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, p, 0, 0, 0);
                    if (NS == 3) {
                        vec_t al = afrag[NS - 1][gi][rt][lane];
                        vec_t bl = bfr[gi][NS - 1];
                        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, p, 0, 0, 0);
                        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, p, 0, 0, 0);
                        p = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, p, 0, 0, 0);
                    }
                }
                double scr = SplitTraits<T>::scaled ? bsc[gi] * ainv[gi][rt] : 1.0;
#pragma unroll
                for (int v = 0; v < 4; v++) acc[rt][v] = fma(scr, (double)p[v], acc[rt][v]);
            }
        }
    }
    if (node_ok) {
#pragma unroll
        for (int rt = 0; rt < RT; rt++) {
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int r = rt * 16 + 4 * (lane >> 4) + v;
                P[((int64_t)ks * BR + r) * ld + node] = acc[rt][v];
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Fold the K-split partial rows of a GEMM block into split 0, in place and in
// split order: P[0][r][i] = ((P[0] + P[1]) + P[2]) + ..., so a 1D step reads
// one far value per node.  The splits of one (r, i) load as one batch of
// independent loads.  (Folding inside the GEMM by its last-arriving
// workgroup per tile needs an agent-scope release per workgroup, an L2
// write-back on gfx950, and measured slower than this pass.)
__global__ __launch_bounds__(256) void k_part_reduce(double* __restrict__ P, int splits, int rows,
                                                     int64_t ld, int nodes) {
    constexpr int W = 16;
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= nodes || r >= rows) return;
    double s = P[(int64_t)r * ld + i];
    for (int k0 = 1; k0 < splits; k0 += W) {
        double v[W];
#pragma unroll
        for (int j = 0; j < W; j++) v[j] = (k0 + j < splits) ? P[((int64_t)(k0 + j) * rows + r) * ld + i] : 0.0;
#pragma unroll
        for (int j = 0; j < W; j++)
            if (k0 + j < splits) s += v[j];
    }
    P[(int64_t)r * ld + i] = s;
}

fpde_status launch_part_reduce(double* P, int splits, int rows, int64_t ld, int nodes, hipStream_t st) {
    if (splits <= 1 || nodes <= 0) return FPDE_OK;
    dim3 grid((unsigned)((nodes + 255) / 256), (unsigned)rows);
    hipLaunchKernelGGL(k_part_reduce, grid, dim3(256), 0, st, P, splits, rows, ld, nodes);
    return FPDE_OK;
}

// ---------------------------------------------------------------------------
// Host launchers
// t-split plan of the plain stream: enough workgroups to fill the chip
// (~4096 with the node tiles), each streaming >= 32 levels; at most
// kPlainMaxSplits partial rows.  Returns the number of partial rows.
int hist_plain_plan(int nodes, int t_hi, int* chunk_out) {
    if (t_hi <= 0) { if (chunk_out) *chunk_out = 0; return 0; }
    int tiles = (nodes + 511) / 512;
    int want = (4096 + tiles - 1) / tiles;
    if (want > kPlainMaxSplits) want = kPlainMaxSplits;
    if (want < 1) want = 1;
    int chunk = (t_hi + want - 1) / want;
    if (chunk < 32) chunk = 32;
    if (chunk_out) *chunk_out = chunk;
    return (t_hi + chunk - 1) / chunk;
}

fpde_status launch_hist_plain(const double* D, int64_t ld, int nodes, int n, int t_hi,
                              const double* omega, double* part, int* splits_out,
                              hipStream_t st) {
    int chunk = 0;
    const int splits = hist_plain_plan(nodes, t_hi, &chunk);
    *splits_out = splits;
    if (splits == 0) return FPDE_OK;
    int tiles = (nodes + 511) / 512;
    dim3 grid(tiles, splits);
    hipLaunchKernelGGL(k_hist_plain, grid, dim3(256), 0, st, D, ld, nodes, n, t_hi, chunk, omega, part);
    return FPDE_OK;
}

fpde_status launch_ref_sum(double* S, const double* old, int64_t old_size, int64_t os,
                           int64_t nodes, const double* tau, const double* kb, int sum_alg,
                           double sum_param, int bs, hipStream_t st) {
    if (nodes <= 0) return FPDE_OK;
    dim3 grid((unsigned)((nodes + 255) / 256));
    if (bs < 1) bs = 1;
    hipLaunchKernelGGL(k_ref_sum, grid, dim3(256), 0, st, S, old, old_size, os, nodes, tau, kb,
                       sum_alg, sum_param, bs);
    return FPDE_OK;
}

// K splits for a GEMM over t < tc: enough workgroups to cover the chip
// (>= ~2048 with the node tiles), at most 16 partial rows per output.
int gemm_max_splits(int64_t nodes) {
    int64_t tiles = (nodes + 63) / 64;
    int64_t s = (2048 + tiles - 1) / tiles;
    return (int)(s < 1 ? 1 : (s > 16 ? 16 : s));
}

static void gemm_split_plan(int tc, int64_t nodes, int* kchunk, int* splits, int max_splits) {
    int ks = gemm_max_splits(nodes);
    if (max_splits > 0 && ks > max_splits) ks = max_splits;
    int ch = (tc + ks - 1) / ks;
    ch = ((ch + 31) / 32) * 32;
    if (ch < 32) ch = 32;
    *kchunk = ch;
    *splits = (tc + ch - 1) / ch;
}

fpde_status launch_gemm_f64_dense(const double* D, int64_t ld, int nodes, int tc, int block, const double* W,
                                  int64_t ldw, double* P, int* splits_out, hipStream_t st, int max_splits) {
    if (tc <= 0) { *splits_out = 0; return FPDE_OK; }
    int kchunk, splits;
    gemm_split_plan(tc, nodes, &kchunk, &splits, max_splits);
    dim3 grid((unsigned)((nodes + 63) / 64), (unsigned)splits);
    switch (block) {
        case 16: hipLaunchKernelGGL(k_gemm_f64_dense<16>, grid, dim3(256), 0, st, D, ld, nodes, tc, kchunk, W, ldw, P); break;
        case 32: hipLaunchKernelGGL(k_gemm_f64_dense<32>, grid, dim3(256), 0, st, D, ld, nodes, tc, kchunk, W, ldw, P); break;
        case 64: hipLaunchKernelGGL(k_gemm_f64_dense<64>, grid, dim3(256), 0, st, D, ld, nodes, tc, kchunk, W, ldw, P); break;
        case 128: hipLaunchKernelGGL(k_gemm_f64_dense<128>, grid, dim3(256), 0, st, D, ld, nodes, tc, kchunk, W, ldw, P); break;
        default: set_error("block must be 16, 32, 64 or 128"); return FPDE_ERR_ARG;
    }
    *splits_out = splits;
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status launch_gemm_f64(const double* D, int64_t ld, int nodes, int n0, int tc, int block,
                            const double* omega, double* P, int* splits_out, hipStream_t st, int max_splits) {
    if (tc <= 0) { *splits_out = 0; return FPDE_OK; }
    int kchunk, splits;
    gemm_split_plan(tc, nodes, &kchunk, &splits, max_splits);
    dim3 grid((nodes + 63) / 64, splits);
    size_t lds = 0;
    switch (block) {
        case 16: hipLaunchKernelGGL(k_gemm_f64<16>, grid, dim3(256), lds, st, D, ld, nodes, n0, tc, kchunk, omega, P); break;
        case 32: hipLaunchKernelGGL(k_gemm_f64<32>, grid, dim3(256), lds, st, D, ld, nodes, n0, tc, kchunk, omega, P); break;
        case 64: hipLaunchKernelGGL(k_gemm_f64<64>, grid, dim3(256), lds, st, D, ld, nodes, n0, tc, kchunk, omega, P); break;
        case 128: hipLaunchKernelGGL(k_gemm_f64<128>, grid, dim3(256), lds, st, D, ld, nodes, n0, tc, kchunk, omega, P); break;
        default: set_error("block must be 16, 32, 64 or 128"); return FPDE_ERR_ARG;
    }
    *splits_out = splits;
    return FPDE_OK;
}

template <typename T, int NS>
static fpde_status pack_t(const double* D, int64_t ld, int nodes, int g_lo, int g_hi, void* c0,
                          void* c1, void* c2, float* scale, hipStream_t st) {
    dim3 grid((unsigned)((ld + 255) / 256), g_hi - g_lo);
    hipLaunchKernelGGL((k_pack_split<T, NS>), grid, dim3(256), 0, st, D, ld, nodes, g_lo, g_hi, (T*)c0,
                       (T*)c1, (T*)c2, scale);
    return FPDE_OK;
}

fpde_status launch_pack_split(int mode, const double* D, int64_t ld, int nodes, int g_lo, int g_hi,
                              void* c0, void* c1, void* c2, float* scale, hipStream_t st) {
    if (g_hi <= g_lo) return FPDE_OK;
    switch (mode) {
        case HIST_SPLIT_BF16: return pack_t<__bf16, 2>(D, ld, nodes, g_lo, g_hi, c0, c1, c2, scale, st);
        case HIST_SPLIT_F16: return pack_t<_Float16, 2>(D, ld, nodes, g_lo, g_hi, c0, c1, c2, scale, st);
        case HIST_SPLIT_BF16X3: return pack_t<__bf16, 3>(D, ld, nodes, g_lo, g_hi, c0, c1, c2, scale, st);
        default: set_error("bad split mode"); return FPDE_ERR_ARG;
    }
}

template <typename T, int NS, int BR>
static void gemm_split_t(dim3 grid, hipStream_t st, const void* c0, const void* c1, const void* c2,
                         const float* scale, int64_t ld, int nodes, int n0, int tc, int kchunk,
                         const double* omega, const double* W, int64_t ldw, void* afrag, double* ainv, double* P) {
    typedef typename SplitTraits<T>::vec vec_t;
    if (W)
        hipLaunchKernelGGL((k_afrag_split<T, NS, BR, true>), dim3(tc / kGroup), dim3(256), 0, st, n0, omega, W, ldw,
                           (vec_t*)afrag, ainv);
    else
        hipLaunchKernelGGL((k_afrag_split<T, NS, BR, false>), dim3(tc / kGroup), dim3(256), 0, st, n0, omega, W, ldw,
                           (vec_t*)afrag, ainv);
    hipLaunchKernelGGL((k_gemm_split<T, NS, BR>), grid, dim3(256), 0, st, (const T*)c0, (const T*)c1, (const T*)c2,
                       scale, ld, nodes, tc, kchunk, (const vec_t*)afrag, (const double*)ainv, P);
}

template <typename T, int NS>
static fpde_status gemm_split_b(int block, dim3 grid, hipStream_t st, const void* c0, const void* c1,
                                const void* c2, const float* scale, int64_t ld, int nodes, int n0,
                                int tc, int kchunk, const double* omega, const double* W, int64_t ldw,
                                void* afrag, double* ainv, double* P) {
    switch (block) {
        case 16: gemm_split_t<T, NS, 16>(grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        case 32: gemm_split_t<T, NS, 32>(grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        case 64: gemm_split_t<T, NS, 64>(grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        case 128: gemm_split_t<T, NS, 128>(grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        default: set_error("block must be 16, 32, 64 or 128"); return FPDE_ERR_ARG;
    }
    return FPDE_OK;
}

size_t split_afrag_bytes(int mode, int block, int groups) {
    const size_t ns = (mode == HIST_SPLIT_BF16X3) ? 3 : 2;
    return (size_t)groups * ns * (size_t)(block / 16) * 64 * 16;
}

fpde_status launch_gemm_split(int mode, const void* c0, const void* c1, const void* c2,
                              const float* scale, int64_t ld, int nodes, int n0, int tc, int block,
                              const double* omega, const double* W, int64_t ldw, void* afrag, double* ainv,
                              double* P, int* splits_out, hipStream_t st, int max_splits) {
    if (tc <= 0) { *splits_out = 0; return FPDE_OK; }
    if (tc % kGroup != 0) { set_error("split GEMM needs a group-aligned cut"); return FPDE_ERR_ARG; }
    if (W && ldw < tc) { set_error("split GEMM: dense weight rows shorter than the cut"); return FPDE_ERR_ARG; }
    if (!afrag || !ainv) { set_error("split GEMM: no fragment workspace"); return FPDE_ERR_ARG; }
    int kchunk, splits;
    gemm_split_plan(tc, nodes, &kchunk, &splits, max_splits);
    dim3 grid((nodes + 63) / 64, splits);
    fpde_status rc;
    switch (mode) {
        case HIST_SPLIT_BF16: rc = gemm_split_b<__bf16, 2>(block, grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        case HIST_SPLIT_F16: rc = gemm_split_b<_Float16, 2>(block, grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        case HIST_SPLIT_BF16X3: rc = gemm_split_b<__bf16, 3>(block, grid, st, c0, c1, c2, scale, ld, nodes, n0, tc, kchunk, omega, W, ldw, afrag, ainv, P); break;
        default: set_error("bad split mode"); return FPDE_ERR_ARG;
    }
    *splits_out = splits;
    return rc;
}

}  // namespace fpde
