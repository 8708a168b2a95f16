// Host side of the solver: device state of one solve, the per-step launch
// plan (history kernels + step kernels), optional hipGraph capture of the
// step loop, profiling, and the C ABI (include/fpde_hip.h, group 3).
//
// The step loop mirrors H_solver::calc_step (fpde_hpc_1d.cpp:5100-5803):
// init_time_step (weights; here generated once, on the device) -> history sum
// (time_summing / fr_ocl_call_sum) -> tridiagonal step(s) -> push the new level.
#include "fpde_internal.h"

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <vector>

using namespace fpde;

namespace {

thread_local std::string g_err;

// Per-device argument tables of the batched step launches: allocated on a
// device's first batch, freed by fpde_shutdown (a hipFree per batch would
// wait for every stream of the device).  One batch at a time uses a table.
std::mutex g_rows_mu;
std::map<int, Step1DArgs*> g_rows_tab;

int64_t round_up(int64_t v, int64_t m) { return (v + m - 1) / m * m; }

bool is_split(int mode) { return mode == HIST_SPLIT_BF16 || mode == HIST_SPLIT_F16 || mode == HIST_SPLIT_BF16X3; }
int n_comp(int mode) { return mode == HIST_SPLIT_BF16X3 ? 3 : 2; }

struct Plan {
    int64_t nodes = 0;     // unknowns per level
    int64_t ld = 0;        // padded level stride
    int64_t levels = 0;    // history levels kept
    int part_rows = 0;     // rows of one partial-sum buffer
    int part_bufs = 1;     // 2: one per block parity (far part computed one block ahead)
};

// 1D solves whose far part can be computed one history block ahead, beside
// the current block's cooperative step launch: a GEMM history form over the
// full history on the uniform Toeplitz grid (not restricted or series
// summing, no prescribed grid or generalized kernel, no space-fractional
// term, cooperative launches requested).
bool lookahead_spec(const fpde_spec& s) {
    return s.dims == 1 && s.step_mode == 0 && s.hist_mode != HIST_PLAIN_F64 && s.sum_alg == 0 && !s.taus &&
           !s.var_tau && s.fk == 0 && !(s.gamma > 0.0 && s.gamma < 1.0);
}

// Chunks of a 1D line (one workgroup each in the partitioned solve).  The
// automatic size is 192 nodes: with three waves of rows per workgroup the
// chunk PCR levels and the near-field history loads cost less per workgroup,
// and the reduced system (2C rows) stays within seven PCR levels up to
// N = 12288; the headline solve (N = 10^4) measured 0.387 s at 160-192
// nodes against 0.409 s at 256.  Where 192-node chunks would number more
// than 256, 256-node ones are taken.
int chunk_count(const fpde_spec& s) {
    if (s.chunk_rows > 0) return (s.nodes + s.chunk_rows - 1) / s.chunk_rows;
    const int c = (s.nodes + 191) / 192;
    return c <= kMaxChunks ? c : (s.nodes + 255) / 256;
}

fpde_status validate(const fpde_spec& s) {
    if (s.dims != 1 && s.dims != 3) { set_error("dims must be 1 or 3"); return FPDE_ERR_ARG; }
    if (s.steps < 1) { set_error("steps must be >= 1"); return FPDE_ERR_ARG; }
    if (s.dims == 1 && (s.nodes < 4 || s.nodes > 65536)) { set_error("1D nodes must be in [4, 65536]"); return FPDE_ERR_ARG; }
    if (s.dims == 3 && (s.nodes < 4 || s.nodes > 256)) { set_error("3D nodes per axis must be in [4, 256]"); return FPDE_ERR_ARG; }
    if (!(s.tau > 0.0) || !(s.length > 0.0)) { set_error("tau and length must be positive"); return FPDE_ERR_ARG; }
    if (s.do_q <= 0 && !(s.alpha > 0.0 && s.alpha <= 1.0)) { set_error("alpha must be in (0,1]"); return FPDE_ERR_ARG; }
    if (s.picard_max < 1 || s.picard_max > 16) { set_error("picard_max must be in [1,16]"); return FPDE_ERR_ARG; }
    if (s.hist_mode < 0 || s.hist_mode > 4) { set_error("hist_mode must be in [0,4]"); return FPDE_ERR_ARG; }
    if (s.hist_mode != HIST_PLAIN_F64) {
        int b = s.block;
        if (b != 16 && b != 32 && b != 64 && b != 128) { set_error("block must be 16, 32, 64 or 128"); return FPDE_ERR_ARG; }
    }
    if (s.step_mode != 0 && s.step_mode != 1) { set_error("step_mode must be 0 or 1"); return FPDE_ERR_ARG; }
    if (!(s.gamma >= 0.0 && s.gamma <= 1.0)) { set_error("gamma must be in [0,1] (0 or 1: classical)"); return FPDE_ERR_ARG; }

    if (s.dims == 3 && s.picard_max != 1) { set_error("3D LOD uses lagged coefficients (picard_max must be 1)"); return FPDE_ERR_ARG; }
    if (s.sum_alg < 0 || s.sum_alg > 2) {
        set_error("sum_alg must be 0 (full sum), 1 (restricted sum) or 2 (series moments)");
        return FPDE_ERR_ARG;
    }
    if (s.sum_alg == 1 && !(s.sum_param >= 0.0 && std::isfinite(s.sum_param))) {
        set_error("sum_param must be finite and >= 0 for the restricted sum");
        return FPDE_ERR_ARG;
    }
    if (s.var_tau) {
        if (s.var_tau != 1 || s.dims != 1 || s.hist_mode != HIST_PLAIN_F64 || s.sum_alg != 0 || s.taus) {
            set_error("var_tau solvers are 1D, hist_mode 0, sum_alg 0, without a prescribed grid");
            return FPDE_ERR_ARG;
        }
    }
    if (s.chunk_rows != 0 && (s.chunk_rows < 32 || s.chunk_rows > 256)) {
        set_error("chunk_rows must be 0 (automatic) or in [32, 256]");
        return FPDE_ERR_ARG;
    }
    if (s.dims == 1 && chunk_count(s) > kMaxChunks) {
        set_error("chunk_rows too small for this line: at most 256 chunks");
        return FPDE_ERR_ARG;
    }
    if (s.fk < 0 || s.fk > 3) { set_error("fk must be 0 (g(t)=t), 1 (sqrt t), 2 (t^2) or 3 (t^fp)"); return FPDE_ERR_ARG; }
    if (s.fk == 3 && !(s.fp >= 0.05 && s.fp <= 20.0)) { set_error("fk 3: fp must be in [0.05, 20]"); return FPDE_ERR_ARG; }
    if (s.fk != 0) {
        // generalized kernel: per-step weight rows (the grid path), exact integrals
        if (s.weight_mode != 0) {
            set_error("a generalized kernel (fk != 0) takes the exact weight integrals: weight_mode must be 0");
            return FPDE_ERR_ARG;
        }
        if (s.sum_alg != 0) { set_error("a generalized kernel (fk != 0) needs sum_alg 0"); return FPDE_ERR_ARG; }
    }
    if (s.taus) {
        if (s.dims != 1) { set_error("a prescribed time grid (taus) is for 1D solves"); return FPDE_ERR_ARG; }
        if (s.sum_alg != 0) { set_error("a prescribed time grid needs sum_alg 0"); return FPDE_ERR_ARG; }
        for (int k = 0; k < s.steps; k++)
            if (!(s.taus[k] > 0.0 && std::isfinite(s.taus[k]))) { set_error("taus must be positive and finite"); return FPDE_ERR_ARG; }
    }
    if (s.sum_alg == 2) {
        if (!(s.sum_param >= 1.0 && s.sum_param <= kMaxSeriesTerms && s.sum_param == std::floor(s.sum_param))) {
            set_error("sum_alg 2: sum_param is the number of series terms, an integer in [1, 64]");
            return FPDE_ERR_ARG;
        }
        if (s.do_q > 0) { set_error("sum_alg 2 is for a single order (do_q must be 0)"); return FPDE_ERR_ARG; }
        if (s.hist_mode != HIST_PLAIN_F64) {
            set_error("sum_alg 2 keeps no history levels: hist_mode must be 0 (plain)");
            return FPDE_ERR_ARG;
        }
    }
    return FPDE_OK;
}

Plan make_plan(const fpde_spec& s) {
    Plan p;
    p.nodes = (s.dims == 1) ? s.nodes : (int64_t)s.nodes * s.nodes * s.nodes;
    p.ld = round_up(p.nodes, 64);
    // sum_alg 2 keeps moments instead of levels: D is a ring of the last two
    p.levels = (s.sum_alg == 2) ? 2 : round_up(s.steps + 1, kGroup);
    int rows_plain = 64;
    int rows_gemm = (s.hist_mode == HIST_PLAIN_F64) ? 0 : gemm_max_splits(p.nodes) * s.block;
    p.part_rows = (s.sum_alg == 2) ? 1 : std::max(rows_plain, rows_gemm);
    if (lookahead_spec(s)) {
        // the older part's K splits, folded into the first B rows, then the
        // newer part's single split in the next B rows
        p.part_rows = std::max(p.part_rows, 2 * s.block);
        p.part_bufs = 2;
    }
    return p;
}

size_t footprint_bytes(const fpde_spec& s) {
    Plan p = make_plan(s);
    size_t b = 0;
    b += sizeof(double) * (size_t)p.levels * p.ld;               // FP64 history
    if (is_split(s.hist_mode)) {
        b += (size_t)n_comp(s.hist_mode) * 2 * p.levels * p.ld + sizeof(float) * (p.levels / kGroup) * p.ld;
        b += split_afrag_bytes(s.hist_mode, s.block, (int)(p.levels / kGroup)) +
             sizeof(double) * (size_t)(p.levels / kGroup) * (s.block / 16);
    }
    b += sizeof(double) * (size_t)p.part_rows * p.part_bufs * p.ld;   // partial sums
    b += sizeof(double) * (size_t)p.ld * 8;                      // U, S, xyz, ...
    if (s.hist_mode != HIST_PLAIN_F64 && s.sum_alg != 2)
        b += sizeof(double) * (size_t)s.block * p.ld;                 // folded near-field groups
    if (s.dims == 3) b += sizeof(double) * (size_t)p.ld * 6;     // Uprev, K, iC, Al, Bt
    b += sizeof(double) * (size_t)(s.steps + 512) * 3;           // weights
    if (s.sum_alg == 2) {
        size_t P = (size_t)s.sum_param;
        b += sizeof(double) * P * p.ld;                              // moments
        b += sizeof(double) * 2 * P * (size_t)s.steps;               // coefficient tables
    }
    if (s.gamma > 0.0 && s.gamma < 1.0) {                            // space-fractional sum
        b += sizeof(double) * (3 * (size_t)s.nodes + p.ld);
        size_t ldp = (size_t)round_up(s.nodes, kSpTile);
        if (s.dims == 1) b += sizeof(double) * ldp * (1 + space_chunks((int)(ldp / kSpTile) - 1));
    }
    if (s.taus || s.var_tau || s.fk) {                               // grid, weight rows
        size_t M = (size_t)s.steps, B = (size_t)s.block;
        b += sizeof(double) * (3 * M + 64 + B * (round_up(M + 64, 64) + B + 64));
    }
    return b;
}

}  // namespace

namespace fpde {
void set_error(const std::string& msg) { g_err = msg; }
}  // namespace fpde

struct fpde_solver {
    fpde_spec spec{};
    Plan plan;
    hipStream_t stream = nullptr;
    DeviceWeights w;
    double* U = nullptr;
    double* D = nullptr;
    double* S = nullptr;
    double* xyz = nullptr;
    double* red = nullptr;
    double* part = nullptr;
    double* Uit = nullptr;
    double* dmax = nullptr;
    double* mid = nullptr;     // near-field groups folded ahead [block][ld] (1D cooperative launches, 3D combine)
    double* H = nullptr;       // sum_alg 2 moments [P][ld]
    double* Uit0 = nullptr;    // initial Picard iterate
    double* uend = nullptr;    // chunk-end values by step parity
    // prescribed grid: host copy of the step lengths, weight rows of the
    // current step (plain) or block (far: direct, near: reversed per step)
    std::vector<double> taus_h;
    std::vector<double> Time_h;          // var_tau: the grid so far
    double* Wrow = nullptr;
    double* Wfar = nullptr;
    double* Wnear = nullptr;
    int64_t ldwf = 0, ldwn = 0;
    const double* omega_cur = nullptr;   // weights of the step being enqueued
    // space-fractional scheme: F of U^n [sp_ldp], partial sums [chunks][sp_ldp], E [ld]
    double* spF = nullptr;
    double* spP = nullptr;
    double* spE = nullptr;
    int64_t sp_ldp = 0;
    unsigned long long* xch = nullptr;  // grid exchange words of the cooperative 1D block launches
    int* err = nullptr;        // grid-exchange time-out flag, Picard solve count
    bool block_mode = false;   // 1D steps run as cooperative block launches
    int64_t coop_launches = 0; // cooperative launches so far (the next one's exchange tag)
    int* flags = nullptr;
    // graph-captured step loop: the executable graph of the last captured chunk
    hipGraphExec_t gexec = nullptr;
    int64_t graph_launches = 0, graph_instantiations = 0;
    void* comp[3] = {nullptr, nullptr, nullptr};
    float* scale = nullptr;
    void* afrag = nullptr;       // A fragments of a split GEMM block
    double* ainv = nullptr;
    // 3D
    double* Uprev = nullptr;
    double* utop = nullptr;
    double* K = nullptr;
    double* iC = nullptr;
    double* Al = nullptr;
    double* Bt = nullptr;
    double h = 0.0, utop1d = 0.0;
    int C = 0;
    // 3D: host copy of the initial field (the halo slots of the results3d dump)
    std::vector<double> u0_h;
    // progress
    int step = 0;
    int64_t launches = 0;
    int last_k = -1;
    bool pending = false;
    bool initialized = false;
    // GEMM block state
    int n0 = -1, tc = 0, gsplits = 0, packed_groups = 0;
    double* part_cur = nullptr;  // partial rows the current block's steps read
    // far part one block ahead (1D cooperative form, lookahead_spec): the
    // older levels [0, tc of this block) of the next block's GEMM run on
    // `side` during this block's step launch, into the other parity's
    // buffer; the next block adds only its newer levels on the main stream
    bool la = false;
    hipStream_t side = nullptr;
    hipEvent_t ev_far = nullptr;   // main: this block's far part is enqueued
    hipEvent_t ev_side = nullptr;  // side: the next block's older part is done
    int la_n0 = -1;                // the block the pending older part is for
    int la_tc = 0;                 // its cut (levels [0, la_tc))
    int la_par = 0;                // parity of the current block's buffer
    // profiling
    int prof = 0;
    struct EvPair { hipEvent_t a, b; int cls; };
    std::vector<EvPair> evs;
    size_t ev_used = 0;
    fpde_stats stats{};
    int picard_dev_base = 0;      // err[1] (Picard solves of cooperative launches) at the last reset
    std::vector<void*> allocs;
    // server-owned memory and stream (solver_create_in); else its own
    DeviceArena* arena = nullptr;
    bool own_stream = true;
    int64_t wait_ticks = 0;      // grid-barrier bound of the cooperative launches
};

namespace {

fpde_status dalloc(fpde_solver* s, void** p, size_t bytes) {
    if (bytes == 0) bytes = 8;
    if (s->arena) {
        *p = s->arena->alloc(bytes);
        if (!*p) { set_error("device arena: no free block of that size within the budget"); return FPDE_ERR_NOMEM; }
    } else {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            set_error(std::string("hipMalloc failed: ") + hipGetErrorString(e));
            *p = nullptr;
            return FPDE_ERR_NOMEM;
        }
    }
    s->allocs.push_back(*p);
    s->stats.device_bytes += bytes;
    return FPDE_OK;
}
#define DALLOC(ptr, bytes)                                              \
    do {                                                                \
        fpde_status rc_ = dalloc(s, (void**)&(ptr), (bytes));           \
        if (rc_ != FPDE_OK) return rc_;                                 \
    } while (0)

// profiling event pairs: cls 0 = history kernels, 1 = step kernels
void ev_begin(fpde_solver* s, int cls, size_t* slot) {
    *slot = (size_t)-1;
    if (!(s->prof & (1 << cls))) return;
    if (s->ev_used == s->evs.size()) {
        fpde_solver::EvPair p;
        hipEventCreate(&p.a);
        hipEventCreate(&p.b);
        s->evs.push_back(p);
    }
    *slot = s->ev_used++;
    s->evs[*slot].cls = cls;
    hipEventRecord(s->evs[*slot].a, s->stream);
}
void ev_end(fpde_solver* s, size_t slot) {
    if (slot == (size_t)-1) return;
    hipEventRecord(s->evs[slot].b, s->stream);
}
void ev_collect(fpde_solver* s) {
    if (s->side) hipStreamSynchronize(s->side);    // far parts computed ahead record there
    for (size_t i = 0; i < s->ev_used; i++) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, s->evs[i].a, s->evs[i].b);
        if (s->evs[i].cls == 0) s->stats.ms_history += ms;
        else s->stats.ms_step += ms;
    }
    s->ev_used = 0;
}

fpde_status solver_alloc(fpde_solver* s) {
    const fpde_spec& sp = s->spec;
    const Plan& p = s->plan;
    if (!s->stream) FPDE_HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    int M = sp.steps;
    s->w.M = M;
    if (sp.var_tau) {
        // the grid is filled step by step (fpde_solver_advance_tau)
        s->w.grid = true;
        s->Time_h.assign((size_t)M + 2, 0.0);
        s->taus_h.assign((size_t)M + 1, 0.0);
        DALLOC(s->w.Time, sizeof(double) * (M + 2));
        DALLOC(s->w.taus, sizeof(double) * (M + 1));
        DALLOC(s->w.Lv, sizeof(double) * M);
        FPDE_HIP_TRY(hipMemsetAsync(s->w.Time, 0, sizeof(double) * (M + 2), s->stream));
        FPDE_HIP_TRY(hipMemsetAsync(s->w.taus, 0, sizeof(double) * (M + 1), s->stream));
        FPDE_HIP_TRY(hipMemsetAsync(s->w.Lv, 0, sizeof(double) * M, s->stream));
        DALLOC(s->Wrow, sizeof(double) * (size_t)(M + 64));
    }
    if (sp.taus || (sp.fk != 0 && !sp.var_tau)) {
        // a grid known at create: prescribed step lengths, or uniform ones
        // with a generalized kernel (its weights are not Toeplitz either)
        s->w.grid = true;
        std::vector<double>& T = s->w.hTime;
        std::vector<double> tt((size_t)M + 1);
        T.assign((size_t)M + 2, 0.0);
        for (int k = 0; k <= M; k++) tt[k] = sp.taus ? sp.taus[k < M ? k : M - 1] : sp.tau;
        for (int k = 1; k <= M + 1; k++) T[k] = T[k - 1] + tt[k - 1];       // init_time_step (:3944)
        DALLOC(s->w.Time, sizeof(double) * (M + 2));
        DALLOC(s->w.taus, sizeof(double) * (M + 1));
        DALLOC(s->w.Lv, sizeof(double) * M);
        FPDE_HIP_TRY(hipMemcpyAsync(s->w.Time, T.data(), sizeof(double) * (M + 2), hipMemcpyHostToDevice, s->stream));
        FPDE_HIP_TRY(hipMemcpyAsync(s->w.taus, tt.data(), sizeof(double) * (M + 1), hipMemcpyHostToDevice, s->stream));
        DALLOC(s->Wrow, sizeof(double) * (size_t)(M + 64));
        if (sp.hist_mode != HIST_PLAIN_F64) {      // dense GEMM rows: FP64 MFMA and split forms
            s->ldwf = round_up(M + 64, 64);
            s->ldwn = sp.block + 64;
            DALLOC(s->Wfar, sizeof(double) * (size_t)sp.block * s->ldwf);
            DALLOC(s->Wnear, sizeof(double) * (size_t)sp.block * s->ldwn);
        }
        FPDE_HIP_TRY(hipStreamSynchronize(s->stream));    // the host vector tt goes out of scope
    }
    DALLOC(s->w.omega, sizeof(double) * (size_t)(M + 512));
    DALLOC(s->w.corr, sizeof(double) * 2 * (size_t)(M + 1));
    FPDE_HIP_TRY(hipMemsetAsync(s->w.omega, 0, sizeof(double) * (size_t)(M + 512), s->stream));
    if (sp.sum_alg == 2) {
        int P = (int)sp.sum_param;
        s->w.ser_P = P;
        DALLOC(s->w.ser_c, sizeof(double) * (size_t)M * P);
        DALLOC(s->w.ser_v, sizeof(double) * (size_t)M * P);
        DALLOC(s->H, sizeof(double) * (size_t)P * p.ld);
    }
    DALLOC(s->U, sizeof(double) * p.ld);
    DALLOC(s->D, sizeof(double) * (size_t)p.levels * p.ld);
    FPDE_HIP_TRY(hipMemsetAsync(s->D, 0, sizeof(double) * (size_t)p.levels * p.ld, s->stream));
    DALLOC(s->S, sizeof(double) * p.ld);
    DALLOC(s->part, sizeof(double) * (size_t)p.part_rows * p.part_bufs * p.ld);
    s->part_cur = s->part;
    if (is_split(sp.hist_mode)) {
        for (int c = 0; c < n_comp(sp.hist_mode); c++) {
            DALLOC(s->comp[c], 2 * (size_t)p.levels * p.ld);
        }
        DALLOC(s->scale, sizeof(float) * (size_t)(p.levels / kGroup) * p.ld);
        DALLOC(s->afrag, split_afrag_bytes(sp.hist_mode, sp.block, (int)(p.levels / kGroup)));
        DALLOC(s->ainv, sizeof(double) * (size_t)(p.levels / kGroup) * (sp.block / 16));
    }
    if (sp.dims == 1) {
        s->C = chunk_count(sp);
        DALLOC(s->xyz, sizeof(double) * 3 * p.ld);
        DALLOC(s->red, sizeof(double) * 2 * 2 * s->C * 4);
        DALLOC(s->Uit, sizeof(double) * p.ld);
        DALLOC(s->Uit0, sizeof(double) * p.ld);
        DALLOC(s->uend, sizeof(double) * 2 * 2 * s->C);
        DALLOC(s->xch, kXchBytes);
        FPDE_HIP_TRY(hipMemsetAsync(s->xch, 0, kXchBytes, s->stream));
        DALLOC(s->err, 64);
        FPDE_HIP_TRY(hipMemsetAsync(s->err, 0, 64, s->stream));
        s->wait_ticks = barrier_wait_ticks();
        // tests of the time-out path force a bound of a few ticks: every
        // workgroup then ends at its first unmet wait and the solve reports it
        if (const char* e = getenv("FPDE_TEST_WAIT_TICKS")) {
            const long long v = atoll(e);
            if (v > 0) s->wait_ticks = v;
        }
        // cooperative block launches pay off where a launch covers many steps:
        // the GEMM history forms (one launch per block) and the series form;
        // the plain form needs its history kernel between every two steps
        bool blocked_history = sp.hist_mode != HIST_PLAIN_F64 || sp.sum_alg == 2;
        const bool space = sp.gamma > 0.0 && sp.gamma < 1.0;
        if (space) {
            s->sp_ldp = round_up(sp.nodes, kSpTile);
            DALLOC(s->w.sp_omega, sizeof(double) * (size_t)sp.nodes);
            DALLOC(s->w.sp_corr, sizeof(double) * 2 * (size_t)sp.nodes);
            DALLOC(s->spF, sizeof(double) * (size_t)s->sp_ldp);
            DALLOC(s->spP, sizeof(double) * (size_t)s->sp_ldp * space_chunks((int)(s->sp_ldp / kSpTile) - 1));
            DALLOC(s->spE, sizeof(double) * p.ld);
        }
        // the space-fractional sum needs the whole U^n between two steps: its
        // kernels run between the per-iteration launches (a cooperative form
        // with helper workgroups for the sum measured no faster)
        s->block_mode = (sp.step_mode == 0) && blocked_history && !space && step1d_block_supported(s->C);
        s->la = s->block_mode && lookahead_spec(sp);
        if (s->block_mode && s->w.ser_P == 0) DALLOC(s->mid, sizeof(double) * (size_t)sp.block * p.ld);
        DALLOC(s->dmax, sizeof(double) * 2 * s->C);
        DALLOC(s->flags, sizeof(int) * 16);
    } else {
        int n = sp.nodes;
        if (sp.gamma > 0.0 && sp.gamma < 1.0) {
            DALLOC(s->w.sp_omega, sizeof(double) * (size_t)n);
            DALLOC(s->w.sp_corr, sizeof(double) * 2 * (size_t)n);
            DALLOC(s->spE, sizeof(double) * p.ld);
        }
        DALLOC(s->Uprev, sizeof(double) * p.ld);
        DALLOC(s->utop, sizeof(double) * (size_t)n * n);
        DALLOC(s->K, sizeof(double) * p.ld);
        DALLOC(s->iC, sizeof(double) * p.ld);
        DALLOC(s->Al, sizeof(double) * (size_t)(n + 1) * n * n);
        DALLOC(s->Bt, sizeof(double) * (size_t)(n + 1) * n * n);
        // the history combine folds its near field in 16-level groups (GEMM forms)
        if (sp.hist_mode != HIST_PLAIN_F64 && s->w.ser_P == 0)
            DALLOC(s->mid, sizeof(double) * (size_t)sp.block * p.ld);
    }
    s->h = sp.length / (sp.nodes - 1);
    fpde_status rc = launch_weights(sp, s->w, s->stream);
    if (rc != FPDE_OK) return rc;
    rc = launch_space_weights(sp, s->w, s->h, s->stream);        // along every line of a 3D sweep
    if (rc != FPDE_OK) return rc;
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    return FPDE_OK;
}

// ---- far part one block ahead (lookahead_spec solves) ---------------------
// The GEMM of block rows n0 .. n0+B-1 over the levels [t_lo, t_hi) on the
// uniform Toeplitz grid into P (K splits <= max_splits when > 0), on stream st;
// in the solve's history form, or in FP64 (k_gemm_f64 on D) when fp64.
fpde_status gemm_range(fpde_solver* s, int n0, int t_lo, int t_hi, double* P, int max_splits, int* splits,
                       hipStream_t st, bool fp64 = false) {
    const fpde_spec& sp = s->spec;
    const Plan& p = s->plan;
    *splits = 0;
    if (t_hi <= t_lo) return FPDE_OK;
    if (is_split(sp.hist_mode) && !fp64) {
        size_t cb = 2 * (size_t)t_lo * p.ld;     // 2-byte components, group-major
        auto cp = [&](int c) { return s->comp[c] ? (const void*)((const char*)s->comp[c] + cb) : nullptr; };
        return launch_gemm_split(sp.hist_mode, cp(0), cp(1), cp(2), s->scale + (size_t)(t_lo / kGroup) * p.ld, p.ld,
                                 (int)p.nodes, n0 - t_lo, t_hi - t_lo, sp.block, s->w.omega, nullptr, 0, s->afrag,
                                 s->ainv, P, splits, st, max_splits);
    }
    return launch_gemm_f64(s->D + (size_t)t_lo * p.ld, p.ld, (int)p.nodes, n0 - t_lo, t_hi - t_lo, sp.block,
                           s->w.omega, P, splits, st, max_splits);
}

// The older part of block n0's far sums, levels [0, tcA), folded into the
// first B rows of P; returns its split count (0 or 1).  The split forms pack
// the levels it needs first (those not packed yet), on the same stream.
// (With a single K split, to draw HBM more slowly beside a step launch, the
// FP64 GEMM lost its occupancy and the split one gained nothing.)
fpde_status older_part(fpde_solver* s, int n0, int tcA, double* P, hipStream_t st, int* splits) {
    const fpde_spec& sp = s->spec;
    fpde_status rc = FPDE_OK;
    if (is_split(sp.hist_mode) && tcA / kGroup > s->packed_groups) {
        const int ghi = tcA / kGroup;
        rc = launch_pack_split(sp.hist_mode, s->D, s->plan.ld, (int)s->plan.nodes, s->packed_groups, ghi,
                               s->comp[0], s->comp[1], s->comp[2], s->scale, st);
        s->packed_groups = ghi;
        if (rc != FPDE_OK) return rc;
    }
    int sp_ = 0;
    rc = gemm_range(s, n0, 0, tcA, P, 0, &sp_, st);
    if (rc == FPDE_OK && sp_ > 1) rc = launch_part_reduce(P, sp_, s->spec.block, s->plan.ld, (int)s->plan.nodes, st);
    *splits = sp_ > 0 ? 1 : 0;
    if (rc == FPDE_OK && tcA > 0) {
        s->stats.history_launches++;
        s->stats.history_alg_bytes += 8.0 * tcA * (double)s->plan.nodes * s->spec.block;
        s->stats.history_alg_flops += 2.0 * tcA * (double)s->plan.nodes * s->spec.block;
    }
    return rc;
}

// Block n0 starts (cut tc; the previous block's cut tcA <= tc): its far sums
// are the older part [0, tcA) -- computed ahead on the side stream during
// the previous block's step launch when lookahead_issue ran for n0, else
// here -- plus the newer part [tcA, tc) as one more split, on the main
// stream.  The newer part (about B levels) is an FP64 MFMA GEMM on D in
// every history form: no packing or fragment kernels on the step chain's
// path; the split forms pack those levels later, with the next older part.
// The two parts' rows sit in this block's parity buffer; the step kernels
// add the two.
fpde_status far_block_lookahead(fpde_solver* s, int n0, int tc, int tcA) {
    const fpde_spec& sp = s->spec;
    const Plan& p = s->plan;
    const int B = sp.block;
    fpde_status rc = FPDE_OK;
    if (!s->ev_far) FPDE_HIP_TRY(hipEventCreateWithFlags(&s->ev_far, hipEventDisableTiming));
    const bool ahead = s->la_n0 == n0 && s->la_tc == tcA;
    if (s->la_n0 >= 0) {
        // the side stream's work is this block's older part, or stale: either
        // way it must be done before this block's buffers are written or read
        FPDE_HIP_TRY(hipStreamWaitEvent(s->stream, s->ev_side, 0));
        s->la_n0 = -1;
    }
    s->la_par ^= 1;
    double* P = s->part + (size_t)s->la_par * p.part_rows * p.ld;
    size_t slot = 0;
    ev_begin(s, 0, &slot);
    int spA = (ahead && tcA > 0) ? 1 : 0;
    if (!ahead) rc = older_part(s, n0, tcA, P, s->stream, &spA);
    int spB = 0;
    if (rc == FPDE_OK) rc = gemm_range(s, n0, tcA, tc, P + (size_t)spA * B * p.ld, 1, &spB, s->stream, true);
    ev_end(s, slot);
    if (rc != FPDE_OK) return rc;
    if (tc > tcA) {
        s->stats.history_launches++;
        s->stats.history_alg_bytes += 8.0 * (tc - tcA) * (double)p.nodes * B;
        s->stats.history_alg_flops += 2.0 * (tc - tcA) * (double)p.nodes * B;
    }
    s->gsplits = spA + spB;
    s->part_cur = P;
    FPDE_HIP_TRY(hipEventRecord(s->ev_far, s->stream));     // the levels [0, tcA) are in memory
    return FPDE_OK;
}

// After the launch of block s->n0 (cut s->tc), ending at n_next = s->n0 + B:
// the older part of block n_next, levels [0, s->tc), on the side stream, into
// the other parity's buffer, while the step launch runs.
fpde_status lookahead_issue(fpde_solver* s, int n_next) {
    if (!s->la || s->tc <= 0 || n_next >= s->spec.steps || !s->ev_far) return FPDE_OK;
    if (!s->side) FPDE_HIP_TRY(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    if (!s->ev_side) FPDE_HIP_TRY(hipEventCreateWithFlags(&s->ev_side, hipEventDisableTiming));
    FPDE_HIP_TRY(hipStreamWaitEvent(s->side, s->ev_far, 0));
    double* P = s->part + (size_t)(s->la_par ^ 1) * s->plan.part_rows * s->plan.ld;
    hipStream_t main = s->stream;
    s->stream = s->side;                      // profiling events go on the side stream
    size_t slot = 0;
    ev_begin(s, 0, &slot);
    int spA = 0;
    fpde_status rc = older_part(s, n_next, s->tc, P, s->side, &spA);
    ev_end(s, slot);
    s->stream = main;
    if (rc != FPDE_OK) return rc;
    FPDE_HIP_TRY(hipEventRecord(s->ev_side, s->side));
    s->la_n0 = n_next;
    s->la_tc = s->tc;
    return FPDE_OK;
}

// Wait for the side stream's pending work (before buffers are reset, copied or freed).
fpde_status lookahead_drain(fpde_solver* s) {
    if (s->side) FPDE_HIP_TRY(hipStreamSynchronize(s->side));
    s->la_n0 = -1;
    return FPDE_OK;
}

// far part of the history for step n (GEMM modes: start a new block when due)
fpde_status far_part(fpde_solver* s, int n, int* nsplit, int64_t* stride, int64_t* off, int* near_lo) {
    const fpde_spec& sp = s->spec;
    const Plan& p = s->plan;
    if (sp.sum_alg == 2) {           // series moments: no history levels to sum
        *nsplit = 0;
        *stride = p.ld;
        *off = 0;
        *near_lo = n;
        return FPDE_OK;
    }
    // levels available in memory: t < avail
    int avail = s->pending ? n - 1 : n;
    if (sp.hist_mode == HIST_PLAIN_F64) {
        int ns = 0;
        size_t slot = 0;
        // restricted summing: lags above jmax have zero weight, t >= n - jmax
        int t_lo = std::max(0, n - s->w.jmax);
        s->omega_cur = s->w.omega;
        if (s->w.grid) {
            // prescribed grid: this step's weight row, omega-style (reversed)
            fpde_status rc = launch_weights_grid(sp, s->w, n, 1, std::max(n, 1), 0, 1, s->Wrow, 0, s->stream);
            if (rc != FPDE_OK) return rc;
            s->omega_cur = s->Wrow;
        }
        if (avail > t_lo) {
            ev_begin(s, 0, &slot);
            fpde_status rc = launch_hist_plain(s->D + (size_t)t_lo * p.ld, p.ld, (int)p.nodes, n - t_lo, avail - t_lo,
                                               s->omega_cur, s->part, &ns, s->stream);
            ev_end(s, slot);
            if (rc != FPDE_OK) return rc;
            s->stats.history_launches++;
            s->stats.history_alg_bytes += 8.0 * (avail - t_lo) * (double)p.nodes;
            s->stats.history_alg_flops += 2.0 * (avail - t_lo) * (double)p.nodes;
        }
        *nsplit = ns;
        *stride = p.ld;
        *off = 0;
        *near_lo = std::max(avail, 0);
        return FPDE_OK;
    }
    int B = sp.block;
    // A block starts where the levels in memory (t < avail) form a multiple of
    // B: the GEMM cut tc = avail is then group aligned and the near field of
    // step n0+r is at most r+1 levels (1D: avail = n-1, since D[n-1] is made by
    // the step launch itself; 3D: avail = n).
    if (s->n0 < 0 || n - s->n0 >= B || (avail > 0 && avail % B == 0 && n > s->n0)) {
        const int tc_prev = s->n0 >= 0 ? s->tc : 0;
        s->n0 = n;
        int tcm = std::max(avail, 0);
        int tc = is_split(sp.hist_mode) ? (tcm / kGroup) * kGroup : tcm;
        s->tc = tc;
        if (s->la) {
            fpde_status rc = far_block_lookahead(s, n, tc, std::min(tc_prev, tc));
            if (rc != FPDE_OK) return rc;
            *nsplit = s->gsplits;
            *stride = (int64_t)B * p.ld;
            *off = 0;
            *near_lo = s->tc;
            s->omega_cur = s->w.omega;
            return FPDE_OK;
        }
        // restricted summing: every row of the block needs t >= n - jmax only
        // (group aligned for the split forms); the GEMM runs on [t_lo, tc)
        int t_lo = std::max(0, n - s->w.jmax);
        if (is_split(sp.hist_mode)) t_lo = (t_lo / kGroup) * kGroup;
        t_lo = std::min(t_lo, tc);
        size_t slot = 0;
        ev_begin(s, 0, &slot);
        fpde_status rc = FPDE_OK;
        if (is_split(sp.hist_mode)) {
            int ghi = tc / kGroup;
            if (ghi > s->packed_groups) {
                rc = launch_pack_split(sp.hist_mode, s->D, p.ld, (int)p.nodes, s->packed_groups, ghi, s->comp[0],
                                       s->comp[1], s->comp[2], s->scale, s->stream);
                s->packed_groups = ghi;
            }
            size_t cb = 2 * (size_t)t_lo * p.ld;     // 2-byte components, group-major
            auto cp = [&](int c) { return s->comp[c] ? (const void*)((const char*)s->comp[c] + cb) : nullptr; };
            if (s->w.grid) {
                // prescribed grid (t_lo = 0: every lag kept): dense far rows
                // for the GEMM, reversed near rows for the step kernels
                if (rc == FPDE_OK)
                    rc = launch_weights_grid(sp, s->w, n, B, std::max(tc, 1), 0, 0, s->Wfar, s->ldwf, s->stream);
                if (rc == FPDE_OK)
                    rc = launch_weights_grid(sp, s->w, n, B, B + 32, tc, 1, s->Wnear, s->ldwn, s->stream);
            }
            if (rc == FPDE_OK)
                rc = launch_gemm_split(sp.hist_mode, cp(0), cp(1), cp(2), s->scale + (size_t)(t_lo / kGroup) * p.ld,
                                       p.ld, (int)p.nodes, n - t_lo, tc - t_lo, B, s->w.omega,
                                       s->w.grid ? s->Wfar : nullptr, s->ldwf, s->afrag, s->ainv, s->part,
                                       &s->gsplits, s->stream);
        } else if (s->w.grid) {
            // prescribed grid: dense weights of the block, far (t < tc) and
            // near (t in [tc, n), one reversed row per step)
            rc = launch_weights_grid(sp, s->w, n, B, std::max(tc, 1), 0, 0, s->Wfar, s->ldwf, s->stream);
            if (rc == FPDE_OK)
                rc = launch_weights_grid(sp, s->w, n, B, B + 32, tc, 1, s->Wnear, s->ldwn, s->stream);
            if (rc == FPDE_OK)
                rc = launch_gemm_f64_dense(s->D, p.ld, (int)p.nodes, tc, B, s->Wfar, s->ldwf, s->part, &s->gsplits,
                                           s->stream);
        } else {
            rc = launch_gemm_f64(s->D + (size_t)t_lo * p.ld, p.ld, (int)p.nodes, n - t_lo, tc - t_lo, B, s->w.omega,
                                 s->part, &s->gsplits, s->stream);
        }
        // 1D: the step kernels read one far value per node (K splits folded
        // here); 3D: the history combine adds the splits itself
        if (rc == FPDE_OK && sp.dims == 1 && s->gsplits > 1) {
            rc = launch_part_reduce(s->part, s->gsplits, B, p.ld, (int)p.nodes, s->stream);
            s->gsplits = 1;
        }
        ev_end(s, slot);
        if (rc != FPDE_OK) return rc;
        if (tc > t_lo) {
            s->stats.history_launches++;
            s->stats.history_alg_bytes += 8.0 * (tc - t_lo) * (double)p.nodes * B;
            s->stats.history_alg_flops += 2.0 * (tc - t_lo) * (double)p.nodes * B;
        }
    }
    *nsplit = s->gsplits;
    *stride = (int64_t)B * p.ld;
    *off = (int64_t)(n - s->n0) * p.ld;
    *near_lo = s->tc;
    s->omega_cur = s->w.grid ? s->Wnear + (size_t)(n - s->n0) * s->ldwn : s->w.omega;
    return FPDE_OK;
}

fpde_status launch_1d(fpde_solver* s, int n, int k, int nsplit, int64_t stride, int64_t off, int near_lo) {
    const fpde_spec& sp = s->spec;
    Step1DArgs a{};
    a.N = sp.nodes;
    a.C = s->C;
    a.h = s->h;
    a.h2inv = 1.0 / (s->h * s->h);
    a.ksat = sp.ksat;
    a.L = (s->w.grid && n < sp.steps) ? s->w.hLv[n] : s->w.local;
    a.utop = s->utop1d;
    a.coef_mode = sp.coef_mode;
    a.n = n;
    a.k = k;
    a.recover = s->pending ? 1 : 0;
    a.parity = (int)(s->launches & 1);
    a.U = s->U;
    a.D = s->D;
    a.ld = s->plan.ld;
    a.omega = s->omega_cur ? s->omega_cur : s->w.omega;
    a.corr = s->w.corr;
    a.part = s->part_cur;
    a.nsplit = nsplit;
    a.part_stride = stride;
    a.part_off = off;
    a.near_lo = near_lo;
    a.Sbuf = s->S;
    a.xyz = s->xyz;
    a.red = s->red;
    a.Uit = s->Uit;
    a.dmax = s->dmax;
    a.flags = s->flags;
    a.tol = sp.picard_tol;
    a.prev_k = s->last_k;
    a.extrap = sp.picard_max > 1 ? 1 : 0;
    a.Uit0 = s->Uit0;
    a.uend = s->uend;
    a.sp_on = s->w.sp_on ? 1 : 0;
    a.sp_kap = s->w.sp_kap;
    a.sp_P = s->spP;
    a.sp_ldp = s->sp_ldp;
    a.sp_F = s->spF;
    a.sp_corr = s->w.sp_corr;
    a.sp_E = s->spE;
    a.ser_P = s->w.ser_P;
    if (a.ser_P > 0) {
        size_t row = (size_t)std::min(n, sp.steps - 1) * a.ser_P;   // the finalize launch reads no row
        a.H = s->H;
        a.ser_c = s->w.ser_c + row;
        a.ser_v = s->w.ser_v + row;
    }
    size_t slot = 0;
    ev_begin(s, 1, &slot);
    fpde_status rc = launch_step1d(a, s->stream);
    ev_end(s, slot);
    if (rc != FPDE_OK) return rc;
    s->launches++;
    s->stats.step_launches++;
    s->pending = (k >= 0);
    s->last_k = k;
    return FPDE_OK;
}

fpde_status finalize_1d(fpde_solver* s);
void count_steps(fpde_solver* s, int n, int n_end, int near_lo);

// space-fractional sum of step n: U^n complete in memory, then F and the
// partial sums of E (read by the step launches)
fpde_status space_part(fpde_solver* s) {
    // a pending solve's finalize launch also forms F of U^n; otherwise (the
    // first step, or U^n finished by an earlier call) the flux kernel does,
    // which also keeps the padding past N at zero
    const bool fused = s->pending;
    fpde_status rc = finalize_1d(s);
    if (rc != FPDE_OK) return rc;
    const fpde_spec& sp = s->spec;
    size_t slot = 0;
    ev_begin(s, 1, &slot);
    if (!fused) rc = launch_space_flux(s->U, sp.nodes, s->h, sp.ksat, sp.coef_mode, s->spF, s->sp_ldp, s->stream);
    if (rc == FPDE_OK) rc = launch_space_apply(s->spF, sp.nodes, s->w.sp_omega, s->spP, s->sp_ldp, s->stream);
    ev_end(s, slot);
    if (rc != FPDE_OK) return rc;
    s->stats.step_launches += fused ? 1 : 2;
    return FPDE_OK;
}

fpde_status enqueue_1d_step(fpde_solver* s, int n) {
    int nsplit = 0, near_lo = 0;
    int64_t stride = 0, off = 0;
    if (s->w.sp_on) {
        fpde_status rc = space_part(s);
        if (rc != FPDE_OK) return rc;
    }
    fpde_status rc = far_part(s, n, &nsplit, &stride, &off, &near_lo);
    if (rc != FPDE_OK) return rc;
    for (int k = 0; k < s->spec.picard_max; k++) {
        rc = launch_1d(s, n, k, nsplit, stride, off, near_lo);
        if (rc != FPDE_OK) return rc;
        s->stats.picard_iters_total++;
    }
    count_steps(s, n, n + 1, near_lo);
    return FPDE_OK;
}

// Tag base of the solver's next cooperative launch: the launch count modulo
// 65535, plus one, in the upper 16 bits (never 0, so no tag equals a zeroed
// exchange word; consecutive launches differ), the barrier index + 1 modulo
// 65536 in the lower 16 (set by the kernel).
unsigned launch_tag(const fpde_solver* s) { return (unsigned)(s->coop_launches % 65535 + 1) << 16; }

// Arguments of the cooperative block launch for the steps [n, n_end).
Step1DArgs block_args(fpde_solver* s, int n, int n_end, int n0, int nsplit, int64_t stride, int near_lo) {
    const fpde_spec& sp = s->spec;
    Step1DArgs a{};
    a.N = sp.nodes;
    a.C = s->C;
    a.h = s->h;
    a.h2inv = 1.0 / (s->h * s->h);
    a.ksat = sp.ksat;
    a.L = s->w.local;
    a.utop = s->utop1d;
    a.coef_mode = sp.coef_mode;
    a.n = n;
    a.n_end = n_end;
    a.n0 = n0;
    a.blk = sp.block;
    a.mid = s->mid;
    a.picard_max = sp.picard_max;
    a.tol = sp.picard_tol;
    a.extrap = sp.picard_max > 1 ? 1 : 0;
    a.U = s->U;
    a.D = s->D;
    a.ld = s->plan.ld;
    a.omega = s->w.grid ? (sp.hist_mode == HIST_PLAIN_F64 ? s->Wrow : s->Wnear) : s->w.omega;
    a.omega_stride = (s->w.grid && sp.hist_mode != HIST_PLAIN_F64) ? s->ldwn : 0;
    a.Lv = s->w.grid ? s->w.Lv : nullptr;
    a.corr = s->w.corr;
    a.part = s->part_cur;
    a.nsplit = nsplit;
    a.part_stride = stride;
    a.near_lo = near_lo;
    a.red = s->red;
    a.dmax = s->dmax;
    a.xch = s->xch;
    a.xtag = launch_tag(s);
    a.err = s->err;
    a.wait_ticks = s->wait_ticks;
    a.ser_P = s->w.ser_P;
    if (a.ser_P > 0) {
        a.H = s->H;
        a.ser_c = s->w.ser_c;
        a.ser_v = s->w.ser_v;
    }
    return a;
}

// Algorithmic bytes of the 1D step kernels for the steps [n, n_end): 100 B
// per node and step (coefficients, rows, line solve) and 8 B per near-field
// history term, t in [near_lo, m) for step m (the series form keeps P
// moments instead: 16 B each, read and written).
void count_steps(fpde_solver* s, int n, int n_end, int near_lo) {
    const double N = (double)s->spec.nodes;
    double terms = 0.0;
    for (int m = n; m < n_end; m++) terms += (double)std::max(0, m - near_lo);
    const double per_step = s->w.ser_P > 0 ? 100.0 + 16.0 * s->w.ser_P : 100.0;
    s->stats.step_alg_bytes += N * (per_step * (n_end - n) + (s->w.ser_P > 0 ? 0.0 : 8.0 * terms));
    s->stats.steps_done += n_end - n;
}

// Cooperative block launch for the steps [n, n_end) (no pending solve
// between launches: each launch leaves U^{n_end} and D[n_end-1] in memory).
fpde_status launch_1d_block(fpde_solver* s, int n, int n_end, int n0, int nsplit, int64_t stride, int near_lo) {
    const fpde_spec& sp = s->spec;
    const Step1DArgs a = block_args(s, n, n_end, n0, nsplit, stride, near_lo);
    size_t slot = 0;
    ev_begin(s, 1, &slot);
    fpde_status rc = launch_step1d_block(a, s->stream);
    ev_end(s, slot);
    if (rc != FPDE_OK) return rc;
    s->launches++;
    s->coop_launches++;
    s->stats.step_launches++;
    count_steps(s, n, n_end, near_lo);      // Picard iterations: counted on the device (err[1])
    return FPDE_OK;
}

// The step range [n, n_end) of the next cooperative launch and the first
// step n0 of its history block: one launch per history block (GEMM forms),
// per step (plain: its history kernel runs between steps), or per 256 steps
// (series moments, no history kernels).
void next_block(const fpde_solver* s, int n, int end, int* n_end, int* n0) {
    const fpde_spec& sp = s->spec;
    if (sp.sum_alg == 2) { *n_end = std::min(end, n + 256); *n0 = n; }
    else if (sp.hist_mode == HIST_PLAIN_F64) { *n_end = n + 1; *n0 = n; }   // step_mode 0 never selects this
    else { *n_end = std::min(end, s->n0 + sp.block); *n0 = s->n0; }
}

// 1D steps [n, end) in cooperative block launches.
fpde_status enqueue_1d_blocks(fpde_solver* s, int n, int end) {
    while (n < end) {
        int nsplit = 0, near_lo = 0;
        int64_t stride = 0, off = 0;
        fpde_status rc = far_part(s, n, &nsplit, &stride, &off, &near_lo);
        if (rc != FPDE_OK) return rc;
        int n_end, n0;
        next_block(s, n, end, &n_end, &n0);
        rc = launch_1d_block(s, n, n_end, n0, nsplit, stride, near_lo);
        if (rc != FPDE_OK) return rc;
        // the block is complete: the next one's older far part starts now,
        // beside this launch
        if (s->la && n_end == s->n0 + s->spec.block && (rc = lookahead_issue(s, n_end)) != FPDE_OK) return rc;
        n = n_end;
        s->step = n;
    }
    return FPDE_OK;
}

fpde_status check_barrier(fpde_solver* s) {
    if (!s->block_mode) return FPDE_OK;
    int e = 0;
    FPDE_HIP_TRY(hipMemcpy(&e, s->err, sizeof(int), hipMemcpyDeviceToHost));
    if (e) {
        set_error("1D cooperative step: a grid barrier timed out (workgroups not co-resident); "
                  "recreate the solver with step_mode 1");
        return FPDE_ERR_HIP;
    }
    return FPDE_OK;
}

fpde_status finalize_1d(fpde_solver* s) {
    if (!s->pending) return FPDE_OK;
    return launch_1d(s, s->step, -1, 0, 0, 0, s->step);
}

// Weight rows of a 3D step's history term: the GEMM forms give the row of
// the block's first step (one Toeplitz vector, or on a prescribed grid the
// block's rows ldwn apart); the plain form the weights of step n itself
// (far_part set them).
struct HistRows { const double* omega; int64_t stride; int n0; };
HistRows hist_rows(const fpde_solver* s, int n) {
    if (s->spec.hist_mode != HIST_PLAIN_F64)
        return s->w.grid ? HistRows{s->Wnear, s->ldwn, s->n0} : HistRows{s->w.omega, 0, s->n0};
    return HistRows{(s->w.grid && s->omega_cur) ? s->omega_cur : s->w.omega, 0, n};
}

fpde_status enqueue_3d_step(fpde_solver* s, int n) {
    const fpde_spec& sp = s->spec;
    const Plan& p = s->plan;
    int nsplit = 0, near_lo = 0;
    int64_t stride = 0, off = 0;
    fpde_status rc = far_part(s, n, &nsplit, &stride, &off, &near_lo);
    if (rc != FPDE_OK) return rc;
    int nn = sp.nodes;
    double h2inv = 1.0 / (s->h * s->h);
    size_t slot = 0;
    const bool series = s->w.ser_P > 0;
    // split3d odd: every sweep carries a third of the history term (Om1
    // fpde_hpc_1d.cpp:4172-4173, :4258-4263); even: the y sweep all of it
    const bool split = (sp.split3d & 1) != 0;
    const double hscale = split ? 1.0 / 3.0 : 1.0;
    ev_begin(s, 0, &slot);
    if (series) {
        size_t row = (size_t)n * s->w.ser_P;
        const double* dnew = (n >= 1) ? s->D + (size_t)((n - 1) & 1) * p.ld : nullptr;
        rc = launch_series3d(s->S, s->H, s->w.ser_P, dnew, s->w.ser_c + row, s->w.ser_v + row, p.nodes, p.ld,
                             hscale, s->stream);
    } else {
        const HistRows hr = hist_rows(s, n);
        rc = launch_hist_combine(s->S, s->part, nsplit, stride, off, s->D, p.ld, p.nodes, n, near_lo, hr.omega,
                                 hr.stride, hr.n0, sp.block, s->mid, s->w.corr, hscale, s->stream);
    }
    ev_end(s, slot);
    if (rc != FPDE_OK) return rc;
    ev_begin(s, 1, &slot);
    // space-fractional scheme: the line sums of each sweep from the field it
    // starts from (K of that field: coef3d precedes every sweep then)
    const bool space = s->w.sp_on;
    const double kap = space ? s->w.sp_kap : 1.0;
    double* E = space ? s->spE : nullptr;
    auto space_sums = [&](int axis) {
        return space ? launch_space3d(s->U, s->K, nn, axis, s->h, s->w.sp_omega, s->w.sp_corr, s->spE, s->stream)
                     : FPDE_OK;
    };
    // local coefficient of this step: per step on a grid (generalized kernel)
    const double Lc = (s->w.grid && n < sp.steps) ? s->w.hLv[n] : s->w.local;
    // the partitioned lateral sweeps evaluate their coefficients themselves;
    // a coefficient pass precedes the one-thread-per-line sweeps and the
    // space-fractional line sums (which need K)
    const bool pass = space || !lateral3d_partitioned(nn, Lc);
    const double* Kc = pass ? s->K : nullptr;
    const double* iCc = pass ? s->iC : nullptr;
    auto coef_pass = [&]() {
        return pass ? launch_coef3d(s->U, nn, s->h, sp.ksat, sp.coef_mode, s->K, s->iC, s->stream) : FPDE_OK;
    };
    if ((rc = coef_pass()) != FPDE_OK) return rc;
    if ((rc = space_sums(2)) != FPDE_OK) return rc;
    if ((rc = launch_lateral3d(s->U, Kc, iCc, s->S, nn, 2, h2inv, Lc, s->Al, s->Bt, kap, E, s->h, sp.ksat,
                               sp.coef_mode, s->stream)) != FPDE_OK) return rc;
    if ((rc = coef_pass()) != FPDE_OK) return rc;
    if ((rc = space_sums(1)) != FPDE_OK) return rc;
    if ((rc = launch_lateral3d(s->U, Kc, iCc, split ? s->S : nullptr, nn, 1, h2inv, Lc, s->Al, s->Bt, kap, E, s->h,
                               sp.ksat, sp.coef_mode, s->stream)) != FPDE_OK) return rc;
    if (space) {
        if ((rc = launch_coef3d(s->U, nn, s->h, sp.ksat, sp.coef_mode, s->K, s->iC, s->stream)) != FPDE_OK) return rc;
        if ((rc = space_sums(0)) != FPDE_OK) return rc;
    }
    // the depth sweep reuses the coefficient pass of the space-fractional sums
    if ((rc = launch_zline3d(s->U, s->Uprev, s->utop, nn, s->h, h2inv, sp.ksat, sp.coef_mode, Lc,
                             s->D + (size_t)(series ? (n & 1) : n) * p.ld, kap, E, space ? s->K : nullptr,
                             space ? s->iC : nullptr, split ? s->S : nullptr, s->stream)) != FPDE_OK) return rc;
    ev_end(s, slot);
    s->stats.step_launches += space ? 9 : (pass ? 5 : 3);
    s->stats.picard_iters_total++;
    s->stats.step_alg_bytes += 100.0 * (double)p.nodes;     // the three sweeps' coefficients, rows and solves
    s->stats.steps_done++;
    return FPDE_OK;
}

}  // namespace

namespace fpde {
// Whether a solver advances in the batched cooperative launches, and its
// chunk workgroups (the server groups such solvers into batches).
bool solver_batchable(const fpde_solver* s) { return s && s->spec.dims == 1 && s->block_mode; }
int solver_chunks(const fpde_solver* s) { return s ? s->C : 0; }

// 3D solves that take the batched LOD step: the partitioned sweeps (n <= 128),
// the full or restricted history (not the series moments), classical diffusion.
bool solver_3d_batchable(const fpde_solver* s) {
    return s && s->spec.dims == 3 && s->spec.nodes <= 128 && s->w.ser_P == 0 && !s->w.sp_on && s->w.local > 0.0;
}

// Up to kMax3DBatch such solves advanced by nsteps (<= 0: to the end) in
// lockstep: per step, each solve's far part (its history GEMM when a block
// starts) and then one launch per kernel of the LOD step for all of them, on
// the first solve's stream.  Asynchronous: every solve's own stream is made to
// wait for the batch, so fpde_solver_sync / query cover it.
fpde_status advance_3d_batch(fpde_solver* const* v, int count, int nsteps) {
    if (count < 1 || count > kMax3DBatch) { set_error("3D batch: 1 to 8 solvers"); return FPDE_ERR_ARG; }
    std::vector<int> end(count);
    for (int j = 0; j < count; j++) {
        if (!solver_3d_batchable(v[j]) || !v[j]->initialized) { set_error("3D batch: solver not batchable"); return FPDE_ERR_ARG; }
        end[j] = nsteps <= 0 ? v[j]->spec.steps : std::min(v[j]->spec.steps, v[j]->step + nsteps);
    }
    hipStream_t bst = v[0]->stream;
    std::vector<hipEvent_t> ev(count + 1, nullptr);
    fpde_status rc = FPDE_OK;
    auto hip_ok = [&](hipError_t e) {
        if (e == hipSuccess) return true;
        set_error(std::string("3D batch: ") + hipGetErrorString(e));
        rc = FPDE_ERR_HIP;
        return false;
    };
    for (int j = 0; j <= count && rc == FPDE_OK; j++) hip_ok(hipEventCreateWithFlags(&ev[j], hipEventDisableTiming));
    // the batch follows each solve's earlier work
    for (int j = 1; j < count && rc == FPDE_OK; j++)
        if (v[j]->stream != bst && hip_ok(hipEventRecord(ev[j], v[j]->stream))) hip_ok(hipStreamWaitEvent(bst, ev[j], 0));
    while (rc == FPDE_OK) {
        Lod3DBatch b{};
        int k = 0;
        for (int j = 0; j < count && rc == FPDE_OK; j++) {
            fpde_solver* s = v[j];
            const int n = s->step;
            if (n >= end[j]) continue;
            const fpde_spec& sp = s->spec;
            const Plan& p = s->plan;
            int nsplit = 0, near_lo = 0;
            int64_t stride = 0, off = 0;
            hipStream_t own = s->stream;
            s->stream = bst;                       // the far part goes on the batch's stream
            rc = far_part(s, n, &nsplit, &stride, &off, &near_lo);
            s->stream = own;
            if (rc != FPDE_OK) break;
            Lod3DRow& r = b.row[k];
            r.U = s->U; r.S = s->S; r.Uprev = s->Uprev;
            r.Dlev = s->D + (size_t)n * p.ld;
            r.part = s->part; r.D = s->D;
            const HistRows hr = hist_rows(s, n);
            r.omega = hr.omega; r.omega_stride = hr.stride; r.n0 = hr.n0; r.blk = sp.block; r.mid = s->mid;
            r.corr = s->w.corr; r.utop = s->utop;
            r.part_stride = stride; r.part_off = off; r.ld = p.ld; r.nodes = p.nodes;
            r.n = sp.nodes; r.nsplit = nsplit; r.step = n; r.near_lo = near_lo; r.coef_mode = sp.coef_mode;
            r.split = (sp.split3d & 1) ? 1 : 0;
            r.h = s->h; r.h2inv = 1.0 / (s->h * s->h); r.ksat = sp.ksat;
            r.L = (s->w.grid && n < sp.steps) ? s->w.hLv[n] : s->w.local;
            r.hscale = r.split ? 1.0 / 3.0 : 1.0;
            r.kap = 1.0; r.E = nullptr; r.Kz = nullptr; r.iCz = nullptr; r.Kl = nullptr; r.iCl = nullptr;
            const int lines = sp.nodes * sp.nodes;
            b.first_node[k + 1] = b.first_node[k] + (int)((p.nodes + 255) / 256);
            b.first_lat[k + 1] = b.first_lat[k] + (lines + 31) / 32;
            b.first_z[k + 1] = b.first_z[k] + (lines + kZLinesHost - 1) / kZLinesHost;
            k++;
        }
        if (rc != FPDE_OK || k == 0) break;
        b.count = k;
        rc = launch_lod3d_batch(b, bst);
        if (rc != FPDE_OK) break;
        for (int j = 0; j < count; j++) {
            fpde_solver* s = v[j];
            if (s->step >= end[j]) continue;
            s->step++;
            s->stats.step_launches += 4;
            s->stats.picard_iters_total++;
            s->stats.step_alg_bytes += 100.0 * (double)s->plan.nodes;
            s->stats.steps_done++;
        }
    }
    // each solve's own stream waits for the batch
    if (rc == FPDE_OK && hip_ok(hipEventRecord(ev[count], bst)))
        for (int j = 1; j < count && rc == FPDE_OK; j++)
            if (v[j]->stream != bst) hip_ok(hipStreamWaitEvent(v[j]->stream, ev[count], 0));
    for (hipEvent_t e : ev)
        if (e) hipEventDestroy(e);
    return rc;
}

// Advance the 3D solves among `v` that take the batched step, in groups of
// kMax3DBatch, and every other solve on its own stream (asynchronous).
fpde_status advance_riders(fpde_solver* const* v, int count, int nsteps) {
    std::vector<fpde_solver*> grp;
    for (int i = 0; i < count; i++) {
        if (solver_3d_batchable(v[i])) {
            grp.push_back(v[i]);
            if ((int)grp.size() == kMax3DBatch) {
                fpde_status rc = advance_3d_batch(grp.data(), (int)grp.size(), nsteps);
                if (rc != FPDE_OK) return rc;
                grp.clear();
            }
            continue;
        }
        fpde_status rc = fpde_solver_advance(v[i], nsteps);
        if (rc != FPDE_OK) return rc;
    }
    if (!grp.empty()) return advance_3d_batch(grp.data(), (int)grp.size(), nsteps);
    return FPDE_OK;
}
}  // namespace fpde

// ===========================================================================
// C ABI
extern "C" {

const char* fpde_last_error(void) { return g_err.c_str(); }

fpde_status fpde_shutdown(void) {
    std::lock_guard<std::mutex> lock(g_rows_mu);
    fpde_status rc = FPDE_OK;
    int cur = 0;
    const bool have_dev = hipGetDevice(&cur) == hipSuccess;
    for (auto& kv : g_rows_tab) {
        if (!kv.second) continue;
        if (hipSetDevice(kv.first) == hipSuccess) {
            if (hipDeviceSynchronize() != hipSuccess) rc = FPDE_ERR_HIP;
            if (hipFree(kv.second) != hipSuccess) rc = FPDE_ERR_HIP;
        } else {
            rc = FPDE_ERR_HIP;
        }
        kv.second = nullptr;
    }
    g_rows_tab.clear();
    if (have_dev) {
        (void)hipSetDevice(cur);
        if (hipDeviceSynchronize() != hipSuccess) rc = FPDE_ERR_HIP;
    }
    if (rc != FPDE_OK) set_error("fpde_shutdown: a device call failed while releasing the library's tables");
    return rc;
}
int fpde_version(void) { return 100; }
int fpde_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

size_t fpde_solver_footprint(const fpde_spec* spec) {
    if (!spec) return 0;
    return footprint_bytes(*spec);
}

fpde_status fpde_solver_create(const fpde_spec* spec, fpde_solver** out) {
    return solver_create_in(spec, nullptr, nullptr, out);
}

}  // extern "C"

fpde_status fpde::solver_create_in(const fpde_spec* spec, DeviceArena* arena, hipStream_t stream, fpde_solver** out) {
    if (!spec || !out) { set_error("null argument"); return FPDE_ERR_ARG; }
    *out = nullptr;
    fpde_status rc = validate(*spec);
    if (rc != FPDE_OK) return rc;
    auto s = std::make_unique<fpde_solver>();
    s->arena = arena;
    if (stream) {
        s->stream = stream;
        s->own_stream = false;
    }
    s->spec = *spec;
    if (spec->taus) {
        s->taus_h.assign(spec->taus, spec->taus + spec->steps);
        s->spec.taus = s->taus_h.data();
    }
    s->plan = make_plan(*spec);
    rc = solver_alloc(s.get());
    if (rc != FPDE_OK) {
        fpde_solver_destroy(s.release());
        return rc;
    }
    *out = s.release();
    return FPDE_OK;
}

extern "C" {

fpde_status fpde_solver_destroy(fpde_solver* s) {
    if (!s) return FPDE_OK;
    if (s->stream) hipStreamSynchronize(s->stream);
    if (s->side) hipStreamSynchronize(s->side);
    if (s->gexec) hipGraphExecDestroy(s->gexec);
    for (void* p : s->allocs) {
        if (s->arena) s->arena->release(p);
        else hipFree(p);
    }
    for (auto& e : s->evs) { hipEventDestroy(e.a); hipEventDestroy(e.b); }
    if (s->ev_far) hipEventDestroy(s->ev_far);
    if (s->ev_side) hipEventDestroy(s->ev_side);
    if (s->side) hipStreamDestroy(s->side);
    if (s->stream && s->own_stream) hipStreamDestroy(s->stream);
    delete s;
    return FPDE_OK;
}

fpde_status fpde_solver_set_initial(fpde_solver* s, const double* u0) {
    if (!s || !u0) { set_error("null argument"); return FPDE_ERR_ARG; }
    if (s->step != 0) { set_error("initial state can only be set before the first step"); return FPDE_ERR_STATE; }
    const Plan& p = s->plan;
    FPDE_HIP_TRY(hipMemsetAsync(s->U, 0, sizeof(double) * p.ld, s->stream));
    FPDE_HIP_TRY(hipMemcpyAsync(s->U, u0, sizeof(double) * p.nodes, hipMemcpyHostToDevice, s->stream));
    if (s->H) FPDE_HIP_TRY(hipMemsetAsync(s->H, 0, sizeof(double) * (size_t)s->w.ser_P * p.ld, s->stream));
    if (s->spec.dims == 1) {
        s->utop1d = u0[0];
    } else {
        int n = s->spec.nodes;
        std::vector<double> top((size_t)n * n);
        for (int y = 0; y < n; y++)
            for (int x = 0; x < n; x++) top[(size_t)y * n + x] = u0[(size_t)n * (x + (size_t)n * y)];
        FPDE_HIP_TRY(hipMemcpyAsync(s->Uprev, s->U, sizeof(double) * p.ld, hipMemcpyDeviceToDevice, s->stream));
        FPDE_HIP_TRY(hipMemcpyAsync(s->utop, top.data(), sizeof(double) * top.size(), hipMemcpyHostToDevice, s->stream));
        s->u0_h.assign(u0, u0 + p.nodes);
    }
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    s->initialized = true;
    return FPDE_OK;
}

fpde_status fpde_solver_reset(fpde_solver* s, const double* u0) {
    if (!s || !u0) { set_error("null argument"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    if (fpde_status rc = lookahead_drain(s); rc != FPDE_OK) return rc;
    ev_collect(s);
    s->step = 0;
    s->launches = 0;
    s->last_k = -1;
    s->pending = false;
    s->n0 = -1;
    s->tc = 0;
    s->gsplits = 0;
    s->packed_groups = 0;
    s->part_cur = s->part;
    s->la_par = 0;
    if (s->xch) FPDE_HIP_TRY(hipMemsetAsync(s->xch, 0, kXchBytes, s->stream));
    if (s->err) FPDE_HIP_TRY(hipMemsetAsync(s->err, 0, 64, s->stream));
    s->picard_dev_base = 0;
    s->coop_launches = 0;
    return fpde_solver_set_initial(s, u0);
}

fpde_status fpde_solver_advance(fpde_solver* s, int nsteps) {
    if (!s) { set_error("null solver"); return FPDE_ERR_ARG; }
    if (!s->initialized) { set_error("set_initial must be called first"); return FPDE_ERR_STATE; }
    int M = s->spec.steps;
    int end = (nsteps <= 0) ? M : std::min(M, s->step + nsteps);
    if (s->spec.dims == 1 && s->block_mode) {
        fpde_status rc = enqueue_1d_blocks(s, s->step, end);
        if (rc != FPDE_OK) return rc;
        FPDE_HIP_TRY(hipGetLastError());
        return FPDE_OK;
    }
    for (int n = s->step; n < end; n++) {
        fpde_status rc = (s->spec.dims == 1) ? enqueue_1d_step(s, n) : enqueue_3d_step(s, n);
        if (rc != FPDE_OK) return rc;
        s->step = n + 1;
    }
    if (s->spec.dims == 1 && s->step == M) {
        fpde_status rc = finalize_1d(s);
        if (rc != FPDE_OK) return rc;
    }
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

// Many solvers advanced together.  Per history block: the far part of each
// 1D block-mode solver on its own stream (they overlap), then one cooperative
// launch of k_step1d_batch that runs the block's steps of all of them.  The
// other solvers advance on their own streams meanwhile.  Returns when every
// solver has reached its end step.
fpde_status fpde_solver_advance_batch(fpde_solver* const* solvers, int count, int nsteps) {
    if (!solvers || count < 1 || count > kMaxBatch) { set_error("advance_batch: 1 to 32 solvers"); return FPDE_ERR_ARG; }
    for (int i = 0; i < count; i++) {
        if (!solvers[i]) { set_error("advance_batch: null solver"); return FPDE_ERR_ARG; }
        if (!solvers[i]->initialized) { set_error("set_initial must be called first"); return FPDE_ERR_STATE; }
        for (int j = 0; j < i; j++)
            if (solvers[j] == solvers[i]) { set_error("advance_batch: a solver appears twice"); return FPDE_ERR_ARG; }
    }
    std::vector<fpde_solver*> blk;
    std::vector<int> ends;
    int64_t wgs = 0;
    for (int i = 0; i < count; i++) {
        fpde_solver* s = solvers[i];
        if (s->spec.dims == 1 && s->block_mode) {
            if (s->C < 1 || s->C > kMaxChunks || s->w.sp_on) {   // what the block launch itself refuses
                set_error("advance_batch: solver outside the cooperative block form's limits");
                return FPDE_ERR_ARG;
            }
            blk.push_back(s);
            ends.push_back(nsteps <= 0 ? s->spec.steps : std::min(s->spec.steps, s->step + nsteps));
            wgs += s->C;
        }
    }
    if (!blk.empty() && !step1d_batch_supported(wgs)) {
        set_error("advance_batch: the batch's " + std::to_string(wgs) +
                  " chunk workgroups exceed what the GPU holds at once; split the batch");
        return FPDE_ERR_ARG;
    }
    // the others: batched 3D LOD steps, or their own streams
    std::vector<fpde_solver*> riders;
    for (int i = 0; i < count; i++)
        if (!(solvers[i]->spec.dims == 1 && solvers[i]->block_mode)) riders.push_back(solvers[i]);
    if (!riders.empty()) {
        fpde_status rc = advance_riders(riders.data(), (int)riders.size(), nsteps);
        if (rc != FPDE_OK) return rc;
    }
    if (blk.empty()) {
        for (int i = 0; i < count; i++) FPDE_HIP_TRY(hipStreamSynchronize(solvers[i]->stream));
        return FPDE_OK;
    }
    const int B = (int)blk.size();
    // the argument rows of the batch (pointers and constants of each solve),
    // in the device's table (g_rows_tab)
    std::lock_guard<std::mutex> lock(g_rows_mu);
    int dev = 0;
    FPDE_HIP_TRY(hipGetDevice(&dev));
    Step1DArgs*& drows = g_rows_tab[dev];
    if (!drows) FPDE_HIP_TRY(hipMalloc((void**)&drows, sizeof(Step1DArgs) * kMaxBatch));
    std::vector<Step1DArgs> rows(B);
    for (int j = 0; j < B; j++) rows[j] = block_args(blk[j], blk[j]->step, blk[j]->step, 0, 0, 0, 0);
    for (fpde_solver* s : blk) FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    std::vector<hipEvent_t> ev(B + 1, nullptr);
    fpde_status rc = FPDE_OK;
    hipStream_t bst = blk[0]->stream;
    auto hip_ok = [&](hipError_t e) {
        if (e == hipSuccess) return true;
        set_error(std::string("advance_batch: ") + hipGetErrorString(e));
        rc = FPDE_ERR_HIP;
        return false;
    };
    // the previous batch's launches have finished: that call synchronised
    // its streams before returning
    if (hip_ok(hipMemcpy(drows, rows.data(), sizeof(Step1DArgs) * B, hipMemcpyHostToDevice))) {
        for (int j = 0; j <= B && rc == FPDE_OK; j++) hip_ok(hipEventCreateWithFlags(&ev[j], hipEventDisableTiming));
    }
    while (rc == FPDE_OK) {
        BatchTick t{};
        t.count = B;
        int first = 0;
        for (int j = 0; j < B && rc == FPDE_OK; j++) {
            fpde_solver* s = blk[j];
            t.first[j] = first;
            t.n[j] = t.n_end[j] = s->step;
            if (s->step >= ends[j]) continue;                // done: no workgroups
            int nsplit = 0, near_lo = 0, n_end = 0, n0 = 0;
            int64_t stride = 0, off = 0;
            rc = far_part(s, s->step, &nsplit, &stride, &off, &near_lo);
            if (rc != FPDE_OK) break;
            next_block(s, s->step, ends[j], &n_end, &n0);
            t.n_end[j] = n_end;
            t.n0[j] = n0;
            t.nsplit[j] = nsplit;
            t.near_lo[j] = near_lo;
            t.part_stride[j] = stride;
            t.part[j] = s->part_cur;
            t.xtag[j] = launch_tag(s);
            first += s->C;
            if (s->stream != bst && hip_ok(hipEventRecord(ev[j], s->stream))) hip_ok(hipStreamWaitEvent(bst, ev[j], 0));
        }
        if (rc != FPDE_OK) break;
        t.first[B] = first;
        if (first == 0) break;
        rc = launch_step1d_batch(drows, t, bst);
        if (rc != FPDE_OK) break;
        if (!hip_ok(hipEventRecord(ev[B], bst))) break;
        for (int j = 0; j < B; j++) {
            fpde_solver* s = blk[j];
            if (t.n_end[j] == t.n[j]) continue;
            if (s->stream != bst && !hip_ok(hipStreamWaitEvent(s->stream, ev[B], 0))) break;
            s->launches++;
            s->coop_launches++;
            s->stats.step_launches++;
            count_steps(s, t.n[j], t.n_end[j], t.near_lo[j]);
            s->step = t.n_end[j];
            // (no far part ahead here: beside a batched launch, which already
            // fills the GPU, the solves' side GEMMs slowed the batch more than
            // they hid: 6 x N=1e4 took 0.097 s against 0.087 s)
        }
    }
    // the argument rows stay alive until the last launch has run
    for (int i = 0; i < count; i++) {
        hipError_t e = hipStreamSynchronize(solvers[i]->stream);
        if (rc == FPDE_OK) hip_ok(e);
    }
    for (hipEvent_t e : ev)
        if (e) hipEventDestroy(e);
    if (rc != FPDE_OK) return rc;
    for (fpde_solver* s : blk) {
        rc = check_barrier(s);
        if (rc != FPDE_OK) return rc;
    }
    return FPDE_OK;
}

fpde_status fpde_autotune_block(const fpde_spec* spec, const double* u0, int trial_steps, int* block_out,
                                double* ms_per_step_out) {
    if (!spec || !u0 || !block_out) { set_error("null argument"); return FPDE_ERR_ARG; }
    fpde_spec base = *spec;
    if (base.hist_mode != HIST_PLAIN_F64 && base.block == 0) base.block = 64;   // 0: "pick one"
    fpde_status rc = validate(base);
    if (rc != FPDE_OK) return rc;
    *block_out = base.block;
    if (ms_per_step_out) *ms_per_step_out = 0.0;
    if (base.hist_mode == HIST_PLAIN_F64) return FPDE_OK;   // no history GEMM, no block
    const int cands[4] = {16, 32, 64, 128};
    int steps = trial_steps > 0 ? trial_steps : 2048;
    steps = std::min(base.steps, std::max(steps, 2 * 128 + 1));
    double best = -1.0;
    for (int b : cands) {
        fpde_spec t = base;
        t.block = b;
        t.steps = steps;                        // a prescribed grid runs its first `steps` entries
        fpde_solver* sv = nullptr;
        rc = fpde_solver_create(&t, &sv);
        if (rc != FPDE_OK) return rc;
        rc = fpde_solver_set_initial(sv, u0);
        if (rc == FPDE_OK) rc = fpde_solver_sync(sv);
        auto t0 = std::chrono::steady_clock::now();
        if (rc == FPDE_OK) rc = fpde_solver_advance(sv, 0);
        if (rc == FPDE_OK) rc = fpde_solver_sync(sv);
        double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        fpde_solver_destroy(sv);
        if (rc != FPDE_OK) return rc;
        if (best < 0.0 || ms < best) { best = ms; *block_out = b; }
    }
    if (ms_per_step_out) *ms_per_step_out = best / steps;
    return FPDE_OK;
}

fpde_status fpde_solver_advance_graph(fpde_solver* s, int nsteps, int steps_per_graph) {
    if (!s) { set_error("null solver"); return FPDE_ERR_ARG; }
    if (!s->initialized) { set_error("set_initial must be called first"); return FPDE_ERR_STATE; }
    if (steps_per_graph < 1 || steps_per_graph > 4096) { set_error("steps_per_graph must be in [1,4096]"); return FPDE_ERR_ARG; }
    if (s->spec.dims == 1 && s->block_mode) {
        set_error("graph capture: the cooperative block launches (step_mode 0) are not captured; use step_mode 1");
        return FPDE_ERR_STATE;
    }
    if (s->spec.var_tau) { set_error("graph capture: var_tau solvers step one length at a time"); return FPDE_ERR_STATE; }
    const int M = s->spec.steps;
    const int end = (nsteps <= 0) ? M : std::min(M, s->step + nsteps);
    const int prof = s->prof;
    s->prof = 0;                                   // no event records inside a capture
    fpde_status rc = FPDE_OK;
    while (rc == FPDE_OK && s->step < end) {
        const int g_end = std::min(end, s->step + steps_per_graph);
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) { set_error(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e)); rc = FPDE_ERR_HIP; break; }
        for (int n = s->step; n < g_end && rc == FPDE_OK; n++) {
            rc = (s->spec.dims == 1) ? enqueue_1d_step(s, n) : enqueue_3d_step(s, n);
            if (rc == FPDE_OK) s->step = n + 1;
        }
        if (rc == FPDE_OK && s->spec.dims == 1 && s->step == M) rc = finalize_1d(s);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(s->stream, &graph);
        if (rc != FPDE_OK) { if (graph) hipGraphDestroy(graph); break; }
        if (e != hipSuccess || !graph) { set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e)); rc = FPDE_ERR_HIP; break; }
        // reuse the executable graph when the chunk has the same topology
        bool updated = false;
        if (s->gexec) {
            hipGraphExecUpdateResult res;
            hipGraphNode_t bad = nullptr;
            updated = hipGraphExecUpdate(s->gexec, graph, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) { hipGetLastError(); hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
        }
        if (!updated) {
            e = hipGraphInstantiate(&s->gexec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) { s->gexec = nullptr; hipGraphDestroy(graph); set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e)); rc = FPDE_ERR_HIP; break; }
            s->graph_instantiations++;
        }
        e = hipGraphLaunch(s->gexec, s->stream);
        hipGraphDestroy(graph);
        if (e != hipSuccess) { set_error(std::string("hipGraphLaunch: ") + hipGetErrorString(e)); rc = FPDE_ERR_HIP; break; }
        s->graph_launches++;
    }
    s->prof = prof;
    if (rc != FPDE_OK) {
        // the host state has advanced over steps whose launches were not run:
        // the solver needs fpde_solver_reset before further steps
        s->initialized = false;
        return rc;
    }
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status fpde_solver_advance_tau(fpde_solver* s, double tau) {
    if (!s) { set_error("null solver"); return FPDE_ERR_ARG; }
    if (!s->spec.var_tau) { set_error("advance_tau needs a var_tau solver"); return FPDE_ERR_STATE; }
    if (!s->initialized) { set_error("set_initial must be called first"); return FPDE_ERR_STATE; }
    if (!(tau > 0.0 && std::isfinite(tau))) { set_error("tau must be positive and finite"); return FPDE_ERR_ARG; }
    int n = s->step;
    if (n >= s->spec.steps) { set_error("var_tau solver: step capacity reached"); return FPDE_ERR_STATE; }
    s->taus_h[n] = tau;
    s->Time_h[n + 1] = s->Time_h[n] + tau;                         // init_time_step (:3944)
    s->w.hLv[n] = grid_local(s->spec, s->Time_h[n], s->Time_h[n + 1]);
    fpde_status rc = launch_set_grid_point(s->w, n, s->Time_h[n + 1], tau, s->w.hLv[n], s->stream);
    if (rc != FPDE_OK) return rc;
    rc = enqueue_1d_step(s, n);
    if (rc != FPDE_OK) return rc;
    s->step = n + 1;
    rc = finalize_1d(s);
    if (rc != FPDE_OK) return rc;
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status fpde_solver_copy(fpde_solver* dst, const fpde_solver* src) {
    if (!dst || !src) { set_error("null solver"); return FPDE_ERR_ARG; }
    if (dst == src) return FPDE_OK;
    const fpde_spec &a = dst->spec, &b = src->spec;
    if (!a.var_tau || !b.var_tau || a.nodes != b.nodes || a.dims != 1 || dst->C != src->C) {
        set_error("copy needs two var_tau 1D solvers of the same nodes and chunks");
        return FPDE_ERR_ARG;
    }
    if (src->step > a.steps) { set_error("copy: destination capacity is below the source's steps"); return FPDE_ERR_ARG; }
    if (src->pending) { set_error("copy: the source has a pending solve"); return FPDE_ERR_STATE; }
    FPDE_HIP_TRY(hipStreamSynchronize(src->stream));
    FPDE_HIP_TRY(hipStreamSynchronize(dst->stream));
    const Plan& p = dst->plan;
    int n = src->step;
    FPDE_HIP_TRY(hipMemcpyAsync(dst->U, src->U, sizeof(double) * p.ld, hipMemcpyDeviceToDevice, dst->stream));
    if (n > 0)
        FPDE_HIP_TRY(hipMemcpyAsync(dst->D, src->D, sizeof(double) * (size_t)n * p.ld, hipMemcpyDeviceToDevice,
                                    dst->stream));
    FPDE_HIP_TRY(hipMemcpyAsync(dst->uend, src->uend, sizeof(double) * 2 * 2 * dst->C, hipMemcpyDeviceToDevice,
                                dst->stream));
    size_t np = (size_t)std::min(a.steps, b.steps);
    FPDE_HIP_TRY(hipMemcpyAsync(dst->w.Time, src->w.Time, sizeof(double) * (np + 2), hipMemcpyDeviceToDevice, dst->stream));
    FPDE_HIP_TRY(hipMemcpyAsync(dst->w.taus, src->w.taus, sizeof(double) * (np + 1), hipMemcpyDeviceToDevice, dst->stream));
    FPDE_HIP_TRY(hipMemcpyAsync(dst->w.Lv, src->w.Lv, sizeof(double) * np, hipMemcpyDeviceToDevice, dst->stream));
    std::copy(src->Time_h.begin(), src->Time_h.begin() + np + 2, dst->Time_h.begin());
    std::copy(src->taus_h.begin(), src->taus_h.begin() + np + 1, dst->taus_h.begin());
    std::copy(src->w.hLv.begin(), src->w.hLv.begin() + np, dst->w.hLv.begin());
    FPDE_HIP_TRY(hipStreamSynchronize(dst->stream));
    dst->step = n;
    dst->pending = false;
    dst->last_k = -1;
    dst->utop1d = src->utop1d;
    dst->initialized = src->initialized;
    return FPDE_OK;
}

fpde_status fpde_varT_step(fpde_solver* s0, fpde_solver* s2, fpde_solver* s3, double* tau, double eps,
                           double tau_max, double* err_out) {
    if (!s0 || !s2 || !s3 || !tau) { set_error("null argument"); return FPDE_ERR_ARG; }
    const int64_t nn = s0->plan.nodes;
    std::vector<double> u0((size_t)nn), u2((size_t)nn);
    // solve_T_half_step_and_check (:8778-8793) and solution_difference (:8723-8749)
    auto check = [&](double t, double* err) -> fpde_status {
        fpde_status rc = fpde_solver_copy(s3, s0);
        if (rc == FPDE_OK) rc = fpde_solver_advance_tau(s0, t);
        if (rc == FPDE_OK) rc = fpde_solver_advance_tau(s2, t / 2.0);
        if (rc == FPDE_OK) rc = fpde_solver_advance_tau(s2, t / 2.0);
        if (rc == FPDE_OK) rc = fpde_solver_get_solution(s0, u0.data());
        if (rc == FPDE_OK) rc = fpde_solver_get_solution(s2, u2.data());
        if (rc != FPDE_OK) return rc;
        double e = 0.0;
        for (int64_t i = 0; i < nn; i++) e += (u0[i] - u2[i]) * (u0[i] - u2[i]);
        *err = e;
        return FPDE_OK;
    };
    auto restore = [&]() -> fpde_status {
        fpde_status rc = fpde_solver_copy(s0, s3);
        if (rc == FPDE_OK) rc = fpde_solver_copy(s2, s3);
        return rc;
    };
    double t = *tau, err = 0.0;
    fpde_status rc = check(t, &err);
    if (rc != FPDE_OK) return rc;
    if (err > eps) {
        // decrease the step until the check passes (:8985-9004)
        const double t_first = t;
        while (err > eps) {
            if ((rc = restore()) != FPDE_OK) return rc;
            t /= 2.0;
            if ((rc = check(t, &err)) != FPDE_OK) return rc;
            if (t < 1e-10) {
                // the reference then takes one more step of the first length
                // (:8996-9000, with init_tau an int there); here of t_first
                t = t_first;
                if ((rc = fpde_solver_advance_tau(s0, t)) != FPDE_OK) return rc;
                break;
            }
        }
    } else if (t < tau_max) {
        // increase the step while the check passes (:9006-9021); the step of
        // the last check is kept, also when it failed or passed tau_max
        while (err <= eps) {
            if ((rc = restore()) != FPDE_OK) return rc;
            t *= 2.0;
            if ((rc = check(t, &err)) != FPDE_OK) return rc;
            if (t > tau_max) break;
        }
    }
    *tau = t;
    if (err_out) *err_out = err;
    return FPDE_OK;
}

fpde_status fpde_solver_sync(fpde_solver* s) {
    if (!s) { set_error("null solver"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->side) FPDE_HIP_TRY(hipStreamSynchronize(s->side));
    ev_collect(s);
    return check_barrier(s);
}

int fpde_solver_step(const fpde_solver* s) { return s ? s->step : -1; }

int fpde_solver_query(fpde_solver* s) {
    if (!s) return -1;
    hipError_t e = hipStreamQuery(s->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    set_error(hipGetErrorString(e));
    return -1;
}

fpde_status fpde_solver_get_solution(fpde_solver* s, double* u_out) {
    if (!s || !u_out) { set_error("null argument"); return FPDE_ERR_ARG; }
    if (s->spec.dims == 1) {
        fpde_status rc = finalize_1d(s);
        if (rc != FPDE_OK) return rc;
    }
    FPDE_HIP_TRY(hipMemcpyAsync(u_out, s->U, sizeof(double) * s->plan.nodes, hipMemcpyDeviceToHost, s->stream));
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    ev_collect(s);
    return check_barrier(s);
}

fpde_status fpde_solver_write_results(fpde_solver* s, const char* path_txt, const char* path_3d) {
    if (!s || !path_txt) { set_error("null argument"); return FPDE_ERR_ARG; }
    const fpde_spec& sp = s->spec;
    std::vector<double> u((size_t)s->plan.nodes);
    fpde_status rc = fpde_solver_get_solution(s, u.data());
    if (rc != FPDE_OK) return rc;
    FILE* f = std::fopen(path_txt, "a");
    if (!f) { set_error(std::string("cannot open ") + path_txt); return FPDE_ERR_ARG; }
    const double t = s->step * sp.tau, h = s->h;
    // wetness() with cv_form 4 (fpde_hpc_1d.cpp:2493-2526) at P = z - u
    auto theta = [](double P) {
        const double n = kVgmN, s0 = kVgmS0, s1 = kVgmS1, a = kVgmA;
        P *= 100.0;
        if (P <= 0.0) return s1;
        return s0 + (s1 - s0) / std::pow(1.0 + std::pow(a * P, n), 1.0 - 1.0 / n);
    };
    const int n = sp.nodes;
    // the reference's centre line idx(i, _M/2, K/2) with _M = K = n - 1 intervals
    const int c = (n - 1) / 2;
    for (int i = 0; i < n; i++) {
        double z = i * h;
        if (sp.dims == 1) {
            std::fprintf(f, "%g %g %g %g %g %g \n", t, z, u[i], 0.0, 0.0, theta(z - u[i]));
        } else {
            size_t id = (size_t)i + (size_t)n * ((size_t)c + (size_t)n * (size_t)c);
            std::fprintf(f, "%g %g %g %g %g %g %g %g \n", t, z, c * h, c * h, u[id], 0.0, 0.0, theta(z - u[id]));
        }
    }
    std::fclose(f);
    if (sp.dims == 3 && path_3d) {
        // b_U as the reference stores it: (N+2)^3 values for N = n - 1
        // intervals per axis, idx(i,j,k) = i + (N+2) j + (N+2)^2 k (depth i,
        // x j, y k; :1181), nodes 0..N plus one halo slot per axis, which no
        // sweep writes and keeps its initial value (here: the initial field
        // at the nearest node)
        const size_t m = (size_t)n + 1;
        std::vector<double> b(m * m * m);
        for (size_t k = 0; k < m; k++)
            for (size_t j = 0; j < m; j++)
                for (size_t i = 0; i < m; i++) {
                    const bool in = i < (size_t)n && j < (size_t)n && k < (size_t)n;
                    const size_t ii = std::min(i, (size_t)n - 1), jj = std::min(j, (size_t)n - 1),
                                 kk = std::min(k, (size_t)n - 1);
                    const size_t id = ii + (size_t)n * (jj + (size_t)n * kk);
                    b[i + m * (j + m * k)] = in ? u[id] : (s->u0_h.empty() ? 0.0 : s->u0_h[id]);
                }
        FILE* g = std::fopen(path_3d, "ab");
        if (!g) { set_error(std::string("cannot open ") + path_3d); return FPDE_ERR_ARG; }
        size_t w = std::fwrite(b.data(), sizeof(double), b.size(), g);
        std::fclose(g);
        if (w != b.size()) { set_error("short write to the 3D results file"); return FPDE_ERR_ARG; }
    }
    return FPDE_OK;
}

fpde_status fpde_solver_get_weights(fpde_solver* s, double* local, double* omega, double* corr) {
    if (!s) { set_error("null solver"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    if (local) *local = s->w.local;
    if (omega) FPDE_HIP_TRY(hipMemcpy(omega, s->w.omega, sizeof(double) * s->w.M, hipMemcpyDeviceToHost));
    if (corr) FPDE_HIP_TRY(hipMemcpy(corr, s->w.corr, sizeof(double) * 2 * s->w.M, hipMemcpyDeviceToHost));
    return FPDE_OK;
}

fpde_status fpde_solver_set_profiling(fpde_solver* s, int mask) {
    if (!s) { set_error("null solver"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    ev_collect(s);
    size_t keep = s->stats.device_bytes;
    s->stats = fpde_stats{};          // counters restart with the new mask
    s->stats.device_bytes = keep;
    if (s->err) FPDE_HIP_TRY(hipMemcpy(&s->picard_dev_base, s->err + 1, sizeof(int), hipMemcpyDeviceToHost));
    s->prof = mask;
    return FPDE_OK;
}

fpde_status fpde_solver_get_stats(fpde_solver* s, fpde_stats* out) {
    if (!s || !out) { set_error("null argument"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipStreamSynchronize(s->stream));
    ev_collect(s);
    *out = s->stats;
    if (s->err) {                     // cooperative launches count their Picard solves on the device
        int dev = 0;
        FPDE_HIP_TRY(hipMemcpy(&dev, s->err + 1, sizeof(int), hipMemcpyDeviceToHost));
        out->picard_iters_total += dev - s->picard_dev_base;
    }
    return FPDE_OK;
}

// --- workload initial conditions (same draws as oracle/fpde_oracle.c) ------
fpde_status fpde_init_1d(int nodes, double length, uint64_t seed, double* u) {
    if (!u || nodes < 2) { set_error("bad arguments"); return FPDE_ERR_ARG; }
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> amp(0.0, 0.05), ph(0.0, 2.0 * M_PI);
    double a[4], f[4];
    for (int k = 0; k < 4; k++) { a[k] = amp(eng); f[k] = ph(eng); }
    double h = length / (nodes - 1);
    for (int i = 0; i < nodes; i++) {
        double z = i * h, v = -1.0;
        for (int k = 0; k < 4; k++) v += a[k] * std::sin((k + 1) * M_PI * z + f[k]);
        u[i] = v;
    }
    return FPDE_OK;
}

fpde_status fpde_init_3d(int n, double length, uint64_t seed, double* u) {
    if (!u || n < 2) { set_error("bad arguments"); return FPDE_ERR_ARG; }
    std::mt19937_64 eng(seed);
    std::uniform_real_distribution<double> cd(0.2, 0.8), sd(0.05, 0.15), ad(0.0, 0.05);
    double c[8][3], sg[8], A[8];
    for (int b = 0; b < 8; b++) {
        for (int d = 0; d < 3; d++) c[b][d] = cd(eng) * length;
        sg[b] = sd(eng) * length;
        A[b] = ad(eng);
    }
    double h = length / (n - 1);
    for (int y = 0; y < n; y++)
        for (int x = 0; x < n; x++)
            for (int z = 0; z < n; z++) {
                double pz = z * h, px = x * h, py = y * h, v = -1.0;
                for (int b = 0; b < 8; b++) {
                    double dx = px - c[b][0], dy = py - c[b][1], dz = pz - c[b][2];
                    v += A[b] * std::exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sg[b] * sg[b]));
                }
                u[z + (size_t)n * (x + (size_t)n * y)] = v;
            }
    return FPDE_OK;
}

double fpde_ksat_for_fourier(double u_top, double h, double tau, double fo) {
    double K, iC;
    vgm_coefs(u_top, 0.0, 1.0, 0, K, iC);
    return fo * h * h / (tau * K * iC);
}

}  // extern "C"
