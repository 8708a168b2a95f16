// Device layer behind the OpenCL_* facade (include/fpde_opencl_facade.h).
//
// Replaces opencl_cuda_base.cpp (OpenCL flavour :20-458, CUDA/NVRTC flavour
// :459-826) for the reference's GPU history sum.  There is no run-time
// compilation: kernels are precompiled for gfx950 and bound by name; the
// program object only records the #define constants the reference prints
// into its kernel text (sum_alg, OS, BS: fpde_hpc_1d.cpp:1771-1775).
// Errors are returned as codes (the reference prints and calls exit(0),
// opencl_cuda_base.cpp:12-17); sizes are 64-bit.
#include "fpde_internal.h"

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

using namespace fpde;

struct fpde_cl_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct fpde_cl_buf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

struct fpde_cl_prog {
    fpde_cl_ctx* ctx = nullptr;
    int sum_alg = 0;
    int64_t os = 0;
    int bs = 16;
    // fr_do.cpp's program (kernels Fill, DOC, Sum, eq_F, eq_Solve) and its
    // element type, told apart from fpde_hpc_1d.cpp's by the kernel names
    bool frdo = false;
    bool frdo_double = true;
};

struct fpde_cl_kernel {
    fpde_cl_prog* prog = nullptr;
    std::string name;
    struct Arg {
        bool set = false;
        bool is_buf = false;
        fpde_cl_buf* buf = nullptr;
        std::vector<unsigned char> bytes;
    };
    std::vector<Arg> args;
};

namespace {

// value of "#define NAME <int>" in a kernel source text, or dflt
int64_t parse_define(const char* src, const char* name, int64_t dflt) {
    std::string s(src ? src : "");
    std::string key = std::string("#define ") + name;
    size_t p = 0;
    while ((p = s.find(key, p)) != std::string::npos) {
        size_t q = p + key.size();
        if (q < s.size() && (s[q] == ' ' || s[q] == '\t')) {
            return std::strtoll(s.c_str() + q, nullptr, 10);
        }
        p = q;
    }
    return dflt;
}

fpde_status arg_double(const fpde_cl_kernel* k, int idx, double* out) {
    if (idx >= (int)k->args.size() || !k->args[idx].set || k->args[idx].is_buf) {
        set_error("kernel argument " + std::to_string(idx) + " not set (scalar expected)");
        return FPDE_ERR_ARG;
    }
    const auto& b = k->args[idx].bytes;
    if (b.size() == sizeof(double)) { std::memcpy(out, b.data(), sizeof(double)); return FPDE_OK; }
    if (b.size() == sizeof(float)) { float f; std::memcpy(&f, b.data(), sizeof(float)); *out = f; return FPDE_OK; }
    set_error("kernel argument " + std::to_string(idx) + ": unsupported scalar size");
    return FPDE_ERR_ARG;
}

fpde_status arg_int(const fpde_cl_kernel* k, int idx, int64_t* out) {
    if (idx >= (int)k->args.size() || !k->args[idx].set || k->args[idx].is_buf) {
        set_error("kernel argument " + std::to_string(idx) + " not set (scalar expected)");
        return FPDE_ERR_ARG;
    }
    const auto& b = k->args[idx].bytes;
    if (b.size() == 4) { int32_t v; std::memcpy(&v, b.data(), 4); *out = v; return FPDE_OK; }
    if (b.size() == 8) { int64_t v; std::memcpy(&v, b.data(), 8); *out = v; return FPDE_OK; }
    set_error("kernel argument " + std::to_string(idx) + ": unsupported integer size");
    return FPDE_ERR_ARG;
}

fpde_status arg_buf(const fpde_cl_kernel* k, int idx, fpde_cl_buf** out) {
    if (idx >= (int)k->args.size() || !k->args[idx].set || !k->args[idx].is_buf) {
        set_error("kernel argument " + std::to_string(idx) + " not set (buffer expected)");
        return FPDE_ERR_ARG;
    }
    *out = k->args[idx].buf;
    return FPDE_OK;
}

}  // namespace

extern "C" {

fpde_status fpde_cl_ctx_create(int device, fpde_cl_ctx** out) {
    if (!out) { set_error("null out"); return FPDE_ERR_ARG; }
    int n = 0;
    FPDE_HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) { set_error("no such device"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipSetDevice(device));
    auto* c = new fpde_cl_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; set_error("hipStreamCreate failed"); return FPDE_ERR_HIP; }
    *out = c;
    return FPDE_OK;
}

fpde_status fpde_cl_ctx_destroy(fpde_cl_ctx* ctx) {
    if (!ctx) return FPDE_OK;
    hipStreamSynchronize(ctx->stream);
    hipStreamDestroy(ctx->stream);
    delete ctx;
    return FPDE_OK;
}

fpde_status fpde_cl_buf_alloc(fpde_cl_ctx* ctx, size_t bytes, const void* host_init, fpde_cl_buf** out) {
    if (!ctx || !out) { set_error("null argument"); return FPDE_ERR_ARG; }
    auto* b = new fpde_cl_buf;
    b->bytes = bytes;
    hipError_t e = hipMalloc(&b->ptr, bytes ? bytes : 8);
    if (e != hipSuccess) { delete b; set_error(std::string("hipMalloc: ") + hipGetErrorString(e)); return FPDE_ERR_NOMEM; }
    if (host_init && bytes) {
        e = hipMemcpy(b->ptr, host_init, bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) { hipFree(b->ptr); delete b; set_error("initial copy failed"); return FPDE_ERR_HIP; }
    }
    *out = b;
    return FPDE_OK;
}

fpde_status fpde_cl_buf_free(fpde_cl_buf* buf) {
    if (!buf) return FPDE_OK;
    hipFree(buf->ptr);
    delete buf;
    return FPDE_OK;
}

size_t fpde_cl_buf_size(const fpde_cl_buf* buf) { return buf ? buf->bytes : 0; }
void* fpde_cl_buf_device_ptr(const fpde_cl_buf* buf) { return buf ? buf->ptr : nullptr; }

fpde_status fpde_cl_buf_write(fpde_cl_ctx* ctx, fpde_cl_buf* buf, size_t offset, size_t bytes, const void* src) {
    if (!ctx || !buf || !src) { set_error("null argument"); return FPDE_ERR_ARG; }
    if (bytes == 0) bytes = buf->bytes - offset;
    if (offset > buf->bytes || bytes > buf->bytes - offset) { set_error("write out of buffer bounds"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipMemcpyAsync((char*)buf->ptr + offset, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FPDE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FPDE_OK;
}

fpde_status fpde_cl_buf_read(fpde_cl_ctx* ctx, fpde_cl_buf* buf, size_t offset, size_t bytes, void* dst) {
    if (!ctx || !buf || !dst) { set_error("null argument"); return FPDE_ERR_ARG; }
    if (bytes == 0) bytes = buf->bytes - offset;
    if (offset > buf->bytes || bytes > buf->bytes - offset) { set_error("read out of buffer bounds"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipMemcpyAsync(dst, (char*)buf->ptr + offset, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FPDE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FPDE_OK;
}

fpde_status fpde_cl_buf_copy(fpde_cl_ctx* ctx, fpde_cl_buf* dst, size_t dst_offset, fpde_cl_buf* src,
                             size_t src_offset, size_t bytes) {
    if (!ctx || !dst || !src) { set_error("null argument"); return FPDE_ERR_ARG; }
    if (dst_offset > dst->bytes || bytes > dst->bytes - dst_offset || src_offset > src->bytes ||
        bytes > src->bytes - src_offset) {
        set_error("copy out of buffer bounds");
        return FPDE_ERR_ARG;
    }
    FPDE_HIP_TRY(hipMemcpyAsync((char*)dst->ptr + dst_offset, (const char*)src->ptr + src_offset, bytes,
                                hipMemcpyDeviceToDevice, ctx->stream));
    FPDE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FPDE_OK;
}

fpde_status fpde_cl_program_create(fpde_cl_ctx* ctx, const char* source, fpde_cl_prog** out) {
    if (!ctx || !out) { set_error("null argument"); return FPDE_ERR_ARG; }
    auto* p = new fpde_cl_prog;
    p->ctx = ctx;
    p->sum_alg = (int)parse_define(source, "sum_alg", 0);
    p->os = parse_define(source, "OS", 0);
    p->bs = (int)parse_define(source, "BS", 16);
    const std::string src(source ? source : "");
    p->frdo = src.find("eq_Solve") != std::string::npos && src.find("DOC") != std::string::npos;
    // the text is instantiated for T = float or double (cuda_text_macro, fr_do.cpp:301-306)
    p->frdo_double = src.find("float *df") == std::string::npos && src.find("float* df") == std::string::npos;
    *out = p;
    return FPDE_OK;
}

fpde_status fpde_cl_program_destroy(fpde_cl_prog* prog) {
    delete prog;
    return FPDE_OK;
}

fpde_status fpde_cl_kernel_create(fpde_cl_prog* prog, const char* name, fpde_cl_kernel** out) {
    if (!prog || !name || !out) { set_error("null argument"); return FPDE_ERR_ARG; }
    static const char* frdo_names[] = {"Fill", "DOC", "Sum", "eq_F", "eq_Solve"};
    bool known = std::strcmp(name, "Sum") == 0;
    if (prog->frdo)
        for (const char* nm : frdo_names) known = known || std::strcmp(name, nm) == 0;
    if (!known) {
        set_error(std::string("unknown kernel '") + name +
                  (prog->frdo ? "' (available: Fill, DOC, Sum, eq_F, eq_Solve)" : "' (available: Sum)"));
        return FPDE_ERR_NOTFOUND;
    }
    auto* k = new fpde_cl_kernel;
    k->prog = prog;
    k->name = name;
    k->args.resize(15);
    *out = k;
    return FPDE_OK;
}

fpde_status fpde_cl_kernel_destroy(fpde_cl_kernel* k) {
    delete k;
    return FPDE_OK;
}

fpde_status fpde_cl_kernel_set_arg(fpde_cl_kernel* k, int idx, size_t size, const void* value) {
    if (!k || idx < 0 || !value) { set_error("bad argument"); return FPDE_ERR_ARG; }
    if (idx >= (int)k->args.size()) k->args.resize(idx + 1);
    auto& a = k->args[idx];
    a.set = true;
    a.is_buf = false;
    a.buf = nullptr;
    a.bytes.assign((const unsigned char*)value, (const unsigned char*)value + size);
    return FPDE_OK;
}

fpde_status fpde_cl_kernel_set_buffer_arg(fpde_cl_kernel* k, int idx, fpde_cl_buf* buf) {
    if (!k || idx < 0 || !buf) { set_error("bad argument"); return FPDE_ERR_ARG; }
    if (idx >= (int)k->args.size()) k->args.resize(idx + 1);
    auto& a = k->args[idx];
    a.set = true;
    a.is_buf = true;
    a.buf = buf;
    a.bytes.clear();
    return FPDE_OK;
}

}  // extern "C"

namespace {

// fr_do.cpp's kernels, with the argument ABI of its host code (every scalar
// a double; DO_Cder_CUDA :373-419, eq_one_step_cuda :793-818): decode the
// arguments, check every index range the kernel will touch against the
// buffers, then launch.  The global size is the reference's work-item count.
fpde_status frdo_launch(fpde_cl_ctx* ctx, fpde_cl_kernel* k, int64_t nth) {
    FrdoLaunch L;
    L.is_double = k->prog->frdo_double;
    const int64_t es = L.is_double ? 8 : 4;
    fpde_status rc;
    double v[12];
    fpde_cl_buf* bf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    auto dbl = [&](std::initializer_list<int> idx) {
        int q = 0;
        for (int i : idx)
            if ((rc = arg_double(k, i, &v[q++])) != FPDE_OK) return false;
        return true;
    };
    auto bufs = [&](std::initializer_list<int> idx) {
        int q = 0;
        for (int i : idx) {
            if ((rc = arg_buf(k, i, &bf[q])) != FPDE_OK) return false;
            L.b[q] = bf[q]->ptr;
            q++;
        }
        return true;
    };
    auto len = [&](int q) { return (int64_t)(bf[q]->bytes / es); };
    auto bad = [&](const char* what) { set_error(std::string("fr_do ") + k->name + ": " + what); return FPDE_ERR_ARG; };
    const std::string& nm = k->name;
    if (nm == "Fill") {                                   // (tau, na, df, bkk) :373-379
        if (!dbl({0, 1}) || !bufs({2, 3})) return rc;
        L.kernel = FRDO_FILL;
        L.tau = v[0];
        L.na = (int)v[1];
        L.n = nth;
        if (L.na < 1 || len(0) < L.n || len(1) < L.n * L.na) return bad("df needs ntau and bkk ntau*na values");
    } else if (nm == "DOC") {   // (mode, ntau, tau, na, t0, j0, maxt, use_f, F, df, bkk, wa, do1, gs, total_ntau) :383-403
        if (!dbl({0, 1, 2, 3, 4, 5, 6, 7, 14}) || !bufs({8, 9, 10, 12, 13})) return rc;
        L.kernel = FRDO_DOC;
        L.ntau = (int)v[1];
        L.tau = v[2];
        L.na = (int)v[3];
        L.t0 = (int)v[4];
        L.j0 = (int)v[5];
        L.maxt = (int)v[6];
        L.use_f = (int)v[7];
        L.tot = (int64_t)v[8];
        if (L.na < 1 || nth < L.na) return bad("the work items must cover na orders");
        L.nf = (int)(nth / L.na);
        // bf: 0 F, 1 df, 2 bkk, 3 do1, 4 gs
        if (L.t0 < 0 || L.ntau < 0) return bad("negative time range");
        if (len(2) < (int64_t)L.ntau * L.na || len(3) < (int64_t)L.nf * L.na * L.ntau || len(4) < L.na)
            return bad("bkk, do1 or gs too small");
        if (L.use_f && (L.ntau > L.tot || len(0) < (int64_t)L.nf * L.tot)) return bad("F too small for the levels");
        if (!L.use_f && len(1) < L.ntau) return bad("df too small");
    } else if (nm == "Sum" ) {   // (ntau, ntau2, t0, na, append, do1, do2, wa) :405-415
        if (!dbl({0, 1, 2, 3, 4}) || !bufs({5, 6, 7})) return rc;
        L.kernel = FRDO_SUM;
        L.ntau = (int)v[0];
        L.ntau2 = (int)v[1];
        L.t0 = (int)v[2];
        L.na = (int)v[3];
        L.append = (int)v[4];
        const int64_t nt = L.ntau - L.t0;
        if (nt <= 0) return FPDE_OK;
        L.nf = (int)(nth / nt);
        if (L.na < 1 || L.ntau > L.ntau2 || len(0) < (int64_t)L.nf * L.na * L.ntau || len(1) < (int64_t)L.nf * L.ntau2 ||
            len(2) < L.na)
            return bad("do1, do2 or wa too small");
    } else if (nm == "eq_F") {   // (nx, eq_Nt, D, b0, tau, h, j, C, DF, F, ex) :793-808
        if (!dbl({0, 1, 2, 3, 4, 5, 6, 10}) || !bufs({7, 8, 9})) return rc;
        L.kernel = FRDO_EQ_F;
        L.nx = (int)v[0];
        L.Nt = (int)v[1];
        L.D = v[2];
        L.b0 = v[3];
        L.tau = v[4];
        L.h = v[5];
        L.j = (int)v[6];
        L.ex = v[7];
        if (nth < L.nx) L.nx = (int)nth;
        if (L.nx < 1 || L.j < 0 || L.j + 1 >= L.Nt || len(0) < (int64_t)L.nx * L.Nt || len(1) < (int64_t)L.nx * L.Nt ||
            len(2) < L.nx)
            return bad("C and DF need nx*Nt values, F nx, and j+1 < Nt");
    } else if (nm == "eq_Solve") {   // (t, nt, nx, A, C, inv, F) :810-818
        if (!dbl({0, 1, 2, 3}) || !bufs({4, 5, 6})) return rc;
        L.kernel = FRDO_EQ_SOLVE;
        L.j = (int)v[0];
        L.Nt = (int)v[1];
        L.nx = (int)v[2];
        L.A = v[3];
        if (L.nx < 1 || L.j < 0 || L.j >= L.Nt || len(0) < (int64_t)L.nx * L.Nt || len(1) < (int64_t)L.nx * L.nx ||
            len(2) < L.nx)
            return bad("C needs nx*nt values, inv nx*nx, F nx, and t < nt");
    } else {
        set_error("unknown fr_do kernel");
        return FPDE_ERR_NOTFOUND;
    }
    rc = launch_frdo(L, ctx->stream);
    if (rc != FPDE_OK) return rc;
    // the reference synchronises after every launch (opencl_cuda_base.cpp:710-713)
    FPDE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FPDE_OK;
}

}  // namespace

extern "C" {

// "Sum" with the argument ABI of fr_ocl_call_sum (fpde_hpc_1d.cpp:1851-1868):
// 0 S, 1 old, 2 int old_size, 3 tau, 4 Htmp, 5 row2_precalc, 6 kb_3_cache,
// 7 double sum_param, 8 gtj, 9 gt, 10 alpha, 11 int tstep.  The kernels of
// fr_do.cpp's program go to frdo_launch.
fpde_status fpde_cl_launch(fpde_cl_ctx* ctx, fpde_cl_kernel* k, int ndims, const size_t* gdims, const size_t* ldims) {
    (void)ldims;
    if (!ctx || !k || ndims < 1 || !gdims) { set_error("bad launch arguments"); return FPDE_ERR_ARG; }
    if (k->prog->frdo) return frdo_launch(ctx, k, (int64_t)gdims[0]);
    fpde_cl_buf *S, *old, *tau, *kb;
    int64_t old_size;
    double sum_param;
    fpde_status rc;
    if ((rc = arg_buf(k, 0, &S)) || (rc = arg_buf(k, 1, &old)) || (rc = arg_int(k, 2, &old_size)) ||
        (rc = arg_buf(k, 3, &tau)) || (rc = arg_buf(k, 6, &kb)) || (rc = arg_double(k, 7, &sum_param)))
        return rc;
    int sum_alg = k->prog->sum_alg;
    int64_t os = k->prog->os;
    int64_t nodes = (int64_t)gdims[0];
    if (os <= 0) os = nodes;
    // bounds: the launch may be rounded up past OS (fpde_hpc_1d.cpp:1873-1874)
    if (nodes > os) nodes = os;
    if ((int64_t)(S->bytes / sizeof(double)) < nodes) { set_error("S buffer too small"); return FPDE_ERR_ARG; }
    if (sum_alg == 2) {
        // series moments: 4 Htmp [nodes][P], 5 row2_precalc [P], 8 gtj, 9 gt,
        // 10 alpha, 11 tstep (fr_ocl_call_sum, fpde_hpc_1d.cpp:1837-1868)
        fpde_cl_buf *H, *coef;
        double gtj, gt, alpha;
        int64_t tstep;
        if ((rc = arg_buf(k, 4, &H)) || (rc = arg_buf(k, 5, &coef)) || (rc = arg_double(k, 8, &gtj)) ||
            (rc = arg_double(k, 9, &gt)) || (rc = arg_double(k, 10, &alpha)) || (rc = arg_int(k, 11, &tstep)))
            return rc;
        if (!(sum_param >= 1.0 && sum_param <= (double)(1 << 20))) { set_error("sum_alg 2: sum_param (terms) must be >= 1"); return FPDE_ERR_ARG; }
        int P = (int)sum_param;
        if ((int64_t)(H->bytes / sizeof(double)) < nodes * P || (int64_t)(coef->bytes / sizeof(double)) < P) {
            set_error("sum_alg 2: Htmp needs nodes*sum_param and row2_precalc sum_param doubles");
            return FPDE_ERR_ARG;
        }
        if (old_size >= 2 && ((int64_t)(old->bytes / sizeof(double)) < 2 * os || tstep < 0 ||
                              (int64_t)(tau->bytes / sizeof(double)) <= tstep)) {
            set_error("sum_alg 2: old needs two levels and tau must hold tstep");
            return FPDE_ERR_ARG;
        }
        rc = launch_ref_series((double*)S->ptr, (const double*)old->ptr, old_size, os, nodes, (const double*)tau->ptr,
                               (int)tstep, (double*)H->ptr, (const double*)coef->ptr, P, gtj, gt, alpha, ctx->stream);
        if (rc != FPDE_OK) return rc;
        FPDE_HIP_TRY(hipGetLastError());
        return FPDE_OK;
    }
    if (old_size >= 1 && (int64_t)(old->bytes / sizeof(double)) < old_size * os) { set_error("old buffer too small"); return FPDE_ERR_ARG; }
    if (old_size >= 2 && ((int64_t)(kb->bytes / sizeof(double)) < old_size - 1 || (int64_t)(tau->bytes / sizeof(double)) < old_size - 1)) {
        set_error("kb/tau buffers too small");
        return FPDE_ERR_ARG;
    }
    rc = launch_ref_sum((double*)S->ptr, (const double*)old->ptr, old_size, os, nodes, (const double*)tau->ptr,
                        (const double*)kb->ptr, sum_alg, sum_param, k->prog->bs, ctx->stream);
    if (rc != FPDE_OK) return rc;
    FPDE_HIP_TRY(hipGetLastError());
    return FPDE_OK;
}

fpde_status fpde_cl_finish(fpde_cl_ctx* ctx) {
    if (!ctx) { set_error("null ctx"); return FPDE_ERR_ARG; }
    FPDE_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return FPDE_OK;
}

// ---------------------------------------------------------------------------
fpde_status fpde_reference_sum(double* S, const double* old, int64_t old_size, int64_t os, int64_t nodes,
                               const double* tau, const double* kb, int sum_alg, double sum_param) {
    if (!S || !old || (old_size >= 2 && (!tau || !kb)) || nodes < 0 || os < nodes) {
        set_error("bad arguments");
        return FPDE_ERR_ARG;
    }
    if (sum_alg != 0 && sum_alg != 1) { set_error("sum_alg must be 0 or 1"); return FPDE_ERR_ARG; }
    fpde_status rc = launch_ref_sum(S, old, old_size, os, nodes, tau, kb, sum_alg, sum_param, 16, 0);
    if (rc != FPDE_OK) return rc;
    FPDE_HIP_TRY(hipGetLastError());
    FPDE_HIP_TRY(hipDeviceSynchronize());
    return FPDE_OK;
}

fpde_status fpde_history_sum_f64(double* S, const double* D, int64_t n, int64_t nodes, int64_t ld,
                                 const double* omega) {
    if (!S || !D || !omega || n < 0 || nodes < 0 || ld < nodes || (ld & 1) || (nodes > (1 << 30))) {
        set_error("bad arguments (ld must be even and >= nodes)");
        return FPDE_ERR_ARG;
    }
    if (n == 0) {
        FPDE_HIP_TRY(hipMemset(S, 0, sizeof(double) * nodes));
        FPDE_HIP_TRY(hipDeviceSynchronize());
        return FPDE_OK;
    }
    // partials then one combine pass
    double* part = nullptr;
    FPDE_HIP_TRY(hipMalloc(&part, sizeof(double) * 64 * (size_t)ld));
    int splits = 0;
    fpde_status rc = launch_hist_plain(D, ld, (int)nodes, (int)n, (int)n, omega, part, &splits, 0);
    if (rc == FPDE_OK) {
        // combine: S = sum of partial rows (no near levels, no corrections)
        rc = launch_hist_combine(S, part, splits, ld, 0, D, ld, nodes, 0, 0, omega, 0, 0, 0, nullptr, nullptr, 1.0, 0);
    }
    hipError_t e = hipDeviceSynchronize();
    hipFree(part);
    if (rc != FPDE_OK) return rc;
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return FPDE_ERR_HIP; }
    return FPDE_OK;
}

fpde_status fpde_space_sum(double* V, const double* F, int64_t nodes, double h, double gamma, int weight_mode,
                           double series_eps) {
    if (!V || !F || nodes < 4 || nodes > (1 << 24) || !(h > 0.0) || !(gamma > 0.0 && gamma < 1.0) ||
        (weight_mode != 0 && weight_mode != 1) || !(series_eps > 0.0)) {
        set_error("bad arguments (nodes >= 4, h > 0, 0 < gamma < 1, weight_mode 0 or 1, series_eps > 0)");
        return FPDE_ERR_ARG;
    }
    fpde_spec sp{};
    sp.dims = 1;
    sp.nodes = (int)nodes;
    sp.gamma = gamma;
    sp.weight_mode = weight_mode;
    sp.series_eps = series_eps;
    const int64_t ldp = (nodes + kSpTile - 1) / kSpTile * kSpTile;
    const size_t chunks = (size_t)space_chunks((int)(ldp / kSpTile) - 1);
    DeviceWeights w;
    double *Fp = nullptr, *P = nullptr;
    fpde_status rc = FPDE_OK;
    auto grab = [&](double** p, size_t n) {
        if (rc == FPDE_OK && hipMalloc(p, sizeof(double) * n) != hipSuccess) {
            set_error("hipMalloc failed");
            rc = FPDE_ERR_NOMEM;
        }
    };
    grab(&w.sp_omega, (size_t)nodes);
    grab(&w.sp_corr, 2 * (size_t)nodes);
    grab(&Fp, (size_t)ldp);
    grab(&P, chunks * (size_t)ldp);
    if (rc == FPDE_OK) {
        // F on the interior rows, zero elsewhere (calc_v reads F[1..i-1])
        hipError_t e = hipMemset(Fp, 0, sizeof(double) * ldp);
        if (e == hipSuccess)
            e = hipMemcpy(Fp + 1, F + 1, sizeof(double) * (size_t)(nodes - 2), hipMemcpyDeviceToDevice);
        if (e != hipSuccess) { set_error(hipGetErrorString(e)); rc = FPDE_ERR_HIP; }
    }
    if (rc == FPDE_OK) rc = launch_space_weights(sp, w, h, 0);
    if (rc == FPDE_OK) rc = launch_space_apply(Fp, (int)nodes, w.sp_omega, P, ldp, 0);
    if (rc == FPDE_OK) rc = launch_space_reduce(V, P, Fp, w.sp_corr, (int)nodes, ldp, 0);
    hipError_t e = hipDeviceSynchronize();
    hipFree(w.sp_omega);
    hipFree(w.sp_corr);
    hipFree(Fp);
    hipFree(P);
    if (rc != FPDE_OK) return rc;
    if (e != hipSuccess) { set_error(hipGetErrorString(e)); return FPDE_ERR_HIP; }
    return FPDE_OK;
}

fpde_status fpde_history_block_gemm(int mode, const double* D, int64_t ld, int64_t nodes, int64_t levels, int n0,
                                    int tc, int block, const double* omega, double* P) {
    if (!D || !omega || !P || nodes <= 0 || ld < nodes || ld % 64 || tc < 0 || tc > n0 || tc > levels ||
        (block != 16 && block != 32 && block != 64 && block != 128) || mode < 1 || mode > 4) {
        set_error("bad arguments");
        return FPDE_ERR_ARG;
    }
    bool split = mode >= 2;
    if (split && tc % kGroup) { set_error("split modes need tc % 32 == 0"); return FPDE_ERR_ARG; }
    int ks = gemm_max_splits(nodes);
    double* part = nullptr;
    void* comp[3] = {nullptr, nullptr, nullptr};
    float* scale = nullptr;
    void* afrag = nullptr;
    double* ainv = nullptr;
    fpde_status rc = FPDE_OK;
    int splits = 0;
    int64_t lv = ((levels + kGroup - 1) / kGroup) * kGroup;
    hipError_t e = hipMalloc(&part, sizeof(double) * (size_t)ks * block * ld);
    if (e == hipSuccess && split) {
        for (int c = 0; c < 3 && e == hipSuccess; c++) e = hipMalloc(&comp[c], 2 * (size_t)lv * ld);
        if (e == hipSuccess) e = hipMalloc(&scale, sizeof(float) * (size_t)(lv / kGroup) * ld);
        if (e == hipSuccess) e = hipMalloc(&afrag, split_afrag_bytes(mode, block, (int)(lv / kGroup)));
        if (e == hipSuccess) e = hipMalloc(&ainv, sizeof(double) * (size_t)(lv / kGroup) * (block / 16));
    }
    if (e != hipSuccess) { rc = FPDE_ERR_NOMEM; set_error("hipMalloc failed"); }
    if (rc == FPDE_OK && split) rc = launch_pack_split(mode, D, ld, (int)nodes, 0, tc / kGroup, comp[0], comp[1], comp[2], scale, 0);
    if (rc == FPDE_OK) {
        if (split) rc = launch_gemm_split(mode, comp[0], comp[1], comp[2], scale, ld, (int)nodes, n0, tc, block, omega, nullptr, 0, afrag, ainv, part, &splits, 0);
        else rc = launch_gemm_f64(D, ld, (int)nodes, n0, tc, block, omega, part, &splits, 0);
    }
    if (rc == FPDE_OK) {
        // P[r] = sum over K splits (combine kernel, one row at a time)
        for (int r = 0; r < block && rc == FPDE_OK; r++)
            rc = launch_hist_combine(P + (size_t)r * ld, part, splits, (int64_t)block * ld, (int64_t)r * ld, D, ld,
                                     nodes, 0, 0, omega, 0, 0, 0, nullptr, nullptr, 1.0, 0);
    }
    e = hipDeviceSynchronize();
    if (rc == FPDE_OK && e != hipSuccess) { rc = FPDE_ERR_HIP; set_error(hipGetErrorString(e)); }
    hipFree(part);
    for (int c = 0; c < 3; c++) hipFree(comp[c]);
    hipFree(scale);
    hipFree(afrag);
    hipFree(ainv);
    return rc;
}

}  // extern "C"
