// Allocation-free plain FP64 history sum: fpde_history_sum_f64 with its
// t-split partial rows in a file-scope device array instead of a hipMalloc'd
// buffer, for callers that cannot allocate (tight serving loops, harnesses
// that forbid hipMalloc).  Synchronous, as fpde_history_sum_f64 is.
//
// The array `scratch` holds kNoAllocRows partial rows of at most
// kNoAllocMaxLd doubles (32 MiB); every GPU the library runs on holds its own
// copy until the process ends.  In host code `scratch` is only a symbol: its
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and kept in g_scratch[device]; kernels get that
// address as an argument.  All calls on one GPU share its array: calls that
// overlap on one GPU (other threads, other streams) race on it.  Each call
// rewrites every partial row it reads, so nothing has to be cleared between
// calls.
#include "fpde_internal.h"

#include <climits>
#include <mutex>

using namespace fpde;

namespace {

constexpr int64_t kNoAllocMaxLd = 65536;   // padded level stride, doubles
constexpr int64_t kNoAllocRows = kPlainMaxSplits;

__device__ __align__(256) double scratch[kNoAllocRows * kNoAllocMaxLd];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// Device address of this device's `scratch` (looked up on the first call on
// each device, with that device current).
fpde_status scratch_of_current_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { set_error(std::string("hipGetDevice: ") + hipGetErrorString(e)); return FPDE_ERR_HIP; }
    if (dev < 0 || dev >= kMaxDevices) { set_error("device index out of range"); return FPDE_ERR_ARG; }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return FPDE_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return FPDE_OK;
}

}  // namespace

extern "C" fpde_status fpde_history_sum_f64_noalloc(double* S, const double* D, int64_t n, int64_t nodes,
                                                    int64_t ld, const double* omega) {
    if (!S || !D || !omega || n < 0 || n > INT_MAX || nodes < 0 || ld < nodes || (ld & 1)) {
        set_error("bad arguments (ld must be even and >= nodes)");
        return FPDE_ERR_ARG;
    }
    if (ld > kNoAllocMaxLd) {
        set_error("fpde_history_sum_f64_noalloc: ld must be at most 65536 (its partial rows live in a fixed "
                  "32 MiB device array); use fpde_history_sum_f64 for larger levels");
        return FPDE_ERR_ARG;
    }
    double* part = nullptr;
    fpde_status rc = scratch_of_current_device(&part);
    if (rc != FPDE_OK) return rc;
    // the same kernels as fpde_history_sum_f64: partial rows of <= 64
    // t-chunks, then one combine pass (with no levels, n = 0, the combine
    // writes S = 0)
    if (hist_plain_plan((int)nodes, (int)n, nullptr) > kNoAllocRows) {
        set_error("partial rows exceed the scratch array");
        return FPDE_ERR_STATE;
    }
    int splits = 0;
    rc = launch_hist_plain(D, ld, (int)nodes, (int)n, (int)n, omega, part, &splits, 0);
    if (rc != FPDE_OK) return rc;
    rc = launch_hist_combine(S, part, splits, ld, 0, D, ld, nodes, 0, 0, omega, 0, 0, 0, nullptr, nullptr, 1.0, 0);
    if (rc != FPDE_OK) return rc;
    FPDE_HIP_TRY(hipGetLastError());
    FPDE_HIP_TRY(hipDeviceSynchronize());
    return FPDE_OK;
}
