/*
 * Drop-in replacement of the reference's GPU abstraction (opencl_cuda_class.h:4-67)
 * backed by libfpde_hip.so (include/fpde_hip.h, group 1).
 *
 * Same class names and member functions as the reference, so the call sites
 * of H_solver::init_opencl / fr_ocl_check_and_resize / fr_ocl_call_sum
 * (fpde_hpc_1d.cpp:1715-1883) compile unchanged against this header:
 *   OpenCL_program(int gpu), create_queue, create_program(source),
 *   create_kernel(prg,"Sum"), create_buffer(flags,size,host_ptr),
 *   OpenCL_kernel::SetArg / SetBufferArg,
 *   OpenCL_commandqueue::ExecuteKernel / EnqueueWriteBuffer / EnqueueBuffer / Finish.
 * Differences, by design:
 *   - sizes and offsets are size_t (the reference's int limits buffers to 2 GiB,
 *     opencl_cuda_class.h:8,44-45);
 *   - no run-time compilation: create_program() only reads the #define
 *     constants (sum_alg, OS, BS) the reference prints into its kernel text;
 *     kernels are precompiled for gfx950 and bound by name;
 *   - a failing call prints the library's message and exits with status 1
 *     (the reference exits with status 0 on errors, opencl_cuda_base.cpp:12-17);
 *   - deleting a buffer unregisters it from its program, so the
 *     delete-then-program-destructor sequence of fr_ocl_check_and_resize
 *     (fpde_hpc_1d.cpp:1802-1812) frees each buffer once;
 *   - every transfer is blocking, as in the reference; ExecuteKernel is
 *     asynchronous until Finish() or the next transfer.
 * Only the cl_* typedefs and CL_MEM_* constants are taken from <CL/cl.h>; no
 * OpenCL runtime is used.
 */
#ifndef FPDE_OPENCL_FACADE_H
#define FPDE_OPENCL_FACADE_H

#ifndef CL_TARGET_OPENCL_VERSION
#define CL_TARGET_OPENCL_VERSION 120
#endif
#include <CL/cl.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fpde_hip.h"

namespace fpde_facade_detail {
inline void die(fpde_status rc, const char* what) {
    if (rc != FPDE_OK) {
        std::fprintf(stderr, "fpde: %s failed (%d): %s\n", what, (int)rc, fpde_last_error());
        std::exit(1);
    }
}
}  // namespace fpde_facade_detail

class OpenCL_program;

class OpenCL_buffer {
public:
    fpde_cl_buf* buffer;
    size_t size;
    OpenCL_program* owner;
    OpenCL_buffer(fpde_cl_ctx* ct, cl_mem_flags flags, size_t sz, void* mem, OpenCL_program* prg)
        : buffer(nullptr), size(sz), owner(prg) {
        (void)flags;   // read/write access is not enforced; COPY_HOST_PTR <=> mem != NULL
        fpde_facade_detail::die(fpde_cl_buf_alloc(ct, sz, mem, &buffer), "create_buffer");
    }
    ~OpenCL_buffer();
};

class OpenCL_prg {
public:
    fpde_cl_prog* hProgram;
    OpenCL_prg(fpde_cl_ctx* ct, const char* source) : hProgram(nullptr) {
        fpde_facade_detail::die(fpde_cl_program_create(ct, source, &hProgram), "create_program");
    }
    ~OpenCL_prg() { fpde_cl_program_destroy(hProgram); }
};

class OpenCL_kernel {
public:
    fpde_cl_kernel* hKernel;
    OpenCL_kernel(OpenCL_prg* hprog, const char* name) : hKernel(nullptr) {
        fpde_facade_detail::die(fpde_cl_kernel_create(hprog->hProgram, name, &hKernel), "create_kernel");
    }
    ~OpenCL_kernel() { fpde_cl_kernel_destroy(hKernel); }
    cl_int SetBufferArg(OpenCL_buffer* buf, int idx) {
        return fpde_cl_kernel_set_buffer_arg(hKernel, idx, buf->buffer) == FPDE_OK ? CL_SUCCESS : CL_INVALID_ARG_VALUE;
    }
    cl_int SetArg(int idx, int size, void* buf) {
        return fpde_cl_kernel_set_arg(hKernel, idx, (size_t)size, buf) == FPDE_OK ? CL_SUCCESS : CL_INVALID_ARG_VALUE;
    }
};

class OpenCL_commandqueue {
    fpde_cl_ctx* ctx;
public:
    explicit OpenCL_commandqueue(fpde_cl_ctx* c) : ctx(c) {}
    cl_event ExecuteKernel(OpenCL_kernel* krnl, int ndims, size_t* dims, size_t* ldims = NULL,
                           int nwaitkernels = 0, cl_event* evs = 0, int noevent = 0) {
        (void)nwaitkernels; (void)evs; (void)noevent;   // one in-order stream per queue
        fpde_facade_detail::die(fpde_cl_launch(ctx, krnl->hKernel, ndims, dims, ldims), "ExecuteKernel");
        return (cl_event)0;
    }
    /* read: device buffer -> host (opencl_cuda_base.cpp:210-219); size 0 = whole buffer */
    cl_int EnqueueBuffer(OpenCL_buffer* b, void* mem, size_t offset = 0, size_t size = 0) {
        fpde_facade_detail::die(fpde_cl_buf_read(ctx, b->buffer, offset, size, mem), "EnqueueBuffer");
        return CL_SUCCESS;
    }
    /* write: host -> device buffer (opencl_cuda_base.cpp:220-228) */
    cl_int EnqueueWriteBuffer(OpenCL_buffer* b, void* mem, size_t offset = 0, size_t size = 0) {
        fpde_facade_detail::die(fpde_cl_buf_write(ctx, b->buffer, offset, size, mem), "EnqueueWriteBuffer");
        return CL_SUCCESS;
    }
    void ReleaseEvent(cl_event) {}
    void Finish() { fpde_facade_detail::die(fpde_cl_finish(ctx), "Finish"); }
};

class OpenCL_program {
    fpde_cl_ctx* ctx;
    std::vector<OpenCL_buffer*> buffers;
    std::vector<OpenCL_kernel*> kernels;
    std::vector<OpenCL_commandqueue*> queues;
    std::vector<OpenCL_prg*> progs;
public:
    /* gpu != 0 selects a GPU device (the reference falls back to any device, :356-406) */
    explicit OpenCL_program(int gpu) : ctx(nullptr) {
        (void)gpu;
        fpde_facade_detail::die(fpde_cl_ctx_create(0, &ctx), "OpenCL_program");
    }
    ~OpenCL_program() {
        while (!buffers.empty()) delete buffers.back();   // each unregisters itself
        for (auto* k : kernels) delete k;
        for (auto* q : queues) delete q;
        for (auto* p : progs) delete p;
        fpde_cl_ctx_destroy(ctx);
    }
    OpenCL_buffer* create_buffer(cl_mem_flags flags, size_t sz, void* mem) {
        auto* b = new OpenCL_buffer(ctx, flags, sz, mem, this);
        buffers.push_back(b);
        return b;
    }
    OpenCL_prg* create_program(const char* source) {
        auto* p = new OpenCL_prg(ctx, source);
        progs.push_back(p);
        return p;
    }
    OpenCL_kernel* create_kernel(OpenCL_prg* hprog, const char* name) {
        auto* k = new OpenCL_kernel(hprog, name);
        kernels.push_back(k);
        return k;
    }
    OpenCL_commandqueue* create_queue(int device, cl_command_queue_properties props = 0) {
        (void)device; (void)props;
        auto* q = new OpenCL_commandqueue(ctx);
        queues.push_back(q);
        return q;
    }
    int get_ndevices() { return fpde_device_count(); }
    cl_device_type get_device_type(int device) { (void)device; return CL_DEVICE_TYPE_GPU; }
    void forget(OpenCL_buffer* b) { buffers.erase(std::remove(buffers.begin(), buffers.end(), b), buffers.end()); }
};

inline OpenCL_buffer::~OpenCL_buffer() {
    fpde_cl_buf_free(buffer);
    if (owner) owner->forget(this);
}

#endif
