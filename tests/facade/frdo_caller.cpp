// A reference-style caller of the OpenCL_* facade for fr_do.cpp's GPU code:
// the distributed-order equation solve of eq_solve_cuda (fr_do.cpp:824-863)
// with DO_CUDA_init / DO_Cder_CUDA (:310-430) and eq_one_step_cuda
// (:781-823), issuing the program's kernels Fill, DOC, Sum, eq_F and
// eq_Solve by name with their argument ABI (every scalar a double).  The
// element type is T = double or float, as the reference builds both.
//   frdo_caller <nx> <na> <tblock> <eNt> <example> <power> <d|f> <out.bin>
// writes the levels C [nx][nt] (node-major, as T) to out.bin and prints
// "nx na tblock nt err": the RMS error against example 0's exact solution
// 100 t^4 sin(2 pi x) at the last level.  Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "fpde_opencl_facade.h"

namespace {

int round_up(int v, int m) { return m * ((v / m) + ((v % m) ? 1 : 0)); }

template <typename T>
struct Run {
    int nx, na, tblock, nt, ex;
    double power, D = 2.0, max_x, max_t;
    T h, tau, Aa, b0, Ss;
    OpenCL_program* prg = nullptr;
    OpenCL_commandqueue* queue = nullptr;
    OpenCL_kernel *kFill = nullptr, *kDO = nullptr, *kSum = nullptr, *kF = nullptr, *kSolve = nullptr;
    OpenCL_buffer *bDF, *bBKK, *bWA, *bDOF, *bgs, *bDO1, *bC, *bF, *binv;

    double w(double a) const { return ex == 0 ? std::tgamma(5.0 - a) : std::pow(a, power) * (1.0 + power); }

    // the program: only its kernel names and element type matter to the facade
    std::string text() const {
        const char* t = sizeof(T) == 8 ? "double" : "float";
        std::string s;
        s += std::string("extern \"C\" __global__ void Fill(double,double,") + t + " *df," + t + " *bkk);\n";
        s += std::string("extern \"C\" __global__ void DOC(/* 8 doubles, 6 buffers, total_ntau */);\n");
        s += std::string("extern \"C\" __global__ void Sum(/* 5 doubles, 3 buffers */);\n");
        s += std::string("extern \"C\" __global__ void eq_F(/* 7 doubles, 3 buffers, ex */);\n");
        s += std::string("extern \"C\" __global__ void eq_Solve(/* 4 doubles, 3 buffers */);\n");
        return s;
    }

    void init() {          // DO_CUDA_init with use_f = nx grid points, ntau = nt
        prg = new OpenCL_program(1);
        queue = prg->create_queue(0, 0);
        OpenCL_prg* prog = prg->create_program(text().c_str());
        kFill = prg->create_kernel(prog, "Fill");
        kDO = prg->create_kernel(prog, "DOC");
        kSum = prg->create_kernel(prog, "Sum");
        kF = prg->create_kernel(prog, "eq_F");
        kSolve = prg->create_kernel(prog, "eq_Solve");
        bDF = prg->create_buffer(CL_MEM_READ_WRITE, nt * sizeof(T), NULL);
        bC = prg->create_buffer(CL_MEM_READ_WRITE, (size_t)nx * nt * sizeof(T), NULL);
        bF = prg->create_buffer(CL_MEM_READ_WRITE, nx * sizeof(T), NULL);
        bBKK = prg->create_buffer(CL_MEM_READ_WRITE, (size_t)nt * na * sizeof(T), NULL);
        bDOF = prg->create_buffer(CL_MEM_READ_WRITE, (size_t)nx * nt * sizeof(T), NULL);
        bDO1 = prg->create_buffer(CL_MEM_READ_WRITE, (size_t)nx * nt * na * sizeof(T), NULL);
        bWA = prg->create_buffer(CL_MEM_READ_WRITE, na * sizeof(T), NULL);
        bgs = prg->create_buffer(CL_MEM_READ_WRITE, na * sizeof(T), NULL);
        std::vector<T> gs(na), WA(na);
        for (int i = 0; i < na; i++) {
            gs[i] = (T)std::tgamma(2.0 - (i / (double)(na - 1)));
            WA[i] = (T)(w(i / (double)(na - 1)) * (1.0 / (na - 1)));
        }
        queue->EnqueueWriteBuffer(bgs, gs.data(), 0, na * sizeof(T));
        queue->EnqueueWriteBuffer(bWA, WA.data(), 0, na * sizeof(T));
    }

    // DO_Cder_CUDA(mode, ntau, tau, na, append, F=C, use_f=nx, maxt, t0, j0, w, init_ntau=nt)
    void do_cder(int mode, int ntau, int append, int maxt, int t0, int j0) {
        if (mode == 1) ntau = round_up(ntau, 16);
        double dtau = tau, dna = na;
        kFill->SetArg(0, sizeof(double), &dtau);
        kFill->SetArg(1, sizeof(double), &dna);
        kFill->SetBufferArg(bDF, 2);
        kFill->SetBufferArg(bBKK, 3);
        size_t nth = ntau, lsize = 1;
        queue->ExecuteKernel(kFill, 1, &nth, &lsize);
        double a[15] = {(double)mode, (double)ntau, dtau, dna, (double)t0, (double)j0, (double)maxt, (double)nx};
        double dtnt = nt;
        for (int i = 0; i < 8; i++) kDO->SetArg(i, sizeof(double), &a[i]);
        kDO->SetBufferArg(bC, 8);
        kDO->SetBufferArg(bDF, 9);
        kDO->SetBufferArg(bBKK, 10);
        kDO->SetBufferArg(bWA, 11);
        kDO->SetBufferArg(bDO1, 12);
        kDO->SetBufferArg(bgs, 13);
        kDO->SetArg(14, sizeof(double), &dtnt);
        nth = (size_t)na * nx;
        lsize = 32;
        queue->ExecuteKernel(kDO, 1, &nth, &lsize);
        double s[5] = {(double)ntau, (double)nt, (double)t0, dna, (double)append};
        for (int i = 0; i < 5; i++) kSum->SetArg(i, sizeof(double), &s[i]);
        kSum->SetBufferArg(bDO1, 5);
        kSum->SetBufferArg(bDOF, 6);
        kSum->SetBufferArg(bWA, 7);
        nth = (size_t)(ntau - t0) * nx;
        lsize = 1;
        queue->ExecuteKernel(kSum, 1, &nth, &lsize);
    }

    void one_step(int j) {  // eq_one_step_cuda: right-hand side of level j from level j-1, then the solve
        double f[8] = {(double)nx, (double)nt, D, (double)b0, (double)tau, (double)h, (double)(j - 1), (double)ex};
        for (int i = 0; i < 7; i++) kF->SetArg(i, sizeof(double), &f[i]);
        kF->SetBufferArg(bC, 7);
        kF->SetBufferArg(bDOF, 8);
        kF->SetBufferArg(bF, 9);
        kF->SetArg(10, sizeof(double), &f[7]);
        size_t nth = nx, lsize = 1;
        queue->ExecuteKernel(kF, 1, &nth, &lsize);
        double q[4] = {(double)j, (double)nt, (double)nx, (double)Aa};
        for (int i = 0; i < 4; i++) kSolve->SetArg(i, sizeof(double), &q[i]);
        kSolve->SetBufferArg(bC, 4);
        kSolve->SetBufferArg(binv, 5);
        kSolve->SetBufferArg(bF, 6);
        lsize = 32;
        queue->ExecuteKernel(kSolve, 1, &nth, &lsize);
    }

    // eq_init's inverse of the Crank-Nicolson matrix, scaled by Aa: interior
    // block = inverse of tridiag(1, -Ss/Aa, 1) (one Thomas solve per column
    // here), first and last rows the identity rows times Aa, first and last
    // columns the coupling of the boundary values (:696-719)
    std::vector<T> inverse() const {
        const int m = nx - 2;
        const double d = -(double)Ss / (double)Aa;
        std::vector<T> inv((size_t)nx * nx, (T)0);
        std::vector<double> cp(m), x(m), rhs(m);
        for (int c = 0; c < m; c++) {
            for (int r = 0; r < m; r++) rhs[r] = (r == c) ? 1.0 : 0.0;
            cp[0] = 1.0 / d;
            x[0] = rhs[0] / d;
            for (int r = 1; r < m; r++) {
                double den = d - cp[r - 1];
                cp[r] = 1.0 / den;
                x[r] = (rhs[r] - x[r - 1]) / den;
            }
            for (int r = m - 2; r >= 0; r--) x[r] -= cp[r] * x[r + 1];
            for (int r = 0; r < m; r++) inv[(size_t)(r + 1) * nx + (c + 1)] = (T)x[r];
        }
        inv[0] = inv[(size_t)(nx - 1) * nx + nx - 1] = Aa;
        for (int i = 1; i < nx - 1; i++) {
            inv[(size_t)i * nx] = -inv[(size_t)i * nx + 1] * Aa;
            inv[(size_t)i * nx + nx - 1] = -inv[(size_t)i * nx + nx - 2] * Aa;
        }
        return inv;
    }

    int solve(const char* out_path) {
        h = (T)(max_x / (nx - 1));
        tau = (T)(max_t / nt);
        Aa = (T)(0.5 * D / ((double)h * h));
        init();
        // eq_b0: trapezoid of WA bk(tau, 0, a) / Gamma(2-a), bk(tau,0,a) = tau^(1-a) (:687-690)
        double bsum = 0.0;
        for (int a = 0; a < na - 1; a++) {
            double a0 = a / (double)(na - 1), a1 = (a + 1) / (double)(na - 1);
            bsum += 0.5 * (w(a0) / (na - 1) * std::pow((double)tau, 1 - a0) / std::tgamma(2 - a0) +
                           w(a1) / (na - 1) * std::pow((double)tau, 1 - a1) / std::tgamma(2 - a1));
        }
        b0 = (T)bsum;
        Ss = (T)(D / ((double)h * h) + (double)b0 / (double)tau);
        std::vector<T> C((size_t)nx * nt, (T)0);
        for (int i = 0; i < nx; i++) C[(size_t)i * nt] = (T)(ex == 1 ? (i * (double)h) * (1.0 - i * (double)h) : 0.0);
        queue->EnqueueWriteBuffer(bC, C.data(), 0, C.size() * sizeof(T));
        std::vector<T> inv = inverse();
        binv = prg->create_buffer(CL_MEM_READ_WRITE, inv.size() * sizeof(T), NULL);
        queue->EnqueueWriteBuffer(binv, inv.data(), 0, inv.size() * sizeof(T));
        for (int i = 0; i < nt; i += tblock) {
            // the block's history before its start, on tensor cores in the reference (mode 1)
            do_cder(1, i + tblock, 0, (i == 0) ? i : i - 1, i, 0);
            for (int j = (i == 0) ? 1 : 0; j < tblock; j++) {
                one_step(i + j);
                if (j != tblock - 1) do_cder(0, i + j + 2, 1, i + j, i + j + 1, i - 1);
            }
        }
        queue->EnqueueBuffer(bC, C.data(), 0, C.size() * sizeof(T));
        queue->Finish();
        double err = 0.0;
        const int last = nt - 1;
        for (int x = 0; x < nx; x++) {
            double a = 100 * std::pow(last * (double)tau, 4) * std::sin(2 * M_PI * x * (double)h);
            double d = (double)C[(size_t)x * nt + last] - a;
            err += d * d;
        }
        err = std::sqrt(err / nx);
        std::printf("%d %d %d %d %.17g\n", nx, na, tblock, nt, err);
        FILE* f = std::fopen(out_path, "wb");
        if (!f) return 1;
        size_t n = std::fwrite(C.data(), sizeof(T), C.size(), f);
        std::fclose(f);
        delete prg;
        return n == C.size() ? 0 : 1;
    }
};

template <typename T>
int run(int argc, char** argv) {
    Run<T> r;
    // eq_solve_cuda rounds the block and the node and order counts (:827-830);
    // eq_init takes nt = (eNt / tblock) * tblock (:669)
    r.tblock = round_up(std::atoi(argv[3]), 16);
    r.na = round_up(std::atoi(argv[2]), 32);
    r.nx = round_up(std::atoi(argv[1]), 32);
    r.nt = (std::atoi(argv[4]) / r.tblock) * r.tblock;
    r.ex = std::atoi(argv[5]);
    r.power = std::atof(argv[6]);
    r.max_x = r.ex == 0 ? 0.5 : 1.0;
    r.max_t = r.ex == 0 ? 0.5 : 1.0;
    return r.solve(argv[8]);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 9) {
        std::fprintf(stderr, "usage: frdo_caller nx na tblock eNt example power d|f out.bin\n");
        return 2;
    }
    return argv[7][0] == 'f' ? run<float>(argc, argv) : run<double>(argc, argv);
}
