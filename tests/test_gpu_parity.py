"""GPU parity: the HIP path (through the C ABI) against the FP64 CPU oracle.

Tolerances (relative = max|u - u_ref| / max|u_ref|):
  FP64 paths (plain streaming, FP64 MFMA GEMM)     <= 1e-10
  split bf16/fp16 MFMA paths                       <= 1e-8
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_F64 = 1e-10
TOL_MIXED = 1e-8


def rel_err(u, ref):
    return float(np.max(np.abs(u - ref)) / np.max(np.abs(ref)))


class Dev:
    """Small helper over the fpde_cl_* buffer API (device memory for direct kernel calls)."""

    def __init__(self, L):
        self.L = L
        self.ctx = ctypes.c_void_p()
        assert L.fpde_cl_ctx_create(0, ctypes.byref(self.ctx)) == 0, L.fpde_last_error()
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = ctypes.c_void_p()
        assert self.L.fpde_cl_buf_alloc(self.ctx, a.nbytes, a.ctypes.data, ctypes.byref(b)) == 0
        self.bufs.append(b)
        return b

    def zeros(self, n):
        return self.put(np.zeros(n))

    def ptr(self, b):
        return self.L.fpde_cl_buf_device_ptr(b)

    def get(self, b, n):
        out = np.zeros(n)
        assert self.L.fpde_cl_buf_read(self.ctx, b, 0, n * 8, out.ctypes.data) == 0
        return out

    def close(self):
        for b in self.bufs:
            self.L.fpde_cl_buf_free(b)
        self.L.fpde_cl_ctx_destroy(self.ctx)


def test_history_sum_f64_matches_numpy(fpde):
    L = fpde.lib()
    rng = np.random.default_rng(7)
    for n, nodes in [(1, 10), (5, 130), (97, 1000), (700, 777)]:
        ld = ((nodes + 63) // 64) * 64
        D = np.zeros((n, ld))
        D[:, :nodes] = rng.standard_normal((n, nodes))
        omega = rng.random(n + 8)
        ref = np.zeros(nodes)
        for j in range(1, n + 1):
            ref += omega[j - 1] * D[n - j, :nodes]
        dv = Dev(L)
        try:
            bD, bw, bS = dv.put(D), dv.put(omega), dv.zeros(ld)
            rc = L.fpde_history_sum_f64(dv.ptr(bS), dv.ptr(bD), n, nodes, ld, dv.ptr(bw))
            assert rc == 0, L.fpde_last_error()
            S = dv.get(bS, nodes)
        finally:
            dv.close()
        assert np.max(np.abs(S - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_reference_sum_kernel_through_facade(fpde, oracle):
    """The reference 'Sum' ABI (fr_ocl_call_sum, fpde_hpc_1d.cpp:1851-1868) vs time_summing."""
    L = fpde.lib()
    rng = np.random.default_rng(11)
    nodes, levels = 333, 41
    OS = 384
    old = np.zeros((levels, OS))
    old[:, :nodes] = np.cumsum(rng.standard_normal((levels, nodes)) * 1e-2, axis=0)
    Time = oracle.time_grid(levels + 1, 1e-3)
    kkk0, kb = oracle.l1_weights(0.7, levels - 1, Time, weight_mode=1)
    tau = np.full(levels + 8, 1e-3)
    ref = oracle.time_summing(old[:, :nodes], kb, tau)
    src = b"#define sum_alg 0\n#define OS 384\n#define BS 16\n__kernel void Sum(...)"
    dv = Dev(L)
    try:
        bS, bold, btau, bK = dv.zeros(OS), dv.put(old), dv.put(tau), dv.put(np.concatenate([kb, np.zeros(8)]))
        bH, bC = dv.zeros(8), dv.zeros(8)
        prog, k = ctypes.c_void_p(), ctypes.c_void_p()
        assert L.fpde_cl_program_create(dv.ctx, src, ctypes.byref(prog)) == 0
        assert L.fpde_cl_kernel_create(prog, b"Sum", ctypes.byref(k)) == 0
        for idx, b in [(0, bS), (1, bold), (3, btau), (4, bH), (5, bC), (6, bK)]:
            assert L.fpde_cl_kernel_set_buffer_arg(k, idx, b) == 0
        iv = ctypes.c_int(levels)
        assert L.fpde_cl_kernel_set_arg(k, 2, 4, ctypes.byref(iv)) == 0
        for idx, v in [(7, 1e-5), (8, 0.0), (9, 0.0), (10, 0.7)]:
            dvv = ctypes.c_double(v)
            assert L.fpde_cl_kernel_set_arg(k, idx, 8, ctypes.byref(dvv)) == 0
        iv2 = ctypes.c_int(levels - 1)
        assert L.fpde_cl_kernel_set_arg(k, 11, 4, ctypes.byref(iv2)) == 0
        g = (ctypes.c_size_t * 1)(OS)
        l = (ctypes.c_size_t * 1)(16)
        assert L.fpde_cl_launch(dv.ctx, k, 1, g, l) == 0, L.fpde_last_error()
        assert L.fpde_cl_finish(dv.ctx) == 0
        S = dv.get(bS, nodes)
        L.fpde_cl_kernel_destroy(k)
        L.fpde_cl_program_destroy(prog)
    finally:
        dv.close()
    assert np.max(np.abs(S - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("ex,power", [(0, 1.5), (1, 1.5), (1, 0.5)])
@pytest.mark.parametrize("T", ["d", "f"])
def test_frdo_equation_through_facade(oracle, tmp_path, ex, power, T):
    """fr_do.cpp's GPU equation solve (eq_solve_cuda, :824-863) by a compiled
    C++ caller issuing the program's kernels Fill, DOC, Sum, eq_F and eq_Solve
    through the OpenCL_* facade, in double and float, against the oracle's
    restatement of eq_solve with the GPU text's source term."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "facade", "_build", "frdo_caller")
    out = tmp_path / "C.bin"
    nx, na, tb, ent = 64, 32, 32, 320
    r = subprocess.run([exe, str(nx), str(na), str(tb), str(ent), str(ex), str(power), T, str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    C = np.fromfile(out, dtype=np.float64 if T == "d" else np.float32).astype(np.float64).reshape(nx, -1)
    ref, err = oracle.doeq_solve(oracle.doeq_params(nx=nx, na=na, tblock=tb, nt=ent, example=ex, power=power,
                                                    gpu_pi=1))
    assert C.shape == ref.shape
    d = rel_err(C, ref)
    print(f"fr_do equation ex {ex} p {power} T={T}: rel diff {d:.2e}, RMS err vs exact {r.stdout.split()[-1]}")
    assert d <= (1e-11 if T == "d" else 2e-4)
    if ex == 0:
        assert abs(float(r.stdout.split()[-1]) - err) <= (1e-12 if T == "d" else 1e-4)


def test_reference_style_cpp_caller_through_facade():
    """The fr_ocl_call_sum call sequence compiled against fpde_opencl_facade.h."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "facade", "_build", "reference_style_caller")
    assert os.path.exists(exe), "build it with __graft_entry__.build()"
    for args in (["1000", "70"], ["37", "20"]):
        r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=60)
        print(r.stdout, r.stderr)
        assert r.returncode == 0


@pytest.mark.parametrize("mode,tol", [(1, 1e-13), (2, 2e-5), (3, 5e-7), (4, 2e-7)])
@pytest.mark.parametrize("block", [16, 64, 128])
def test_block_gemm_matches_numpy(fpde, mode, tol, block):
    """P[r,i] = sum_{t<tc} omega[n0+r-t-1] D[t,i]; tol is relative to sum |w||D|."""
    L = fpde.lib()
    rng = np.random.default_rng(mode * 100 + block)
    nodes, ld, tc, n0 = 300, 320, 224, 230
    levels = 256
    D = np.zeros((levels, ld))
    D[:tc, :nodes] = rng.standard_normal((tc, nodes)) * np.exp(rng.uniform(-8, 0, (tc, 1)))
    omega = np.zeros(n0 + block + 600)
    j = np.arange(1, n0 + block + 1)
    omega[: n0 + block] = (j + 1.0) ** 0.3 - j ** 0.3
    W = np.array([[omega[n0 + r - t - 1] for t in range(tc)] for r in range(block)])
    ref = W @ D[:tc, :nodes]
    scale = np.abs(W) @ np.abs(D[:tc, :nodes])
    dv = Dev(L)
    try:
        bD, bw, bP = dv.put(D), dv.put(omega), dv.zeros(block * ld)
        rc = L.fpde_history_block_gemm(mode, dv.ptr(bD), ld, nodes, levels, n0, tc, block, dv.ptr(bw), dv.ptr(bP))
        assert rc == 0, L.fpde_last_error()
        P = dv.get(bP, block * ld).reshape(block, ld)[:, :nodes]
    finally:
        dv.close()
    err = np.max(np.abs(P - ref) / scale)
    print(f"gemm mode {mode} block {block}: max err/scale {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("weight_mode", [0, 1])
@pytest.mark.parametrize("do", [None, (8, 0), (6, 1)])
def test_weights_match_oracle(fpde, oracle, weight_mode, do):
    M, tau = 300, 1e-3
    kw = {}
    if do:
        kw = dict(do_q=do[0], do_rule=do[1], do_a0=0.1, do_a1=1.0)
    spec = fpde.make_spec(50, M, tau, alpha=0.6, weight_mode=weight_mode, **kw)
    with fpde.Solver(spec) as s:
        local, omega, corr = s.weights()
    p = oracle.make_params(50, M, tau, 0.6, weight_mode=weight_mode, ksat=spec.ksat,
                           do_q=kw.get("do_q", 0), do_rule=kw.get("do_rule", 0))
    lo, om, co = oracle.folded_weights(p)
    assert abs(local - lo) <= 1e-13 * abs(lo)
    assert np.max(np.abs(omega - om) / np.abs(om)) <= 1e-10
    full = om[None, :] * 0  # noqa
    # compare full weights of the two oldest intervals (omega + corr)
    for n in range(1, M):
        for t in range(min(2, n)):
            a = omega[n - t - 1] + corr[n, t]
            b = om[n - t - 1] + co[n, t]
            assert abs(a - b) <= 1e-10 * abs(b)


def _oracle_params(oracle, spec, picard_tol=0.0):
    return oracle.make_params(spec.nodes, spec.steps, spec.tau, spec.alpha, ksat=spec.ksat,
                              coef_mode=spec.coef_mode, picard_max=spec.picard_max, picard_tol=picard_tol,
                              weight_mode=spec.weight_mode, series_eps=spec.series_eps, do_q=spec.do_q,
                              do_rule=spec.do_rule, do_a0=spec.do_a0, do_a1=spec.do_a1,
                              sum_alg=spec.sum_alg, sum_param=spec.sum_param)


def _oracle_1d(oracle, spec, u0, picard_tol=0.0):
    return oracle.solve_1d(_oracle_params(oracle, spec, picard_tol), u0)


CASES_1D = [
    # nodes, steps, alpha, picard, coef_mode
    (300, 200, 0.8, 1, 0),
    (517, 150, 0.7, 2, 0),
    (1000, 300, 0.5, 1, 0),
    (5, 90, 0.9, 1, 0),
    (64, 130, 0.3, 3, 1),
]


@pytest.mark.parametrize("case", CASES_1D)
@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_bf16", "split_f16", "split_bf16x3"])
def test_solve_1d_matches_oracle(fpde, oracle, case, mode):
    nodes, steps, alpha, picard, coef = case
    u0 = fpde.init_1d(nodes, 1000 + nodes)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=alpha, picard_max=picard, coef_mode=coef,
                          hist_mode=mode, block=64, u_top=u0[0])
    u = fpde.solve(spec, u0)
    ref = _oracle_1d(oracle, spec, u0)
    err = rel_err(u, ref)
    tol = TOL_F64 if mode in ("plain", "mfma_f64") else TOL_MIXED
    print(f"1D {case} {mode}: rel err {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_bf16"])
def test_edge_sizes_match_oracle(fpde, oracle, mode):
    """The smallest line (4 nodes: one chunk of two interior rows), a single
    step, lines that end one node past a 192-node chunk (the automatic size)
    or a 256-node one, and runs one step past a history block or a 32-level
    group; 3D at the smallest grid."""
    tol = TOL_F64 if mode in ("plain", "mfma_f64") else TOL_MIXED
    for nodes, steps in [(4, 1), (4, 3), (4, 70), (193, 66), (385, 65), (257, 1), (257, 66), (513, 65),
                         (256, 33)]:
        u0 = fpde.init_1d(nodes, 4000 + nodes + steps)
        spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.75, picard_max=3, picard_tol=1e-12, hist_mode=mode,
                              block=64, u_top=u0[0])
        err = rel_err(fpde.solve(spec, u0), _oracle_1d(oracle, spec, u0, picard_tol=1e-12))
        print(f"edge 1D N={nodes} M={steps} {mode}: rel err {err:.3e}", flush=True)
        assert err <= tol, (nodes, steps)
    for n, steps in [(4, 1), (4, 20)]:
        u0 = fpde.init_3d(n, 4100 + steps)
        spec = fpde.make_spec(n, steps, 1e-3, alpha=0.85, dims=3, hist_mode=mode, block=16, u_top=u0[0])
        ref = oracle.solve_3d(oracle.make_params(n, steps, 1e-3, 0.85, ksat=spec.ksat), u0)
        err = rel_err(fpde.solve(spec, u0), ref)
        print(f"edge 3D n={n} M={steps} {mode}: rel err {err:.3e}", flush=True)
        assert err <= tol, (n, steps)


@pytest.mark.parametrize("rows", [0, 32, 100, 192, 256])
def test_chunk_sizes_match_oracle(fpde, oracle, rows):
    """The partitioned line solve in chunks of 32 to 256 nodes (0: the
    automatic size), on a 2000-node line with Picard to tol 1e-12 in the
    cooperative block form: every chunking agrees with the oracle."""
    nodes, steps = 2000, 130
    u0 = fpde.init_1d(nodes, 4321)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.7, picard_max=3, picard_tol=1e-12, hist_mode="mfma_f64",
                          block=64, u_top=u0[0], chunk_rows=rows)
    err = rel_err(fpde.solve(spec, u0), _oracle_1d(oracle, spec, u0, picard_tol=1e-12))
    print(f"chunk rows {rows}: rel err {err:.3e}", flush=True)
    assert err <= TOL_F64


@pytest.mark.parametrize("tol", [1e-12, 1e-7, 3e-6])
@pytest.mark.parametrize("mode", ["plain", "split_f16"])
def test_picard_tolerance_exit_matches_oracle(fpde, oracle, tol, mode):
    """Picard on the nonlinear coefficient, at most 3 iterations, stopping once
    max|U^k - U^{k-1}| <= tol (the convergence test runs on the device)."""
    nodes, steps = 600, 120
    u0 = fpde.init_1d(nodes, 77)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.8, picard_max=3, picard_tol=tol, hist_mode=mode,
                          u_top=u0[0], fourier=40.0)
    u = fpde.solve(spec, u0)
    ref = _oracle_1d(oracle, spec, u0, picard_tol=tol)
    err = rel_err(u, ref)
    print(f"picard tol {tol} {mode}: rel err {err:.3e}")
    assert err <= (TOL_F64 if mode == "plain" else TOL_MIXED)


@pytest.mark.parametrize("tol", [0.0, 1e-5, 1e-7])
@pytest.mark.parametrize("picard", [2, 3])
def test_picard_chunk_halo_tight(fpde, oracle, tol, picard):
    """Stiff coefficients (mesh Fourier number 40) make every Picard iterate
    sensitive to the coefficient at the node left of each chunk: that node
    must see the same iterate (the extrapolated start, the converged iterate
    when the tolerance was met one launch late) as the chunk's own nodes.
    Checked to 1e-12 over four chunks."""
    nodes, steps = 600, 60
    u0 = fpde.init_1d(nodes, 78)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.8, picard_max=picard, picard_tol=tol, hist_mode="plain",
                          u_top=u0[0], fourier=40.0)
    u = fpde.solve(spec, u0)
    err = rel_err(u, _oracle_1d(oracle, spec, u0, picard_tol=tol))
    print(f"picard {picard} tol {tol}: rel err {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_bf16"])
@pytest.mark.parametrize("picard,tol", [(1, 0.0), (3, 1e-12), (3, 1e-5), (2, 1e-12), (3, 1e-15), (4, 1e-14)])
def test_cooperative_block_launch_equals_per_iteration_launch(fpde, mode, picard, tol):
    """The two 1D launch forms (one cooperative launch per history block with
    grid barriers between Picard iterations; one launch per iteration) run
    the same arithmetic: results agree to rounding, also when advanced in
    pieces that cut blocks.  The cooperative form ends a step tentatively
    when its last Picard iteration is due and decides convergence with the
    next step's first rows; tolerances that are not met (1e-15, 1e-14, and
    picard_max 2) exercise the rollback to the skipped iteration.  The
    reduced systems are solved by different
    PCR variants (one wave for small C per iteration, the workgroup in the
    cooperative form), and at mesh Fourier number 10 their rounding grows to
    ~2e-13 over 150 steps, hence the 1e-12 bound.  (The split forms place
    their GEMM cut one level apart in the two forms, so they agree to the
    split precision.)"""
    nodes, steps = 900, 150
    u0 = fpde.init_1d(nodes, 91)
    out = []
    for step_mode in (0, 1):
        spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.75, hist_mode=mode, block=32, picard_max=picard,
                              picard_tol=tol, u_top=u0[0], fourier=10.0, step_mode=step_mode)
        with fpde.Solver(spec) as s:
            s.set_initial(u0)
            for chunk in (7, 50, 1000):
                s.advance(chunk)
            out.append(s.solution())
            if step_mode == 0:             # Picard solves counted by the cooperative launches themselves
                its = s.stats()["picard_iters_total"]
                assert steps <= its <= picard * steps and (picard > 1 or its == steps), its
    err = rel_err(out[0], out[1])
    print(f"coop vs per-iteration {mode} picard {picard} tol {tol}: {err:.3e}")
    assert err <= (TOL_MIXED if mode.startswith("split") else 1e-12)


@pytest.mark.parametrize("block", [16, 32, 128])
def test_solve_1d_blocks(fpde, oracle, block):
    nodes, steps = 700, 260
    u0 = fpde.init_1d(nodes, 5)
    for mode in ("mfma_f64", "split_f16"):
        spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.75, hist_mode=mode, block=block, u_top=u0[0])
        u = fpde.solve(spec, u0)
        ref = _oracle_1d(oracle, spec, u0)
        assert rel_err(u, ref) <= (TOL_F64 if mode == "mfma_f64" else TOL_MIXED)


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_f16"])
def test_distributed_order_matches_oracle(fpde, oracle, mode):
    nodes, steps = 400, 160
    u0 = fpde.init_1d(nodes, 1002)
    for rule, q in [(0, 8), (1, 6)]:
        spec = fpde.make_spec(nodes, steps, 1e-3, do_q=q, do_rule=rule, do_a0=0.1, do_a1=1.0,
                              hist_mode=mode, u_top=u0[0])
        u = fpde.solve(spec, u0)
        ref = _oracle_1d(oracle, spec, u0)
        assert rel_err(u, ref) <= (TOL_MIXED if mode.startswith("split") else TOL_F64)


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_f16"])
def test_solve_3d_matches_oracle(fpde, oracle, mode):
    n, steps = 12, 70
    u0 = fpde.init_3d(n, 1003)
    spec = fpde.make_spec(n, steps, 1e-3, alpha=0.85, dims=3, hist_mode=mode, block=16, u_top=u0[0])
    u = fpde.solve(spec, u0)
    p = oracle.make_params(n, steps, 1e-3, 0.85, ksat=spec.ksat)
    ref = oracle.solve_3d(p, u0)
    err = rel_err(u, ref)
    print(f"3D {mode}: rel err {err:.3e}")
    assert err <= (TOL_MIXED if mode.startswith("split") else TOL_F64)


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_bf16", "split_f16", "split_bf16x3"])
@pytest.mark.parametrize("fk,fp", [(3, 0.9), (1, 0.6), (2, 0.6)])
def test_generalized_kernel_1d_matches_oracle(fpde, oracle, mode, fk, fp):
    """Generalized Caputo kernel g(t) = t^0.9 (FK 3 FP 0.9, the reference
    scripts' setting, gpu_test.sh:140-141), sqrt(t) (FK 1) and t^2 (FK 2):
    exact weight integrals per step row, every history form (the GEMM forms in
    cooperative block launches), Picard <= 3, across GEMM blocks of 32."""
    nodes, steps, alpha = 300, 150, 0.7
    u0 = fpde.init_1d(nodes, 1300)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=alpha, picard_max=3, picard_tol=1e-12, hist_mode=mode,
                          block=32, u_top=u0[0], weight_mode=0, fk=fk, fp=fp)
    u = fpde.solve(spec, u0)

    def ref(k):
        return oracle.solve_1d(oracle.make_params(nodes, steps, 1e-3, alpha, ksat=spec.ksat, picard_max=3,
                                                  picard_tol=1e-12, weight_mode=0, fk=k, fp=fp), u0)
    r = ref(fk)
    tol = TOL_F64 if mode in ("plain", "mfma_f64") else TOL_MIXED
    assert rel_err(ref(0), r) >= 1e3 * tol               # the kernel changes the solution
    err = rel_err(u, r)
    print(f"1D FK {fk} FP {fp} {mode}: rel err {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_f16"])
def test_generalized_kernel_3d_and_distributed_order(fpde, oracle, mode):
    """FK 3 FP 0.9 in the 3D LOD step (12^3, the per-step weight rows feed the
    history combine and the local coefficient every sweep), and with
    distributed order in 1D."""
    n, steps = 12, 40
    u0 = fpde.init_3d(n, 1003)
    spec = fpde.make_spec(n, steps, 1e-3, alpha=0.85, dims=3, hist_mode=mode, block=16, u_top=u0[0],
                          weight_mode=0, fk=3, fp=0.9)
    u = fpde.solve(spec, u0)
    ref = oracle.solve_3d(oracle.make_params(n, steps, 1e-3, 0.85, ksat=spec.ksat, weight_mode=0, fk=3, fp=0.9),
                          u0)
    tol = TOL_MIXED if mode.startswith("split") else TOL_F64
    err = rel_err(u, ref)
    print(f"3D FK 3 {mode}: rel err {err:.3e}")
    assert err <= tol
    nodes, M = 200, 120
    u1 = fpde.init_1d(nodes, 1302)
    sp1 = fpde.make_spec(nodes, M, 1e-3, picard_max=2, picard_tol=1e-12, hist_mode=mode, block=32, u_top=u1[0],
                         weight_mode=0, fk=3, fp=0.9, do_q=6, do_rule=0, do_a0=0.1, do_a1=1.0)
    r1 = oracle.solve_1d(oracle.make_params(nodes, M, 1e-3, ksat=sp1.ksat, picard_max=2, picard_tol=1e-12,
                                            weight_mode=0, fk=3, fp=0.9, do_q=6, do_rule=0), u1)
    err = rel_err(fpde.solve(sp1, u1), r1)
    print(f"1D FK 3 distributed order {mode}: rel err {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("fk", [0, 3])
@pytest.mark.parametrize("mode", ["mfma_f64", "split_f16"])
def test_folded_near_field_long_blocks_match_oracle(fpde, oracle, mode, fk):
    """History blocks long enough that the near field is folded ahead in
    several 16-level groups: 3D in 64-step blocks (the history combine folds
    three groups per block, each into the block's later steps) and 1D in
    128-step blocks (the cooperative launches fold seven), with the Toeplitz
    weights and with FK 3's per-step weight rows, against the oracle over two
    full blocks and part of a third."""
    kw = dict(weight_mode=0, fk=3, fp=0.9) if fk else {}
    tol = TOL_MIXED if mode.startswith("split") else TOL_F64
    n, steps = 12, 150
    u0 = fpde.init_3d(n, 1777)
    spec = fpde.make_spec(n, steps, 1e-3, alpha=0.8, dims=3, hist_mode=mode, block=64, u_top=u0[0], **kw)
    ref = oracle.solve_3d(oracle.make_params(n, steps, 1e-3, 0.8, ksat=spec.ksat, weight_mode=spec.weight_mode,
                                             fk=fk, fp=0.9), u0)
    err = rel_err(fpde.solve(spec, u0), ref)
    print(f"3D block 64 FK {fk} {mode}: rel err {err:.3e}", flush=True)
    assert err <= tol
    nodes, M = 500, 300
    u1 = fpde.init_1d(nodes, 1778)
    sp1 = fpde.make_spec(nodes, M, 1e-3, alpha=0.7, picard_max=3, picard_tol=1e-12, hist_mode=mode, block=128,
                         u_top=u1[0], **kw)
    r1 = oracle.solve_1d(oracle.make_params(nodes, M, 1e-3, 0.7, ksat=sp1.ksat, picard_max=3, picard_tol=1e-12,
                                            weight_mode=sp1.weight_mode, fk=fk, fp=0.9), u1)
    err = rel_err(fpde.solve(sp1, u1), r1)
    print(f"1D block 128 FK {fk} {mode}: rel err {err:.3e}", flush=True)
    assert err <= tol


@pytest.mark.parametrize("n,mode", [(12, "plain"), (12, "split_f16"), (12, "series"), (129, "mfma_f64")])
def test_solve_3d_split3d_matches_oracle(fpde, oracle, n, mode):
    """split3d odd: each LOD sweep carries a third of the history term (Om1
    fpde_hpc_1d.cpp:4172-4173, :4258-4263), in the partitioned sweeps (12^3)
    and the one-thread-per-line / PCR sweeps (129^3), for a GEMM, a split and
    the series history."""
    steps = 40 if n == 12 else 12
    u0 = fpde.init_3d(n, 1003)
    kw = dict(hist_mode="plain", sum_alg=2, sum_param=16) if mode == "series" else dict(hist_mode=mode)
    spec = fpde.make_spec(n, steps, 1e-3, alpha=0.85, dims=3, block=16, u_top=u0[0], split3d=1, **kw)
    u = fpde.solve(spec, u0)

    def ref(split):
        return oracle.solve_3d(oracle.make_params(n, steps, 1e-3, 0.85, ksat=spec.ksat, split3d=split,
                                                  sum_alg=spec.sum_alg, sum_param=spec.sum_param), u0)
    r1 = ref(1)
    tol = TOL_MIXED if mode.startswith("split") else TOL_F64
    assert rel_err(ref(0), r1) >= 1e3 * tol         # the flag changes the solution
    err = rel_err(u, r1)
    print(f"3D split3d n={n} {mode}: rel err {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("n", [129, 150])
def test_solve_3d_long_lines_matches_oracle(fpde, oracle, n):
    """Lines longer than 128 nodes take the one-thread-per-line lateral sweep
    (Thomas coefficients in HBM) and the PCR depth sweep: FP64 and split
    forms against the oracle, across a history GEMM block, with and without
    the space-fractional sums."""
    steps = 36
    u0 = fpde.init_3d(n, 1003)
    cases = [("plain", TOL_F64, 0.0), ("split_bf16", TOL_MIXED, 0.0)]
    if n == 129:
        cases.append(("mfma_f64", TOL_F64, 0.7))
    for mode, tol, gamma in cases:
        spec = fpde.make_spec(n, steps, 1e-3, alpha=0.85, dims=3, hist_mode=mode, block=16, u_top=u0[0],
                              gamma=gamma)
        u = fpde.solve(spec, u0)
        ref = _oracle_space3d(oracle, spec, u0) if gamma else oracle.solve_3d(
            oracle.make_params(n, steps, 1e-3, 0.85, ksat=spec.ksat), u0)
        err = rel_err(u, ref)
        print(f"3D n={n} {mode} gamma {gamma}: rel err {err:.3e}")
        assert err <= tol


def test_advance_in_pieces_equals_one_shot(fpde):
    nodes, steps = 400, 150
    u0 = fpde.init_1d(nodes, 3)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.6, hist_mode="mfma_f64", u_top=u0[0])
    ref = fpde.solve(spec, u0)
    with fpde.Solver(spec) as s:
        s.set_initial(u0)
        for chunk in (1, 7, 30, 50, 100):
            s.advance(chunk)
            _ = s.solution()
        assert s.step == steps
        u = s.solution()
    assert rel_err(u, ref) <= 1e-13


def test_server_results_same_alone_and_in_mix(fpde):
    rng = np.random.default_rng(1004)
    specs = []
    for i in range(10):
        if i % 4 == 3:
            n = int(rng.integers(8, 14))
            u0 = fpde.init_3d(n, 2000 + i)
            sp = fpde.make_spec(n, int(rng.integers(20, 60)), 1e-3, alpha=float(rng.uniform(0.3, 0.95)),
                                dims=3, hist_mode="split_f16", block=16, u_top=u0[0])
        else:
            n = int(rng.integers(100, 900))
            u0 = fpde.init_1d(n, 2000 + i)
            kw = dict(do_q=4) if i % 5 == 1 else {}
            # the server launches 1D steps per Picard iteration: the same here
            sp = fpde.make_spec(n, int(rng.integers(60, 300)), 1e-3, alpha=float(rng.uniform(0.3, 0.95)),
                                hist_mode="split_f16", u_top=u0[0], step_mode=1, **kw)
        specs.append((sp, u0))
    alone = [fpde.solve(sp, u0) for sp, u0 in specs]
    fps = [fpde.footprint(sp) for sp, _ in specs]
    srv = fpde.Server(mem_budget_bytes=max(sum(fps) // 3, max(fps)), max_active=4)
    ids = [srv.submit(sp, u0) for sp, u0 in specs]
    srv.run()
    for rid, a in zip(ids, alone):
        info = srv.info(rid)
        assert info["status"] == 0
        assert rel_err(srv.result(rid), a) <= 1e-14
    srv.close()


def test_grid_wait_time_out_is_reported(fpde, monkeypatch):
    """A grid exchange whose wait runs out of time ends every workgroup of
    the cooperative launch (each leaves at its first unmet wait) and the
    solve reports it.  The condition is forced from the host: a bound of one
    tick of the device's constant clock (FPDE_TEST_WAIT_TICKS, read when a
    solver is created), so every kernel still exits normally.  A solver made
    without the bound then solves the same spec as usual."""
    u0 = fpde.init_1d(3000, 91)
    spec = fpde.make_spec(3000, 130, 1e-3, alpha=0.7, hist_mode="split_bf16", picard_max=3, picard_tol=1e-12,
                          u_top=u0[0])
    monkeypatch.setenv("FPDE_TEST_WAIT_TICKS", "1")
    with pytest.raises(fpde.FpdeError, match="timed out"):
        fpde.solve(spec, u0)
    monkeypatch.delenv("FPDE_TEST_WAIT_TICKS")
    u = fpde.solve(spec, u0)
    assert np.isfinite(u).all()


def test_server_marks_a_timed_out_request_failed(fpde, monkeypatch):
    """A served 1D request whose grid exchange times out (bound forced from
    the host as above, while the server admits it) ends with status 3
    (failed), gives back its memory and slot, and the server goes on: the
    next request, admitted without the bound, completes and equals its solo
    result."""
    u0 = fpde.init_1d(2500, 92)
    spec = fpde.make_spec(2500, 200, 1e-3, alpha=0.6, hist_mode="split_bf16", picard_max=2, picard_tol=1e-12,
                          u_top=u0[0])
    srv = fpde.Server(mem_budget_bytes=2 * fpde.footprint(spec), max_active=1)
    try:
        monkeypatch.setenv("FPDE_TEST_WAIT_TICKS", "1")
        bad = srv.submit(spec, u0)
        srv.poll()                                   # admitted (solver created) with the bound
        monkeypatch.delenv("FPDE_TEST_WAIT_TICKS")
        srv.run()
        assert srv.info(bad)["status"] == 3
        good = srv.submit(spec, u0)
        srv.run()
        assert srv.info(good)["status"] == 0
        np.testing.assert_array_equal(srv.result(good), fpde.solve(spec, u0))
    finally:
        srv.close()


def test_serving_mix_512_requests(fpde):
    """The serving mix at full size (512 requests, seed 1004: 80% 1D with N
    500..20000 and M 1000..40000, 20% 3D 32^3..96^3, 25% distributed order,
    split bf16): a quarter submitted up front, the rest 8 per scheduling round
    while the server runs.  Every request completes, none is rejected, and the
    smallest 1D, 3D and distributed-order requests equal their solo results."""
    reqs = fpde.serving_mix(512, 16.0, 1004, "split_bf16")
    with fpde.Server(200 << 30, max_active=16) as srv:
        ids, k, rounds = [], 0, 0
        while k < 128:
            ids.append(srv.submit(reqs[k][1], reqs[k][2]))
            k += 1
        while True:
            for _ in range(8):
                if k < len(reqs):
                    ids.append(srv.submit(reqs[k][1], reqs[k][2]))
                    k += 1
            left = srv.poll(256)
            rounds += 1
            if rounds % 25 == 0:
                print(f"serving mix: round {rounds}, {k} submitted, {left} active or queued", flush=True)
            if k >= len(reqs) and left == 0:
                break
        status = [srv.info(i)["status"] for i in ids]
        assert status.count(0) == 512, {s: status.count(s) for s in set(status)}
        kinds = {"1d": lambda sp: sp.dims == 1 and not sp.do_q, "3d": lambda sp: sp.dims == 3,
                 "do": lambda sp: sp.dims == 1 and sp.do_q > 0}
        sample = []
        for name, pred in kinds.items():
            idx = sorted((r[3], j) for j, r in enumerate(reqs) if pred(r[1]))
            sample += [(name, j) for _, j in idx[:2]]
        for name, j in sample:
            _, spec, u0, _ = reqs[j]
            solo = fpde.solve(spec, u0)
            err = rel_err(srv.result(ids[j]), solo)
            print(f"serving mix: request {j} ({name}, N={spec.nodes}, M={spec.steps}) vs solo {err:.1e}", flush=True)
            assert err <= 1e-13


def test_advance_batch_matches_solo(fpde):
    """Solves of uneven N, M, alpha, history form, block size, chunk size and
    time grid advanced as one batch (one cooperative launch per history block for the
    block-mode ones; the plain form on its own stream) give each solo result exactly:
    every solve keeps its own barriers and Picard decisions.  One solver
    starts the batch 37 steps ahead, so the batch's blocks are not aligned;
    another hands a far part already computed ahead over to the batch."""
    cases = [
        dict(n=3000, m=300, alpha=0.6, hist_mode="split_bf16", block=64, picard_max=3, picard_tol=1e-10),
        dict(n=10000, m=200, alpha=0.85, hist_mode="mfma_f64", block=32, picard_max=3, picard_tol=1e-10,
             chunk_rows=256),
        dict(n=777, m=513, alpha=0.35, hist_mode="split_f16", block=128, chunk_rows=96),
        dict(n=2048, m=150, alpha=0.7, hist_mode="split_bf16", block=16, do_q=4),
        dict(n=1200, m=400, alpha=0.75, hist_mode="plain", sum_alg=2, sum_param=10),
        dict(n=1500, m=120, alpha=0.5, hist_mode="plain"),
        dict(n=900, m=200, alpha=0.6, hist_mode="split_f16", block=32, picard_max=2, picard_tol=1e-10,
             taus=0.2 * np.diff((np.arange(201) / 200.0) ** 2)),     # graded prescribed grid
        dict(n=600, m=90, alpha=0.7, hist_mode="split_bf16", gamma=0.8),   # space-fractional: own stream
        dict(n=12, m=40, alpha=0.8, hist_mode="split_f16", block=16, dims=3),  # 3D: own stream
    ]
    specs, u0s = [], []
    for i, c in enumerate(cases):
        u0 = (fpde.init_3d if c.get("dims") == 3 else fpde.init_1d)(c["n"], 3000 + i)
        kw = {k: v for k, v in c.items() if k not in ("n", "m")}
        specs.append(fpde.make_spec(c["n"], c["m"], 1e-3, u_top=u0[0], **kw))
        u0s.append(u0)
    # case 2 starts 37 steps ahead (its blocks are not aligned with the batch's
    # calls); case 0 starts two whole 64-step blocks ahead, so the older far part
    # of its third block is already running on its second stream when the batch
    # takes over
    pre = [37 if i == 2 else (128 if i == 0 else 0) for i in range(len(cases))]

    def solo(sp, u0, p):                           # the batch's schedule, alone
        with fpde.Solver(sp) as s:
            s.set_initial(u0)
            if p:
                s.advance(p)
            s.advance(150)
            s.advance(0)
            return s.solution()

    alone = [solo(sp, u0, p) for sp, u0, p in zip(specs, u0s, pre)]
    solvers = [fpde.Solver(sp) for sp in specs]
    try:
        for s, u0 in zip(solvers, u0s):
            s.set_initial(u0)
        for s, p in zip(solvers, pre):
            if p:
                s.advance(p)
        fpde.advance_batch(solvers, 150)           # part of the way
        assert [s.step for s in solvers] == [min(c["m"], 150 + p) for c, p in zip(cases, pre)]
        fpde.advance_batch(solvers, 0)             # the rest
        got = [s.solution() for s in solvers]
    finally:
        for s in solvers:
            s.close()
    for c, a, b in zip(cases, alone, got):
        d = float(np.abs(b - a).max())
        print(f"batch vs solo {c['hist_mode']} N={c['n']}: max |diff| {d:.3e}", flush=True)
        np.testing.assert_array_equal(b, a)


def test_advance_batch_of_32_small_solves(fpde):
    """The largest batch: 32 solves of 5..700 nodes (one to four chunk
    workgroups each, so the workgroup table has 33 uneven bounds), M from 40 to
    300 (solves leave the batch at different blocks), blocks 16..128, all
    equal to their solo results."""
    rng = np.random.default_rng(3232)
    modes = ("split_bf16", "split_f16", "mfma_f64")
    specs, u0s = [], []
    for i in range(32):
        n = 5 if i < 2 else int(rng.integers(6, 701))
        u0 = fpde.init_1d(n, 4000 + i)
        specs.append(fpde.make_spec(n, int(rng.integers(40, 301)), 1e-3, alpha=float(rng.uniform(0.2, 0.95)),
                                    hist_mode=modes[i % 3], block=(16, 32, 64, 128)[i % 4], picard_max=1 + i % 3,
                                    picard_tol=1e-10, u_top=u0[0]))
        u0s.append(u0)
    alone = [fpde.solve(sp, u0) for sp, u0 in zip(specs, u0s)]
    got = fpde.solve_batch(specs, u0s)
    for a, b in zip(alone, got):
        np.testing.assert_array_equal(b, a)


def test_advance_batch_refuses_what_does_not_fit(fpde):
    """Nine N=65536 solves need 9 x 256 chunk workgroups at once, more than
    the GPU holds for this kernel: refused before any launch.  One of them
    alone (256 workgroups, the longest line of the block form) still runs as
    a batch and equals its solo result."""
    u0 = fpde.init_1d(65536, 5)
    spec = fpde.make_spec(65536, 6, 1e-3, hist_mode="split_bf16", picard_max=2, picard_tol=1e-10, u_top=u0[0])
    solvers = [fpde.Solver(spec) for _ in range(9)]
    try:
        for s in solvers:
            s.set_initial(u0)
        with pytest.raises(fpde.FpdeError, match="exceed"):
            fpde.advance_batch(solvers, 0)
        assert all(s.step == 0 for s in solvers)
        fpde.advance_batch(solvers[:1], 0)
        assert solvers[0].step == 6
        np.testing.assert_array_equal(solvers[0].solution(), fpde.solve(spec, u0))
    finally:
        for s in solvers:
            s.close()


def test_batched_3d_steps_equal_solo(fpde):
    """3D solves of different sizes, orders, history forms and block sizes
    (one with split3d odd, one with the generalized kernel FK 3) advance in
    batched LOD launches (one launch per kernel of the step for all of them,
    fpde_solver_advance_batch), in uneven pieces: each equals its solo run
    bit for bit."""
    cases = [(12, 60, 0.85, "plain", 16, {}), (17, 45, 0.6, "mfma_f64", 16, {}),
             (32, 70, 0.75, "split_bf16", 32, {"split3d": 1}),
             (24, 40, 0.8, "mfma_f64", 16, {"weight_mode": 0, "fk": 3, "fp": 0.9})]
    specs, u0s = [], []
    for n, M, a, mode, blk, kw in cases:
        u0 = fpde.init_3d(n, 900 + n)
        specs.append(fpde.make_spec(n, M, 1e-3, alpha=a, dims=3, hist_mode=mode, block=blk, u_top=u0[0], **kw))
        u0s.append(u0)
    solvers = [fpde.Solver(sp) for sp in specs]
    try:
        for s, u0 in zip(solvers, u0s):
            s.set_initial(u0)
        fpde.advance_batch(solvers, 23)
        fpde.advance_batch(solvers, 19)
        fpde.advance_batch(solvers, 0)
        assert all(s.step == s.spec.steps for s in solvers)
        for s, sp, u0 in zip(solvers, specs, u0s):
            np.testing.assert_array_equal(s.solution(), fpde.solve(sp, u0))
    finally:
        for s in solvers:
            s.close()


def test_near_full_batch_beside_a_3d_rider(fpde):
    """Fourteen N=9000 solves in 256-node chunks (504 chunk workgroups: two
    per CU on 252 of the 256 CUs, the batched kernel's limit) in one batch,
    advanced together with a 64^3 3D solve on its own stream, which is
    enqueued first, so its sweeps hold CUs while the cooperative launches are
    dispatched: every grid exchange completes within its bound and every
    result equals its solo run."""
    u0 = fpde.init_1d(9000, 77)
    specs = [fpde.make_spec(9000, 256, 1e-3, alpha=0.5 + 0.025 * j, hist_mode="split_bf16", picard_max=2,
                            picard_tol=1e-12, u_top=u0[0], chunk_rows=256) for j in range(14)]
    u3 = fpde.init_3d(64, 78)
    sp3 = fpde.make_spec(64, 120, 1e-3, alpha=0.85, dims=3, hist_mode="mfma_f64", block=16, u_top=u3[0])
    solvers = [fpde.Solver(sp) for sp in specs] + [fpde.Solver(sp3)]
    try:
        for s in solvers[:-1]:
            s.set_initial(u0)
        solvers[-1].set_initial(u3)
        fpde.advance_batch(solvers, 0)
        assert all(s.step == s.spec.steps for s in solvers)
        for s, sp in zip(solvers[:-1], specs):
            np.testing.assert_array_equal(s.solution(), fpde.solve(sp, u0))
        np.testing.assert_array_equal(solvers[-1].solution(), fpde.solve(sp3, u3))
    finally:
        for s in solvers:
            s.close()


def test_server_batches_block_form_requests(fpde):
    """1D requests in the cooperative block form (step_mode 0) advance in the
    server's batched launches, beside 3D requests on their own streams; the
    server's quantum (256 steps) cuts the 1D steps at the block bounds of
    blocks 16/32/64, so each 1D result equals the solve alone exactly."""
    rng = np.random.default_rng(77)
    specs = []
    for i in range(10):
        if i % 4 == 3:
            n = int(rng.integers(8, 14))
            u0 = fpde.init_3d(n, 900 + i)
            sp = fpde.make_spec(n, int(rng.integers(20, 60)), 1e-3, alpha=float(rng.uniform(0.3, 0.95)),
                                dims=3, hist_mode="split_f16", block=16, u_top=u0[0])
        else:
            n = int(rng.integers(500, 6000))
            u0 = fpde.init_1d(n, 900 + i)
            kw = dict(do_q=4) if i == 5 else {}
            sp = fpde.make_spec(n, int(rng.integers(100, 400)), 1e-3, alpha=float(rng.uniform(0.3, 0.95)),
                                hist_mode=("split_bf16", "mfma_f64", "split_f16")[i % 3], block=(16, 32, 64)[i % 3],
                                picard_max=2, picard_tol=1e-10, u_top=u0[0], **kw)
        specs.append((sp, u0))
    alone = [fpde.solve(sp, u0) for sp, u0 in specs]
    fps = [fpde.footprint(sp) for sp, _ in specs]
    srv = fpde.Server(mem_budget_bytes=max(sum(fps) // 2, max(fps)), max_active=6)
    ids = [srv.submit(sp, u0) for sp, u0 in specs]
    srv.run()
    for (sp, _), rid, a in zip(specs, ids, alone):
        assert srv.info(rid)["status"] == 0
        if sp.dims == 1:
            np.testing.assert_array_equal(srv.result(rid), a)
        else:
            assert rel_err(srv.result(rid), a) <= 1e-14
    srv.close()


# ------------------------------------------------------------ restricted summing
def _lag_weights(alpha, tau, M):
    j = np.arange(1, M + 1, dtype=np.float64)
    return tau ** (1 - alpha) / (1 - alpha) * ((j + 1) ** (1 - alpha) - j ** (1 - alpha))


def _cut_between(alphas, tau, M, lag):
    """A cut between the weights of lags `lag` and `lag+1` of the first order,
    moved until no order has a weight within 1e-6 (relative) of it, so that
    rounding cannot decide which lags are kept."""
    b = _lag_weights(alphas[0], tau, M)
    cut = float(np.sqrt(b[lag - 1] * b[lag]))
    for _ in range(200):
        near = [np.min(np.abs(_lag_weights(a, tau, M) / cut - 1.0)) for a in alphas]
        if min(near) > 1e-6:
            return cut
        cut *= 1.0 + 3e-4
    raise AssertionError("no cut with a margin")


def test_reference_sum_kernel_restricted_through_facade(fpde, oracle):
    """'Sum' with sum_alg 1 (fpde_hpc_1d.cpp:1259-1283): newest first, in chunks of
    BS levels; the first chunk whose weights are all below sum_param is still
    added, and nothing older.  With the cut between kb[23] and kb[24] and BS 16
    the chunks are t=39..24 (kept), 23..8 (all below: added, then stop)."""
    L = fpde.lib()
    rng = np.random.default_rng(12)
    nodes, levels, OS = 200, 41, 256
    old = np.zeros((levels, OS))
    old[:, :nodes] = np.cumsum(rng.standard_normal((levels, nodes)) * 1e-2, axis=0)
    Time = oracle.time_grid(levels + 1, 1e-3)
    _, kb = oracle.l1_weights(0.7, levels - 1, Time, weight_mode=1)
    cut = 0.5 * (kb[23] + kb[24])
    assert np.all(np.abs(kb[:24]) < cut) and np.all(np.abs(kb[24:]) >= cut)
    tau = np.full(levels + 8, 1e-3)
    ref = oracle.time_summing(old[8:, :nodes], kb[8:], tau[8:])
    src = b"#define sum_alg 1\n#define OS 256\n#define BS 16\n__kernel void Sum(...)"
    dv = Dev(L)
    try:
        bS, bold, btau, bK = dv.zeros(OS), dv.put(old), dv.put(tau), dv.put(np.concatenate([kb, np.zeros(8)]))
        bH, bC = dv.zeros(8), dv.zeros(8)
        prog, k = ctypes.c_void_p(), ctypes.c_void_p()
        assert L.fpde_cl_program_create(dv.ctx, src, ctypes.byref(prog)) == 0
        assert L.fpde_cl_kernel_create(prog, b"Sum", ctypes.byref(k)) == 0
        for idx, b in [(0, bS), (1, bold), (3, btau), (4, bH), (5, bC), (6, bK)]:
            assert L.fpde_cl_kernel_set_buffer_arg(k, idx, b) == 0
        iv = ctypes.c_int(levels)
        assert L.fpde_cl_kernel_set_arg(k, 2, 4, ctypes.byref(iv)) == 0
        for idx, v in [(7, cut), (8, 0.0), (9, 0.0), (10, 0.7)]:
            dvv = ctypes.c_double(v)
            assert L.fpde_cl_kernel_set_arg(k, idx, 8, ctypes.byref(dvv)) == 0
        iv2 = ctypes.c_int(levels - 1)
        assert L.fpde_cl_kernel_set_arg(k, 11, 4, ctypes.byref(iv2)) == 0
        g = (ctypes.c_size_t * 1)(OS)
        l = (ctypes.c_size_t * 1)(16)
        assert L.fpde_cl_launch(dv.ctx, k, 1, g, l) == 0, L.fpde_last_error()
        assert L.fpde_cl_finish(dv.ctx) == 0
        S = dv.get(bS, nodes)
        L.fpde_cl_kernel_destroy(k)
        L.fpde_cl_program_destroy(prog)
    finally:
        dv.close()
    assert np.max(np.abs(S - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("weight_mode", [0, 1])
def test_restricted_weights_match_oracle(fpde, oracle, weight_mode):
    """Toeplitz weights are zero beyond the cut lag; the two oldest intervals
    keep their weight only while every newer interval is kept."""
    M, tau, alpha = 300, 1e-3, 0.6
    cut = _cut_between([alpha], tau, M, 70)
    spec = fpde.make_spec(50, M, tau, alpha=alpha, weight_mode=weight_mode, sum_alg=1, sum_param=cut)
    with fpde.Solver(spec) as s:
        local, omega, corr = s.weights()
    p = oracle.make_params(50, M, tau, alpha, weight_mode=weight_mode, ksat=spec.ksat)
    lo, om, co = oracle.folded_weights(p)
    keep = _lag_weights(alpha, tau, M) >= cut
    assert keep[69] and not keep[70]
    assert np.array_equal(omega[70:], np.zeros(M - 70))
    assert np.max(np.abs(omega[:70] - om[:70]) / np.abs(om[:70])) <= 1e-10
    T = oracle.time_grid(M + 1, tau)
    Da = omega[0] / _lag_weights(alpha, tau, 1)[0]
    for n in range(1, M):
        # newest first: interval t survives when all lags n-t' (t' >= t) pass
        ok = True
        for t in range(n - 1, -1, -1):
            b = oracle.kb_series(alpha, T[t], T[t + 1], T[n + 1])[0] if (weight_mode == 1 and t <= 1) \
                else oracle.kb_closed(alpha, T[t], T[t + 1], T[n + 1])
            ok = ok and abs(b) >= cut
            if t <= 1:
                got = omega[n - t - 1] + corr[n, t]
                want = Da * b if ok else 0.0
                assert abs(got - want) <= 1e-10 * abs(Da * b), (n, t, got, want)


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_bf16", "split_f16"])
def test_restricted_sum_1d_matches_oracle(fpde, oracle, mode):
    nodes, steps, alpha, tau = 500, 400, 0.5, 1e-3
    u0 = fpde.init_1d(nodes, 44)
    for lag in (37, 150):
        cut = _cut_between([alpha], tau, steps, lag)
        spec = fpde.make_spec(nodes, steps, tau, alpha=alpha, hist_mode=mode, block=32, u_top=u0[0],
                              sum_alg=1, sum_param=cut)
        u = fpde.solve(spec, u0)
        ref = _oracle_1d(oracle, spec, u0)
        full = _oracle_1d(oracle, fpde.make_spec(nodes, steps, tau, alpha=alpha, u_top=u0[0]), u0)
        err = rel_err(u, ref)
        print(f"restricted 1D lag {lag} {mode}: rel err {err:.3e}, cut effect {rel_err(ref, full):.3e}")
        assert rel_err(ref, full) > 1e3 * TOL_MIXED     # the cut changes the answer
        assert err <= (TOL_MIXED if mode.startswith("split") else TOL_F64)


def test_restricted_sum_distributed_order_and_3d(fpde, oracle):
    nodes, steps, tau = 300, 200, 1e-3
    u0 = fpde.init_1d(nodes, 45)
    q = 6
    alphas = [0.1 + 0.9 * i / (q - 1) for i in range(q)]
    cut = _cut_between([0.5] + alphas[:-1], tau, steps, 40)
    for mode in ("plain", "mfma_f64"):
        spec = fpde.make_spec(nodes, steps, tau, do_q=q, do_rule=1, do_a0=0.1, do_a1=1.0, hist_mode=mode,
                              block=16, u_top=u0[0], sum_alg=1, sum_param=cut)
        u = fpde.solve(spec, u0)
        assert rel_err(u, _oracle_1d(oracle, spec, u0)) <= TOL_F64
    n = 10
    u3 = fpde.init_3d(n, 46)
    cut3 = _cut_between([0.85], tau, 90, 20)
    for mode in ("plain", "split_f16"):
        spec = fpde.make_spec(n, 90, tau, alpha=0.85, dims=3, hist_mode=mode, block=16, u_top=u3[0],
                              sum_alg=1, sum_param=cut3)
        u = fpde.solve(spec, u3)
        p = oracle.make_params(n, 90, tau, 0.85, ksat=spec.ksat, sum_alg=1, sum_param=cut3)
        err = rel_err(u, oracle.solve_3d(p, u3))
        print(f"restricted 3D {mode}: rel err {err:.3e}")
        assert err <= (TOL_MIXED if mode.startswith("split") else TOL_F64)


# ---------------------------------------------------------- series moments (sum_alg 2)
@pytest.mark.parametrize("P", [1, 10, 64])
@pytest.mark.parametrize("picard", [1, 3])
def test_series_history_1d_matches_oracle(fpde, oracle, P, picard):
    nodes, steps, alpha, tau = 400, 300, 0.7, 1e-3
    u0 = fpde.init_1d(nodes, 50 + P)
    spec = fpde.make_spec(nodes, steps, tau, alpha=alpha, u_top=u0[0], picard_max=picard,
                          picard_tol=1e-9 if picard > 1 else 0.0, sum_alg=2, sum_param=float(P))
    u = fpde.solve(spec, u0)
    ref = _oracle_1d(oracle, spec, u0, picard_tol=spec.picard_tol)
    err = rel_err(u, ref)
    print(f"series P={P} picard={picard}: rel err {err:.3e}")
    assert err <= TOL_F64


def test_series_history_pieces_and_3d(fpde, oracle):
    nodes, steps = 300, 120
    u0 = fpde.init_1d(nodes, 51)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.6, u_top=u0[0], sum_alg=2, sum_param=12.0)
    with fpde.Solver(spec) as s:
        s.set_initial(u0)
        for chunk in (1, 7, 40, 1000):
            s.advance(chunk)
            s.solution()            # finalize between pieces
        u = s.solution()
    assert rel_err(u, _oracle_1d(oracle, spec, u0)) <= TOL_F64
    n = 10
    u3 = fpde.init_3d(n, 52)
    spec3 = fpde.make_spec(n, 60, 1e-3, alpha=0.85, dims=3, u_top=u3[0], sum_alg=2, sum_param=10.0)
    u = fpde.solve(spec3, u3)
    p = oracle.make_params(n, 60, 1e-3, 0.85, ksat=spec3.ksat, sum_alg=2, sum_param=10.0)
    err = rel_err(u, oracle.solve_3d(p, u3))
    print(f"series 3D: rel err {err:.3e}")
    assert err <= TOL_F64


def test_reference_sum_kernel_series_through_facade(fpde, oracle):
    """'Sum' with sum_alg 2 (fr_ocl_call_sum, fpde_hpc_1d.cpp:1837-1868): two
    consecutive steps update the moments Htmp and return the series sum, as
    time_summing's sum_alg 2 branch (:4137-4151) does."""
    L = fpde.lib()
    rng = np.random.default_rng(13)
    nodes, OS, P, alpha, tau = 150, 192, 16, 0.7, 1e-3
    T = oracle.time_grid(60, tau)
    H0 = rng.standard_normal((OS, P)) * 1e-3
    Href = H0[:nodes].copy()
    src = b"#define sum_alg 2\n#define OS 192\n#define BS 16\n__kernel void Sum(...)"
    dv = Dev(L)
    try:
        bS, bH = dv.zeros(OS), dv.put(H0)
        btau, bK = dv.put(np.full(64, tau)), dv.zeros(8)
        prog, k = ctypes.c_void_p(), ctypes.c_void_p()
        assert L.fpde_cl_program_create(dv.ctx, src, ctypes.byref(prog)) == 0
        assert L.fpde_cl_kernel_create(prog, b"Sum", ctypes.byref(k)) == 0
        for n in (40, 41):
            old = rng.standard_normal((2, OS))
            K, v3 = oracle.series_row(alpha, T[n - 1], T[n], T[n + 1], 1e-5, cap=P)
            assert K >= P
            ref = oracle.series_step(Href, v3, K, T[n], T[n + 1], alpha, tau, old[0, :nodes], old[1, :nodes])
            bold, bC = dv.put(old), dv.put(v3)
            for idx, b in [(0, bS), (1, bold), (3, btau), (4, bH), (5, bC), (6, bK)]:
                assert L.fpde_cl_kernel_set_buffer_arg(k, idx, b) == 0
            for idx, v in [(2, 2), (11, n)]:
                iv = ctypes.c_int(v)
                assert L.fpde_cl_kernel_set_arg(k, idx, 4, ctypes.byref(iv)) == 0
            for idx, v in [(7, float(P)), (8, T[n + 1]), (9, T[n]), (10, alpha)]:
                dvv = ctypes.c_double(v)
                assert L.fpde_cl_kernel_set_arg(k, idx, 8, ctypes.byref(dvv)) == 0
            g = (ctypes.c_size_t * 1)(nodes)
            assert L.fpde_cl_launch(dv.ctx, k, 1, g, None) == 0, L.fpde_last_error()
            assert L.fpde_cl_finish(dv.ctx) == 0
            S = dv.get(bS, nodes)
            H = dv.get(bH, OS * P).reshape(OS, P)
            assert np.max(np.abs(S - ref)) <= 1e-12 * np.max(np.abs(ref))
            assert np.allclose(H[:nodes], Href, rtol=1e-13, atol=1e-18)
            assert np.array_equal(H[nodes:], H0[nodes:])
        L.fpde_cl_kernel_destroy(k)
        L.fpde_cl_program_destroy(prog)
    finally:
        dv.close()


def test_results_files_in_reference_format(fpde, oracle, tmp_path):
    """results.txt rows (fpde_hpc_1d.cpp:9498-9526): t z u C T theta (1D) and
    t z x y u C T theta along the centre depth line plus the raw FP64 field
    (3D), theta = VGM wetness at P = z - u."""
    u0 = fpde.init_1d(50, 12)
    spec = fpde.make_spec(50, 20, 1e-3, alpha=0.7, u_top=u0[0])
    txt = tmp_path / "results.txt"
    with fpde.Solver(spec) as s:
        s.set_initial(u0)
        s.advance(0)
        u = s.solution()
        s.write_results(str(txt))
        s.write_results(str(txt))            # appends
    rows = np.loadtxt(txt)
    assert rows.shape == (100, 6)
    z = np.arange(50) / 49.0
    ref = np.column_stack([np.full(50, 20 * 1e-3), z, u, np.zeros(50), np.zeros(50),
                           [oracle.wetness(zz - uu) for zz, uu in zip(z, u)]])
    assert np.allclose(rows[:50], ref, rtol=6e-6, atol=1e-12) and np.array_equal(rows[:50], rows[50:])
    n = 8
    u3 = fpde.init_3d(n, 13)
    spec3 = fpde.make_spec(n, 5, 1e-3, alpha=0.85, dims=3, u_top=u3[0])
    t3, raw = tmp_path / "r3.txt", tmp_path / "results3d.txt"
    with fpde.Solver(spec3) as s:
        s.set_initial(u3)
        s.advance(0)
        u = s.solution()
        s.write_results(str(t3), str(raw))
    rows = np.loadtxt(t3)
    assert rows.shape == (n, 8)
    c = (n - 1) // 2                                       # idx(i, _M/2, K/2), _M = K = n - 1
    line = u.reshape(n, n, n)[c, c, :]                     # y, x, depth
    assert np.allclose(rows[:, 4], line, rtol=6e-6)
    assert np.allclose(rows[:, 2], c / (n - 1)) and np.allclose(rows[:, 3], c / (n - 1))
    # b_U with the halo: (N+2)^3 for N = n - 1, idx = i + (N+2) j + (N+2)^2 k
    b = np.fromfile(raw, dtype=np.float64).reshape(n + 1, n + 1, n + 1)    # k (y), j (x), i (depth)
    assert np.array_equal(b[:n, :n, :n].ravel(), u)
    u03 = u3.reshape(n, n, n)
    assert np.array_equal(b[n, :n, :n], u03[n - 1]) and np.array_equal(b[:n, :n, n], u03[:, :, n - 1])


# ------------------------------------------------- prescribed non-uniform time grid
def _oracle_grid(oracle, spec, u0, taus, picard_tol=0.0):
    p = oracle.make_params(spec.nodes, spec.steps, spec.tau, spec.alpha, ksat=spec.ksat,
                           coef_mode=spec.coef_mode, picard_max=spec.picard_max, picard_tol=picard_tol,
                           weight_mode=spec.weight_mode, series_eps=spec.series_eps, do_q=spec.do_q,
                           do_rule=spec.do_rule, do_a0=spec.do_a0, do_a1=spec.do_a1, taus=taus)
    return oracle.solve_1d(p, u0)


@pytest.mark.parametrize("mode", ["plain", "mfma_f64", "split_bf16", "split_f16", "split_bf16x3"])
@pytest.mark.parametrize("step_mode", [0, 1])
@pytest.mark.parametrize("picard", [1, 3])
def test_nonuniform_grid_matches_oracle(fpde, oracle, mode, step_mode, picard):
    """Non-Toeplitz weights: every step has its own weight row over Time[]
    (plain: a per-step row; GEMM forms: dense far block + near rows; the
    split forms split the dense rows per fragment)."""
    nodes, steps = 300, 150
    rng = np.random.default_rng(60 + picard)
    taus = 1e-3 * (0.4 + 1.2 * rng.random(steps))
    u0 = fpde.init_1d(nodes, 61)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.7, hist_mode=mode, block=32, picard_max=picard,
                          picard_tol=1e-12 if picard > 1 else 0.0, u_top=u0[0], step_mode=step_mode, taus=taus)
    u = fpde.solve(spec, u0)
    err = rel_err(u, _oracle_grid(oracle, spec, u0, taus, picard_tol=spec.picard_tol))
    print(f"grid {mode} step_mode {step_mode} picard {picard}: rel err {err:.3e}")
    assert err <= (TOL_MIXED if mode.startswith("split") else TOL_F64)


def test_graded_grid_and_distributed_order(fpde, oracle):
    nodes, steps = 400, 200
    k = np.arange(steps + 1) / steps
    taus = 0.2 * np.diff(k ** 2)                      # graded towards t = 0
    u0 = fpde.init_1d(nodes, 62)
    for mode, kw in [("mfma_f64", {}), ("plain", dict(do_q=6, do_rule=1, do_a0=0.1, do_a1=1.0)),
                     ("mfma_f64", dict(do_q=8, do_rule=0, do_a0=0.1, do_a1=1.0)), ("split_f16", {}),
                     ("split_bf16", dict(do_q=8, do_rule=0, do_a0=0.1, do_a1=1.0))]:
        spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.6, hist_mode=mode, block=64, u_top=u0[0], taus=taus, **kw)
        u = fpde.solve(spec, u0)
        err = rel_err(u, _oracle_grid(oracle, spec, u0, taus))
        print(f"graded {mode} {kw}: rel err {err:.3e}")
        assert err <= (TOL_MIXED if mode.startswith("split") else TOL_F64)


@pytest.mark.parametrize("mode", ["split_f16", "split_bf16"])
def test_split_dense_weights_across_magnitudes(fpde, oracle, mode):
    """Step lengths spanning three decades (alternating 1e-5 and 1e-2 blocks):
    the dense far rows then vary over orders of magnitude inside one
    fragment, which the fp16 form scales by the fragment maximum."""
    nodes, steps = 320, 260
    taus = np.where((np.arange(steps) // 7) % 2 == 0, 1e-5, 1e-2) * (1.0 + 0.3 * np.sin(np.arange(steps)))
    u0 = fpde.init_1d(nodes, 64)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.75, hist_mode=mode, block=32, picard_max=2,
                          picard_tol=1e-12, u_top=u0[0], taus=taus)
    u = fpde.solve(spec, u0)
    ref = _oracle_grid(oracle, spec, u0, taus, picard_tol=1e-12)
    err = rel_err(u, ref)
    print(f"three-decade grid {mode}: rel err {err:.3e}")
    assert err <= TOL_MIXED


def test_uniform_tau_array_equals_uniform_solve(fpde):
    nodes, steps = 500, 130
    u0 = fpde.init_1d(nodes, 63)
    for mode in ("plain", "mfma_f64", "split_bf16"):
        a = fpde.solve(fpde.make_spec(nodes, steps, 1e-3, alpha=0.8, hist_mode=mode, u_top=u0[0]), u0)
        b = fpde.solve(fpde.make_spec(nodes, steps, 1e-3, alpha=0.8, hist_mode=mode, u_top=u0[0],
                                      taus=np.full(steps, 1e-3)), u0)
        assert rel_err(a, b) <= (TOL_MIXED if mode.startswith("split") else 1e-12)


# ------------------------------------------------- variable time step (step doubling)
def test_var_tau_steps_equal_prescribed_grid(fpde, oracle):
    nodes, steps = 250, 40
    rng = np.random.default_rng(71)
    taus = 1e-3 * (0.5 + rng.random(steps))
    u0 = fpde.init_1d(nodes, 72)
    spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.7, u_top=u0[0], picard_max=3, picard_tol=1e-12,
                          var_tau=1)
    with fpde.Solver(spec) as s:
        s.set_initial(u0)
        for t in taus:
            s.advance_tau(float(t))
        u = s.solution()
    grid = fpde.solve(fpde.make_spec(nodes, steps, 1e-3, alpha=0.7, u_top=u0[0], picard_max=3, picard_tol=1e-12,
                                     taus=taus), u0)
    assert rel_err(u, grid) <= 1e-13
    assert rel_err(u, _oracle_grid(oracle, spec, u0, taus, picard_tol=1e-12)) <= TOL_F64


def test_varT_step_doubling_matches_oracle_model(fpde, oracle):
    """The reference's outer loop (fpde_hpc_1d.cpp:8978-9022): the tau and err
    sequence of the device driver equals the same decisions taken on oracle
    solves (every solver state is its time grid), and s0 ends on that grid."""
    nodes, alpha, tau0, eps, tau_max, K = 200, 0.7, 2e-3, 3e-10, 1.6e-2, 10
    u0 = fpde.init_1d(nodes, 70)
    mk = lambda cap: fpde.make_spec(nodes, cap, tau0, alpha=alpha, u_top=u0[0], fourier=1.0, var_tau=1)
    spec = mk(64)
    ksat = spec.ksat

    def U(grid):
        if not grid:
            return u0.copy()
        p = oracle.make_params(nodes, len(grid), tau0, alpha, ksat=ksat, taus=np.array(grid))
        return oracle.solve_1d(p, u0)

    # oracle model of the driver
    m0, m2, m3 = [], [], []
    tau, model = tau0, []
    for _ in range(K):
        def check(t):
            nonlocal m0, m2, m3
            m3 = list(m0); m0 = m0 + [t]; m2 = m2 + [t / 2, t / 2]
            d = U(m0) - U(m2)
            return float(np.sum(d * d))
        err = check(tau)
        if err > eps:
            first = tau
            while err > eps:
                m0 = list(m3); m2 = list(m3); tau /= 2; err = check(tau)
                if tau < 1e-10:
                    tau = first; m0 = m0 + [tau]; break
        elif tau < tau_max:
            while err <= eps:
                m0 = list(m3); m2 = list(m3); tau *= 2; err = check(tau)
                if tau > tau_max:
                    break
        model.append((tau, err))

    s0, s2, s3 = fpde.Solver(spec), fpde.Solver(mk(128)), fpde.Solver(mk(64))
    try:
        for s in (s0, s2, s3):
            s.set_initial(u0)
        tau, dev = tau0, []
        for _ in range(K):
            tau, err = fpde.varT_step(s0, s2, s3, tau, eps, tau_max)
            dev.append((tau, err))
        u = s0.solution()
    finally:
        for s in (s0, s2, s3):
            s.close()
    assert [t for t, _ in dev] == [t for t, _ in model]
    for (_, a), (_, b) in zip(dev, model):
        assert abs(a - b) <= 1e-6 * b
    assert rel_err(u, U(m0)) <= TOL_F64


# ------------------------------------------------- space-fractional scheme (gamma < 1)
@pytest.mark.parametrize("n", [4, 5, 17, 255, 256, 257, 1100, 4100])
@pytest.mark.parametrize("wm", [0, 1])
def test_space_sum_kernel_matches_oracle(fpde, oracle, n, wm):
    """calc_v on the FP64 matrix cores across tile (256 rows) and
    block-diagonal chunk (1024 columns) boundaries, both weight modes."""
    rng = np.random.default_rng(3 * n + wm)
    F = rng.standard_normal(n)
    h, gam = 1.0 / (n - 1), 0.65
    ref = oracle.space_v(F, h, gam, weight_mode=wm)
    L = fpde.lib()
    dv = Dev(L)
    try:
        bF, bV = dv.put(F), dv.put(np.full(n, 7.0))
        rc = L.fpde_space_sum(dv.ptr(bV), dv.ptr(bF), n, h, gam, wm, 1e-5)
        assert rc == 0, L.fpde_last_error()
        V = dv.get(bV, n)
    finally:
        dv.close()
    assert V[0] == 0.0 and V[-1] == 0.0
    if n > 4:
        assert rel_err(V, ref) <= 1e-12
    else:
        assert V[1] == 0.0 and abs(V[2] - ref[2]) <= 1e-12 * abs(ref[2])


def _oracle_space(oracle, spec, u0):
    p = oracle.make_params(spec.nodes, spec.steps, spec.tau, spec.alpha, ksat=spec.ksat,
                           coef_mode=spec.coef_mode, picard_max=spec.picard_max, picard_tol=spec.picard_tol,
                           weight_mode=spec.weight_mode, series_eps=spec.series_eps, sum_alg=spec.sum_alg,
                           sum_param=spec.sum_param, gamma=spec.gamma)
    return oracle.solve_1d(p, u0)


@pytest.mark.parametrize("mode,tol", [("plain", TOL_F64), ("mfma_f64", TOL_F64), ("split_bf16", TOL_MIXED)])
@pytest.mark.parametrize("gam", [0.5, 0.8])
def test_space_fractional_linear_solve_matches_oracle(fpde, oracle, mode, tol, gam):
    """Linear coefficients, where the space-fractional term changes the
    solution by 10-30%: every history form with the explicit spatial sum."""
    nodes, steps = 1100, 60
    x = np.linspace(0.0, 1.0, nodes)
    u0 = np.sin(3 * x) + 0.3 * np.cos(17 * x)
    spec = fpde.make_spec(nodes, steps, 1e-4, alpha=0.8, coef_mode=1, hist_mode=mode, block=32, gamma=gam)
    u = fpde.solve(spec, u0)
    ref = _oracle_space(oracle, spec, u0)
    err = rel_err(u, ref)
    print(f"space-fractional linear {mode} gamma {gam}: rel err {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("mode,tol", [("plain", TOL_F64), ("mfma_f64", TOL_F64), ("split_bf16", TOL_MIXED)])
@pytest.mark.parametrize("picard", [1, 3])
def test_space_fractional_vgm_solve_matches_oracle(fpde, oracle, mode, tol, picard):
    """VGM coefficients with Picard: the solution and the space-fractional
    effect (the difference from the classical solve) both match."""
    nodes, steps, gam = 700, 80, 0.8
    u0 = fpde.init_1d(nodes, 81)
    kw = dict(alpha=0.8, hist_mode=mode, block=32, picard_max=picard, picard_tol=1e-12 if picard > 1 else 0.0,
              u_top=u0[0])
    spec = fpde.make_spec(nodes, steps, 1e-3, gamma=gam, **kw)
    u = fpde.solve(spec, u0)
    ref = _oracle_space(oracle, spec, u0)
    classical = _oracle_space(oracle, fpde.make_spec(nodes, steps, 1e-3, **kw), u0)
    err, eff = rel_err(u, ref), rel_err(u - classical, ref - classical)
    print(f"space-fractional vgm {mode} picard {picard}: rel err {err:.3e}, effect err {eff:.3e}")
    assert err <= tol
    assert eff <= (1e-6 if tol == TOL_F64 else 1e-4)


@pytest.mark.parametrize("sum_alg,param", [(1, 1e-3), (2, 16)])
def test_space_fractional_with_restricted_and_series_history(fpde, oracle, sum_alg, param):
    nodes, steps = 600, 70
    x = np.linspace(0.0, 1.0, nodes)
    u0 = np.cos(2 * x) - 0.2 * np.sin(11 * x)
    spec = fpde.make_spec(nodes, steps, 1e-4, alpha=0.6, coef_mode=1, hist_mode="plain", gamma=0.7,
                          sum_alg=sum_alg, sum_param=param)
    u = fpde.solve(spec, u0)
    err = rel_err(u, _oracle_space(oracle, spec, u0))
    print(f"space-fractional sum_alg {sum_alg}: rel err {err:.3e}")
    assert err <= TOL_F64


def test_space_fractional_var_tau_steps(fpde, oracle):
    """Run-time step lengths with the space-fractional sum: the var_tau solver
    equals the prescribed-grid oracle solve."""
    nodes, steps = 400, 30
    rng = np.random.default_rng(90)
    taus = 1e-4 * (0.5 + rng.random(steps))
    x = np.linspace(0.0, 1.0, nodes)
    u0 = np.sin(4 * x)
    spec = fpde.make_spec(nodes, steps, 1e-4, alpha=0.7, coef_mode=1, var_tau=1, gamma=0.6)
    with fpde.Solver(spec) as s:
        s.set_initial(u0)
        for t in taus:
            s.advance_tau(float(t))
        u = s.solution()
    p = oracle.make_params(nodes, steps, 1e-4, 0.7, coef_mode=1, taus=taus, gamma=0.6)
    assert rel_err(u, oracle.solve_1d(p, u0)) <= TOL_F64


# ------------------------------------------------- graph-captured step loop
@pytest.mark.parametrize("case", ["plain", "mfma_f64", "split_bf16", "series", "space", "3d"])
def test_graph_step_loop_equals_eager(fpde, case):
    """fpde_solver_advance_graph replays captured chunks of steps (the
    executable graph updated in place between chunks); every replay is
    checked against the eager launches, with eager work (the finalize of
    the solution read) and synchronisation between replays."""
    kw = dict(alpha=0.75, picard_max=3, picard_tol=1e-12, block=32, step_mode=1)
    if case == "3d":
        n, M = 20, 40
        u0 = fpde.init_3d(n, 5)
        spec = fpde.make_spec(n, M, 1e-3, dims=3, alpha=0.85, hist_mode="mfma_f64", block=16, u_top=u0[0])
    else:
        n, M = 900, 100
        u0 = fpde.init_1d(n, 6)
        mode = {"series": "plain", "space": "mfma_f64"}.get(case, case)
        extra = {"series": dict(sum_alg=2, sum_param=12), "space": dict(gamma=0.7)}.get(case, {})
        spec = fpde.make_spec(n, M, 1e-3, hist_mode=mode, u_top=u0[0], **kw, **extra)
    with fpde.Solver(spec) as a, fpde.Solver(spec) as b:
        a.set_initial(u0)
        b.set_initial(u0)
        for chunk in (20, 20, 13, 27, 20):
            a.advance(chunk)
            b.advance_graph(chunk, steps_per_graph=8)
            ua, ub = a.solution(), b.solution()
            assert a.step == b.step
            assert np.array_equal(ua, ub), f"{case}: replay differs at step {a.step}: {rel_err(ub, ua):.3e}"
        assert a.step == M


def test_graph_step_loop_refuses_cooperative_solver(fpde):
    spec = fpde.make_spec(600, 40, 1e-3, hist_mode="split_bf16", block=32, step_mode=0, u_top=-1.0)
    with fpde.Solver(spec) as s:
        s.set_initial(fpde.init_1d(600, 2))
        rc = fpde.lib().fpde_solver_advance_graph(s._h, 10, 8)
        assert rc == 5 and b"cooperative" in fpde.lib().fpde_last_error()


def _oracle_space3d(oracle, spec, u0):
    p = oracle.make_params(spec.nodes, spec.steps, spec.tau, spec.alpha, ksat=spec.ksat, coef_mode=spec.coef_mode,
                           weight_mode=spec.weight_mode, series_eps=spec.series_eps, gamma=spec.gamma)
    return oracle.solve_3d(p, u0)


@pytest.mark.parametrize("n,mode,tol", [(17, "mfma_f64", TOL_F64), (40, "plain", TOL_F64), (64, "mfma_f64", TOL_F64),
                                        (40, "split_bf16", TOL_MIXED)])
@pytest.mark.parametrize("coef", [0, 1])
def test_space_fractional_3d_matches_oracle(fpde, oracle, n, mode, tol, coef):
    """Every LOD sweep with the line sums of its own axis (16-line groups,
    partial groups and row tiles for n not a multiple of 16)."""
    steps = 12
    u0 = fpde.init_3d(n, 7 + n)
    spec = fpde.make_spec(n, steps, 1e-3, dims=3, alpha=0.85, hist_mode=mode, block=16, coef_mode=coef,
                          u_top=u0[0], gamma=0.7, ksat=1.0 if coef else None)
    u = fpde.solve(spec, u0)
    ref = _oracle_space3d(oracle, spec, u0)
    spec0 = fpde.make_spec(n, steps, 1e-3, dims=3, alpha=0.85, hist_mode=mode, block=16, coef_mode=coef,
                           u_top=u0[0], ksat=spec.ksat)
    classical = _oracle_space3d(oracle, spec0, u0)
    err, eff = rel_err(u, ref), rel_err(u - classical, ref - classical)
    print(f"3D space-fractional n={n} {mode} coef {coef}: rel err {err:.3e}, effect err {eff:.3e}")
    assert err <= tol
    assert eff <= (1e-6 if tol == TOL_F64 else 1e-3)


@pytest.mark.parametrize("mode,tol", [("mfma_f64", 1e-12), ("split_bf16", TOL_MIXED)])
@pytest.mark.parametrize("nodes", [300, 2900])
def test_space_fractional_step_modes_agree(fpde, oracle, mode, tol, nodes):
    """With gamma < 1 a step_mode 0 request (cooperative block launches where
    they apply) runs the per-iteration launches, since the spatial sum needs
    U^n complete before each step: both requests give the same solve, and it
    matches the oracle."""
    steps = 150
    u0 = fpde.init_1d(nodes, 91)
    kw = dict(alpha=0.75, hist_mode=mode, block=32, picard_max=3, picard_tol=1e-12, u_top=u0[0], gamma=0.75)
    a = fpde.solve(fpde.make_spec(nodes, steps, 1e-3, step_mode=0, **kw), u0)
    b = fpde.solve(fpde.make_spec(nodes, steps, 1e-3, step_mode=1, **kw), u0)
    spec = fpde.make_spec(nodes, steps, 1e-3, **kw)
    ref = _oracle_space(oracle, spec, u0)
    print(f"space coop {mode} N={nodes}: coop vs launch {rel_err(a, b):.3e}, coop vs oracle {rel_err(a, ref):.3e}")
    assert rel_err(a, b) <= tol
    assert rel_err(a, ref) <= (TOL_F64 if mode == "mfma_f64" else TOL_MIXED)


# ------------------------------------------------- the benchmark configurations at full size
# (the oracle is too slow here: the FP64 GPU forms, pinned against the oracle at
# smaller sizes above, are the reference, plus size-independent properties)
def _full(fpde, name, modes, **kw):
    out = {}
    for mode in modes:
        spec, u0 = fpde.workload(name, hist_mode=mode, **kw)
        out[mode] = fpde.solve(spec, u0)
    return out


def test_small_configuration_full_size_against_oracle(fpde, oracle):
    """The CPU-runnable configuration at full size (1D N=1000, M=2000,
    tau=1e-3, alpha=0.8, Picard <= 3 with tol 1e-12): every history form
    against the FP64 oracle run to the end."""
    ref = None
    for mode, tol in (("plain", TOL_F64), ("mfma_f64", TOL_F64), ("split_bf16", TOL_MIXED),
                      ("split_f16", TOL_MIXED)):
        spec, u0 = fpde.workload("1d_small", hist_mode=mode, picard_max=3, picard_tol=1e-12)
        if ref is None:
            ref = _oracle_1d(oracle, spec, u0, picard_tol=1e-12)
        err = rel_err(fpde.solve(spec, u0), ref)
        print(f"1d_small {mode}: rel err vs oracle {err:.3e}")
        assert err <= tol


def test_headline_size_mixed_and_fp64_mfma_vs_plain(fpde):
    """1D N=10^4, M=2*10^4, Picard <= 3 (tol 1e-12): split bf16 within 1e-8 and
    FP64 MFMA within 1e-10 of the plain FP64 stream."""
    out = _full(fpde, "1d_large", ["plain", "mfma_f64", "split_bf16"], picard_max=3, picard_tol=1e-12)
    e_split, e_f64 = rel_err(out["split_bf16"], out["plain"]), rel_err(out["mfma_f64"], out["plain"])
    print(f"1d_large: split_bf16 {e_split:.3e}, mfma_f64 {e_f64:.3e}")
    assert e_split <= TOL_MIXED
    assert e_f64 <= TOL_F64


def test_headline_size_superposition(fpde):
    """Linear coefficients at N=10^4, M=2*10^4: the solve is linear in the
    initial field (the Dirichlet top is u0[0]), so solve(u + 0.7 v) equals
    solve(u) + 0.7 solve(v) to the precision of each form."""
    for mode, tol in (("mfma_f64", 1e-12), ("split_bf16", TOL_MIXED)):
        spec, u = fpde.workload("1d_large", hist_mode=mode, coef_mode=1, ksat=1.0)
        v = fpde.init_1d(spec.nodes, 2024)
        a, b, ab = fpde.solve(spec, u), fpde.solve(spec, v), fpde.solve(spec, u + 0.7 * v)
        err = rel_err(ab, a + 0.7 * b)
        print(f"superposition at full size, {mode}: {err:.3e}")
        assert err <= tol


def test_3d_and_distributed_order_full_size_vs_plain(fpde):
    out3 = _full(fpde, "3d", ["plain", "split_bf16"], block=16)
    outd = _full(fpde, "do", ["plain", "split_f16"], picard_max=3, picard_tol=1e-12)
    e3, ed = rel_err(out3["split_bf16"], out3["plain"]), rel_err(outd["split_f16"], outd["plain"])
    print(f"3d 128^3 split_bf16 {e3:.3e}; do N=4000 split_f16 {ed:.3e}")
    assert e3 <= TOL_MIXED and ed <= TOL_MIXED


@pytest.mark.parametrize("nodes", [65536, 40001])
def test_longest_lines_both_launch_forms(fpde, oracle, nodes):
    """The largest 1D lines (256 chunk workgroups, a 512-row reduced system,
    256 chunks in every grid exchange) and a ragged one: both launch forms against the oracle,
    across two history blocks."""
    steps = 130
    u0 = fpde.init_1d(nodes, 33)
    outs = []
    for step_mode in (0, 1):
        spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.7, hist_mode="mfma_f64", block=64, picard_max=3,
                              picard_tol=1e-12, u_top=u0[0], step_mode=step_mode)
        outs.append(fpde.solve(spec, u0))
    p = oracle.make_params(nodes, steps, 1e-3, 0.7, ksat=spec.ksat, picard_max=3, picard_tol=1e-12)
    ref = oracle.solve_1d(p, u0)
    e0, e1 = rel_err(outs[0], ref), rel_err(outs[1], ref)
    print(f"N={nodes}: cooperative {e0:.3e}, per-iteration {e1:.3e}")
    assert e0 <= TOL_F64 and e1 <= TOL_F64


def test_autotune_block_picks_a_candidate_that_solves_right(fpde, oracle):
    """The autotuner times every candidate block on the device and returns one
    of them; a solve with the picked block matches the oracle (1D, split and
    FP64 MFMA forms) and the picked block is stable on a 3D spec too."""
    nodes, steps = 900, 300
    u0 = fpde.init_1d(nodes, 77)
    for mode, tol in (("mfma_f64", TOL_F64), ("split_bf16", TOL_MIXED)):
        spec = fpde.make_spec(nodes, steps, 1e-3, alpha=0.7, hist_mode=mode, block=0, picard_max=3,
                              picard_tol=1e-12, u_top=u0[0])
        b, ms = fpde.autotune_block(spec, u0)
        print(f"autotune {mode} N={nodes} M={steps}: block {b}, {ms * 1e3:.1f} us/step")
        assert b in (16, 32, 64, 128) and ms > 0.0
        spec.block = b
        ref = _oracle_1d(oracle, spec, u0, picard_tol=1e-12)
        assert rel_err(fpde.solve(spec, u0), ref) <= tol
    spec3, v0 = fpde.workload("3d", hist_mode="split_f16", nodes=24, steps=280, block=0)
    b3, ms3 = fpde.autotune_block(spec3, v0, trial_steps=280)
    print(f"autotune 3d 24^3: block {b3}, {ms3 * 1e3:.1f} us/step")
    assert b3 in (16, 32, 64, 128) and ms3 > 0.0
