"""Host-side API of the history-sum solver (Python mirror over the C ABI).

`Solver` owns one solve on the device (fpde_solver); `Server` is the
continuous-batching request loop (fpde_server).  `workload()` builds the
configurations of the benchmark suite.  Every call goes to libfpde_hip.so;
there is no CPU path here.
"""
import ctypes
import os
import sys

import numpy as np

from ._lib import ResultInfo, Spec, Stats, check, lib, register_owner

HIST_MODES = {"plain": 0, "mfma_f64": 1, "split_bf16": 2, "split_f16": 3, "split_bf16x3": 4}

_dp = ctypes.POINTER(ctypes.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


def make_spec(nodes, steps, tau, alpha=0.8, dims=1, length=1.0, ksat=None, coef_mode=0,
              picard_max=1, picard_tol=0.0, weight_mode=1, series_eps=1e-5, do_q=0, do_rule=0,
              do_a0=0.1, do_a1=1.0, hist_mode="plain", block=64, step_mode=0, fourier=1.0,
              u_top=-1.0, sum_alg=0, sum_param=1e-5, taus=None, var_tau=0, gamma=0.0, split3d=0, fk=0, fp=0.6,
              chunk_rows=0):
    """fpde_spec; ksat defaults to the value giving mesh Fourier number `fourier`
    (K*tau/(C*h^2)) at suction -u_top.  `taus` (one length per step) makes the
    time grid non-uniform; `gamma` in (0,1) selects the space-fractional
    scheme (0 or 1: classical diffusion); `split3d` odd spreads the 3D history
    term over the three sweeps; `fk`/`fp` select the generalized kernel
    g(t) = t, sqrt(t), t^2 or t^fp (fk 0..3; needs weight_mode 0); `chunk_rows`
    sets the 1D nodes per chunk workgroup (0: 192 for the fastest single solve;
    256 fits more solves into one batched launch)."""
    if ksat is None:
        h = length / (nodes - 1)
        ksat = lib().fpde_ksat_for_fourier(u_top, h, tau, fourier) if coef_mode == 0 else 1.0
    hm = HIST_MODES[hist_mode] if isinstance(hist_mode, str) else int(hist_mode)
    spec = Spec(dims, nodes, steps, tau, alpha, length, ksat, coef_mode, picard_max, picard_tol,
                weight_mode, series_eps, do_q, do_rule, do_a0, do_a1, hm, block, step_mode,
                sum_alg, sum_param, None, var_tau, split3d, gamma, fk, fp, chunk_rows)
    if taus is not None:
        spec._taus = np.ascontiguousarray(taus, dtype=np.float64)   # alive until the solver copies it
        if spec._taus.size != steps:
            raise ValueError("taus must hold one step length per time step")
        spec.taus = spec._taus.ctypes.data
    return spec


def init_1d(nodes, seed, length=1.0):
    u = np.zeros(nodes)
    check(lib().fpde_init_1d(nodes, length, seed, _p(u)))
    return u


def init_3d(n, seed, length=1.0):
    u = np.zeros(n * n * n)
    check(lib().fpde_init_3d(n, length, seed, _p(u)))
    return u


class Solver:
    """One solve kept on the device (fpde_solver_*)."""

    def __init__(self, spec):
        self.spec = spec
        self._h = ctypes.c_void_p()
        check(lib().fpde_solver_create(ctypes.byref(spec), ctypes.byref(self._h)))
        register_owner(self)
        self.unknowns = spec.nodes if spec.dims == 1 else spec.nodes ** 3

    def close(self):
        if self._h:
            lib().fpde_solver_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        # at interpreter shutdown the atexit hook has closed it already; a
        # finaliser must not reach the HIP runtime then
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_initial(self, u0):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        if u0.size != self.unknowns:
            raise ValueError("u0 has the wrong size")
        check(lib().fpde_solver_set_initial(self._h, _p(u0)))

    def reset(self, u0):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        if u0.size != self.unknowns:
            raise ValueError("u0 has the wrong size")
        check(lib().fpde_solver_reset(self._h, _p(u0)))

    def advance(self, nsteps=0):
        check(lib().fpde_solver_advance(self._h, int(nsteps)))

    def advance_graph(self, nsteps=0, steps_per_graph=64):
        """advance() through hipGraphs of steps_per_graph steps each."""
        check(lib().fpde_solver_advance_graph(self._h, int(nsteps), int(steps_per_graph)))

    def sync(self):
        check(lib().fpde_solver_sync(self._h))

    @property
    def step(self):
        return lib().fpde_solver_step(self._h)

    def solution(self):
        u = np.zeros(self.unknowns)
        check(lib().fpde_solver_get_solution(self._h, _p(u)))
        return u

    def advance_tau(self, tau):
        """One step of length tau (var_tau solvers)."""
        check(lib().fpde_solver_advance_tau(self._h, tau))

    def copy_from(self, other):
        """Take other's state (var_tau solvers)."""
        check(lib().fpde_solver_copy(self._h, other._h))

    def write_results(self, path_txt="results.txt", path_3d=None):
        """Append the current solution in the reference driver's results.txt
        (and, in 3D, results3d.txt) formats."""
        check(lib().fpde_solver_write_results(self._h, os.fsencode(path_txt),
                                              None if path_3d is None else os.fsencode(path_3d)))

    def weights(self):
        M = self.spec.steps
        local = ctypes.c_double()
        omega = np.zeros(M)
        corr = np.zeros(2 * M)
        check(lib().fpde_solver_get_weights(self._h, ctypes.byref(local), _p(omega), _p(corr)))
        return local.value, omega, corr.reshape(M, 2)

    def set_profiling(self, mask=3):
        """bit 0: history kernels, bit 1: step kernels (hipEvents per launch)."""
        check(lib().fpde_solver_set_profiling(self._h, int(mask)))

    def stats(self):
        st = Stats()
        check(lib().fpde_solver_get_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in Stats._fields_}


def solve(spec, u0):
    with Solver(spec) as s:
        s.set_initial(u0)
        s.advance(0)
        return s.solution()


def advance_batch(solvers, nsteps=0):
    """Advance up to 32 solvers together (fpde_solver_advance_batch): the 1D
    cooperative-block solvers share one cooperative launch per history block.
    Returns when all of them are done."""
    solvers = list(solvers)
    arr = (ctypes.c_void_p * len(solvers))(*[s._h.value for s in solvers])
    check(lib().fpde_solver_advance_batch(arr, len(solvers), int(nsteps)))


def solve_batch(specs, u0s):
    """Solutions of independent solves run as one batch (advance_batch)."""
    solvers = [Solver(sp) for sp in specs]
    try:
        for s, u0 in zip(solvers, u0s):
            s.set_initial(u0)
        advance_batch(solvers, 0)
        return [s.solution() for s in solvers]
    finally:
        for s in solvers:
            s.close()


class Server:
    """Continuous-batching request server (fpde_server_*)."""

    def __init__(self, mem_budget_bytes, max_active=8):
        self._h = ctypes.c_void_p()
        check(lib().fpde_server_create(int(mem_budget_bytes), int(max_active), ctypes.byref(self._h)))
        register_owner(self)
        self._sizes = {}

    def close(self):
        if self._h:
            lib().fpde_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, spec, u0):
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        rid = ctypes.c_int64()
        check(lib().fpde_server_submit(self._h, ctypes.byref(spec), _p(u0), ctypes.byref(rid)))
        self._sizes[rid.value] = u0.size
        return rid.value

    def cancel(self, rid):
        check(lib().fpde_server_cancel(self._h, rid))

    def poll(self, quantum=256):
        return lib().fpde_server_poll(self._h, quantum)

    def run(self, stop_after=0):
        check(lib().fpde_server_run(self._h, stop_after))

    def clock(self):
        return lib().fpde_server_clock(self._h)

    def info(self, rid):
        inf = ResultInfo()
        check(lib().fpde_server_result(self._h, rid, ctypes.byref(inf), None, 0))
        return {f: getattr(inf, f) for f, _ in ResultInfo._fields_}

    def result(self, rid):
        u = np.zeros(self._sizes[rid])
        check(lib().fpde_server_result(self._h, rid, None, _p(u), u.size))
        return u


def varT_step(s0, s2, s3, tau, eps, tau_max):
    """One outer step of the reference's step doubling; returns (tau, err)."""
    t = ctypes.c_double(tau)
    e = ctypes.c_double(0.0)
    check(lib().fpde_varT_step(s0._h, s2._h, s3._h, ctypes.byref(t), eps, tau_max, ctypes.byref(e)))
    return t.value, e.value


def autotune_block(spec, u0, trial_steps=0):
    """History block size B picked by timed device trials (fpde_autotune_block);
    returns (B, ms per step of its trial)."""
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    b = ctypes.c_int(0)
    ms = ctypes.c_double(0.0)
    check(lib().fpde_autotune_block(ctypes.byref(spec), _p(u0), int(trial_steps), ctypes.byref(b), ctypes.byref(ms)))
    return b.value, ms.value


def footprint(spec):
    return lib().fpde_solver_footprint(ctypes.byref(spec))


# ---------------------------------------------------------------------------
# Workloads of the benchmark suite (1D / distributed order / 3D / serving).
def workload(name, hist_mode="plain", steps=None, nodes=None, **kw):
    """Spec + initial field of a named workload:
    '1d_small'  : N=1000, M=2000, tau=1e-3, alpha=0.8, seed 1000
    '1d_large'  : N=10000, M=20000, tau=5e-4, alpha=0.7, seed 1001
    'do'        : N=4000, M=10000, tau=1e-4... distributed order on [0.1,1],
                  32 Gauss-Legendre nodes, rho(a)=Gamma(3-a)/2, seed 1002
    '3d'        : 128^3, M=1000, alpha=0.85, seed 1003
    """
    if name == "1d_small":
        N, M, tau, a, seed = 1000, 2000, 1e-3, 0.8, 1000
    elif name == "1d_large":
        N, M, tau, a, seed = 10000, 20000, 5e-4, 0.7, 1001
    elif name == "do":
        N, M, tau, a, seed = 4000, 10000, 5e-4, 0.5, 1002
    elif name == "3d":
        N, M, tau, a, seed = 128, 1000, 1e-3, 0.85, 1003
    else:
        raise ValueError(name)
    if nodes is not None:
        N = nodes
    if steps is not None:
        M = steps
    dims = 3 if name == "3d" else 1
    u0 = init_3d(N, seed) if dims == 3 else init_1d(N, seed)
    u_top = u0[0]
    extra = dict(kw)
    if name == "do":
        extra.setdefault("do_q", 32)
        extra.setdefault("do_rule", 0)
        extra.setdefault("do_a0", 0.1)
        extra.setdefault("do_a1", 1.0)
    spec = make_spec(N, M, tau, alpha=a, dims=dims, hist_mode=hist_mode, u_top=u_top, **extra)
    return spec, u0


def serving_mix(n_req=512, rate=16.0, seed=1004, hist_mode="split_bf16"):
    """Requests of the serving mix: Poisson arrivals at `rate` per second; 80%
    1D (N ~ U[500, 20000], M ~ U[1000, 40000], tau 5e-4, block 128), 20% 3D
    (32^3..96^3, M ~ U[200, 1000], tau 1e-3, block 32); alpha ~ U(0.3, 0.95);
    25% distributed order (16 Gauss-Legendre nodes on [0.1, 1]); each with its
    own seeded initial head.  The 1D requests are set for batched launches:
    256-node chunks, so that more of them share a launch, and history blocks
    of 128 steps, which halve their far-part GEMMs (on a batch's critical
    path; a single solve is fastest at 64).  Returns (arrival_s, spec, u0,
    node_steps) tuples in arrival order.  The reference runs such a mix as one process per solve,
    each configured on its command line (mpi_test.sh:12-23)."""
    rng = np.random.default_rng(seed)
    reqs = []
    t = 0.0
    for i in range(n_req):
        t += rng.exponential(1.0 / rate)
        is3d = rng.random() < 0.2
        do = rng.random() < 0.25
        alpha = float(rng.uniform(0.3, 0.95))
        if is3d:
            n = int(rng.integers(32, 97))
            M = int(rng.integers(200, 1001))
            u0 = init_3d(n, 100000 + i)
            kw = dict(dims=3, block=32)
        else:
            n = int(rng.integers(500, 20001))
            M = int(rng.integers(1000, 40001))
            u0 = init_1d(n, 100000 + i)
            kw = dict(dims=1, block=128, chunk_rows=256)
        if do:
            kw.update(do_q=16, do_rule=0, do_a0=0.1, do_a1=1.0)
        tau = 1e-3 if is3d else 5e-4
        spec = make_spec(n, M, tau, alpha=alpha, hist_mode=hist_mode, u_top=u0[0], **kw)
        reqs.append((t, spec, u0, (n ** 3 if is3d else n) * M))
    return reqs
