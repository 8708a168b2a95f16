"""ctypes binding of the FP64 CPU oracle (oracle/fpde_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product (fpde_hpc_amd)
never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


class DoEqParams(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("na", ctypes.c_int), ("tblock", ctypes.c_int), ("nt", ctypes.c_int),
                ("example", ctypes.c_int), ("power", ctypes.c_double), ("D", ctypes.c_double),
                ("max_x", ctypes.c_double), ("max_t", ctypes.c_double), ("gpu_pi", ctypes.c_int)]


class Params(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int), ("steps", ctypes.c_int), ("tau", ctypes.c_double),
        ("alpha", ctypes.c_double), ("length", ctypes.c_double), ("ksat", ctypes.c_double),
        ("coef_mode", ctypes.c_int), ("picard_max", ctypes.c_int), ("picard_tol", ctypes.c_double),
        ("weight_mode", ctypes.c_int), ("series_eps", ctypes.c_double), ("do_q", ctypes.c_int),
        ("do_rule", ctypes.c_int), ("do_a0", ctypes.c_double), ("do_a1", ctypes.c_double),
        ("nthreads", ctypes.c_int), ("sum_alg", ctypes.c_int), ("sum_param", ctypes.c_double),
        ("taus", ctypes.c_void_p), ("gamma", ctypes.c_double), ("split3d", ctypes.c_int),
        ("fk", ctypes.c_int), ("fp", ctypes.c_double),
    ]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_mt64_seed.argtypes = [ctypes.c_void_p, ctypes.c_uint64]
        L.orc_mt64_next.argtypes = [ctypes.c_void_p]
        L.orc_mt64_next.restype = ctypes.c_uint64
        L.orc_mt64_uniform.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_mt64_uniform.restype = ctypes.c_double
        L.orc_kb_closed.argtypes = [ctypes.c_double] * 4
        L.orc_kb_closed.restype = ctypes.c_double
        L.orc_kb_series.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_kb_series.restype = ctypes.c_double
        L.orc_kb.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double]
        L.orc_kb.restype = ctypes.c_double
        L.orc_l1_weights.argtypes = [ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double, _dp]
        L.orc_l1_weights.restype = ctypes.c_double
        L.orc_time_summing.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_double]
        L.orc_series_row.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, _dp]
        L.orc_series_row.restype = ctypes.c_int
        L.orc_series_step.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int] + \
            [ctypes.c_double] * 4 + [_dp, _dp, _dp]
        for f in ("orc_wetness", "orc_inv_dw_dh", "orc_kr"):
            getattr(L, f).argtypes = [ctypes.c_double]
            getattr(L, f).restype = ctypes.c_double
        L.orc_gauss_legendre.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        L.orc_solve_1d.argtypes = [ctypes.POINTER(Params), _dp]
        L.orc_solve_3d.argtypes = [ctypes.POINTER(Params), _dp]
        L.orc_solve_3d_trace.argtypes = [ctypes.POINTER(Params), _dp, _dp]
        L.orc_folded_weights.argtypes = [ctypes.POINTER(Params), _dp, _dp, _dp]
        L.orc_thomas.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_init_1d.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]
        L.orc_init_3d.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]
        L.orc_space_v.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                  _dp, _dp]
        L.orc_space_flux.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp, _dp]
        L.orc_g.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
        L.orc_g.restype = ctypes.c_double
        L.orc_inv_g_der.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_inv_g_der.restype = ctypes.c_double
        L.orc_kb_row_taylor.argtypes = [ctypes.c_double, ctypes.c_int] + [ctypes.c_double] * 5 + \
            [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_kb_row_taylor.restype = ctypes.c_double
        L.orc_kb_general.argtypes = [ctypes.c_double] * 5
        L.orc_kb_general.restype = ctypes.c_double
        L.orc_l1_weights_g.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp,
                                       ctypes.c_int, ctypes.c_double, _dp]
        L.orc_l1_weights_g.restype = ctypes.c_double
        L.orc_ksat_for_fourier.argtypes = [ctypes.c_double] * 4
        L.orc_ksat_for_fourier.restype = ctypes.c_double
        L.orc_doeq_solve.argtypes = [ctypes.POINTER(DoEqParams), _dp]
        L.orc_doeq_solve.restype = ctypes.c_double
        L.orc_do_cder.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                                  ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def mt64_stream(seed, count):
    st = ctypes.create_string_buffer(312 * 8 + 16)
    lib().orc_mt64_seed(st, seed)
    return [lib().orc_mt64_next(st) for _ in range(count)]


def mt64_uniforms(seed, bounds):
    st = ctypes.create_string_buffer(312 * 8 + 16)
    lib().orc_mt64_seed(st, seed)
    return [lib().orc_mt64_uniform(st, a, b) for a, b in bounds]


def kb_closed(alpha, p0, p1, pa):
    return lib().orc_kb_closed(alpha, p0, p1, pa)


def kb_series(alpha, p0, p1, pa, eps=1e-5, max_niter=100000):
    n = ctypes.c_int(0)
    v = lib().orc_kb_series(alpha, p0, p1, pa, eps, max_niter, ctypes.byref(n))
    return v, n.value


def time_grid(steps, tau):
    """Cumulative grid, as init_time_step builds it (fpde_hpc_1d.cpp:3944)."""
    t = np.zeros(steps + 2)
    for k in range(1, steps + 2):
        t[k] = t[k - 1] + tau
    return t


def l1_weights(alpha, n, Time, weight_mode=1, eps=1e-5):
    Time = np.ascontiguousarray(Time, dtype=np.float64)
    kb3 = np.zeros(max(n, 1))
    kkk0 = lib().orc_l1_weights(alpha, n, _p(Time), weight_mode, eps, _p(kb3))
    return kkk0, kb3[:n]


def kb_row_taylor(alpha, fk, fp, t0, t1, gtj, eps=1e-5, max_niter=100000):
    """The reference's _kb_row Taylor series; returns (value, iterations)."""
    n = ctypes.c_int(0)
    v = lib().orc_kb_row_taylor(alpha, fk, fp, t0, t1, gtj, eps, max_niter, ctypes.byref(n))
    return v, n.value


def kb_general(alpha, gp, p0, p1, pa):
    return lib().orc_kb_general(alpha, gp, p0, p1, pa)


def l1_weights_g(alpha, fk, fp, n, Time):
    """(kkk0, kb_3_cache) of step n with the generalized kernel fk (exact weights)."""
    Time = np.ascontiguousarray(Time, dtype=np.float64)
    kb3 = np.zeros(max(n, 1))
    kkk0 = lib().orc_l1_weights_g(alpha, fk, fp, n, _p(Time), 0, 1e-5, _p(kb3))
    return kkk0, kb3[:n]


def time_summing(old, kb3, tau, nthreads=0, sum_alg=0, sum_param=1e-5):
    old = np.ascontiguousarray(old, dtype=np.float64)
    levels, nodes = old.shape
    kb3 = np.ascontiguousarray(kb3, dtype=np.float64)
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    S = np.zeros(nodes)
    lib().orc_time_summing(nodes, levels, _p(old), _p(kb3), _p(tau), _p(S), nthreads, sum_alg, sum_param)
    return S


def series_row(alpha, p0, p1, pa, eps=1e-5, cap=64):
    """(K, v3[:min(K, cap)]) of the sum_alg 2 coefficient row."""
    v3 = np.zeros(max(cap, 1))
    K = lib().orc_series_row(alpha, p0, p1, pa, eps, cap, _p(v3))
    return K, v3[:min(K, cap)]


def series_step(H, v3, K, p1, gtj, alpha, tau, u0, u1):
    """One sum_alg 2 step: updates H [nodes, P] in place, returns S."""
    nodes, P = H.shape
    assert H.flags.c_contiguous and H.dtype == np.float64
    v3 = np.ascontiguousarray(np.concatenate([v3, np.zeros(P)])[:max(P, len(v3))], dtype=np.float64)
    u0 = np.ascontiguousarray(u0, dtype=np.float64)
    u1 = np.ascontiguousarray(u1, dtype=np.float64)
    S = np.zeros(nodes)
    lib().orc_series_step(nodes, P, _p(H), _p(v3), K, p1, gtj, alpha, tau, _p(u0), _p(u1), _p(S))
    return S


def wetness(P):
    return lib().orc_wetness(P)


def inv_dw_dh(P):
    return lib().orc_inv_dw_dh(P)


def kr(P):
    return lib().orc_kr(P)


def gauss_legendre(n, a, b):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre(n, a, b, _p(x), _p(w))
    return x, w


def thomas(a, b, c, d):
    n = len(b)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, d)]
    x = np.zeros(n)
    lib().orc_thomas(n, *[_p(v) for v in arrs], _p(x))
    return x


def make_params(nodes, steps, tau, alpha=0.8, length=1.0, ksat=None, coef_mode=0,
                picard_max=1, picard_tol=1e-12, weight_mode=1, series_eps=1e-5,
                do_q=0, do_rule=0, do_a0=0.1, do_a1=1.0, nthreads=0, fourier=1.0,
                u_top=None, sum_alg=0, sum_param=1e-5, taus=None, gamma=0.0, split3d=0, fk=0, fp=0.6):
    if ksat is None:
        if coef_mode == 0:
            h = length / (nodes - 1)
            ksat = lib().orc_ksat_for_fourier(u_top if u_top is not None else -1.0, h, tau, fourier)
        else:
            ksat = 1.0
    p = Params(nodes, steps, tau, alpha, length, ksat, coef_mode, picard_max, picard_tol,
               weight_mode, series_eps, do_q, do_rule, do_a0, do_a1, nthreads, sum_alg, sum_param, None, gamma,
               split3d, fk, fp)
    if taus is not None:
        p._taus = np.ascontiguousarray(taus, dtype=np.float64)    # kept alive with the struct
        assert p._taus.size == steps
        p.taus = p._taus.ctypes.data
    return p


def init_1d(nodes, seed, length=1.0):
    u = np.zeros(nodes)
    lib().orc_init_1d(nodes, length, seed, _p(u))
    return u


def init_3d(n, seed, length=1.0):
    u = np.zeros(n * n * n)
    lib().orc_init_3d(n, length, seed, _p(u))
    return u


def solve_1d(params, u0):
    u = np.array(u0, dtype=np.float64, copy=True)
    rc = lib().orc_solve_1d(ctypes.byref(params), _p(u))
    if rc != 0:
        raise RuntimeError("orc_solve_1d failed")
    return u


def solve_3d(params, u0):
    u = np.array(u0, dtype=np.float64, copy=True)
    rc = lib().orc_solve_3d(ctypes.byref(params), _p(u))
    if rc != 0:
        raise RuntimeError("orc_solve_3d failed")
    return u


def solve_3d_trace(params, u0):
    """(u^M, fields after the y, x and z sweeps of the last step)."""
    u = np.array(u0, dtype=np.float64, copy=True)
    sw = np.zeros(3 * u.size)
    rc = lib().orc_solve_3d_trace(ctypes.byref(params), _p(u), _p(sw))
    if rc != 0:
        raise RuntimeError("orc_solve_3d failed")
    return u, sw.reshape(3, u.size)


def folded_weights(params):
    M = params.steps
    local = ctypes.c_double(0.0)
    omega = np.zeros(M)
    corr = np.zeros(2 * M)
    lib().orc_folded_weights(ctypes.byref(params), ctypes.byref(local), _p(omega), _p(corr))
    return local.value, omega, corr.reshape(M, 2)


def space_v(F, h, gamma, weight_mode=1, eps=1e-5):
    """calc_v of the explicit space-fractional scheme on a uniform grid."""
    F = np.ascontiguousarray(F, dtype=np.float64)
    V = np.zeros_like(F)
    lib().orc_space_v(F.size, h, gamma, weight_mode, eps, _p(F), _p(V))
    return V


def space_flux(K, u, h):
    K = np.ascontiguousarray(K, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    F = np.zeros_like(u)
    lib().orc_space_flux(u.size, h, _p(K), _p(u), _p(F))
    return F


def ksat_for_fourier(u_top, h, tau, fo=1.0):
    return lib().orc_ksat_for_fourier(u_top, h, tau, fo)


def doeq_params(nx=128, na=64, tblock=32, nt=1280, example=0, power=1.5, D=2.0, max_x=None, max_t=None, gpu_pi=0):
    """fr_do.cpp's equation settings (set_for_ex: [0, 0.5]^2 for example 0, [0, 1]^2 for 1)."""
    if max_x is None:
        max_x = 0.5 if example == 0 else 1.0
    if max_t is None:
        max_t = 0.5 if example == 0 else 1.0
    nt = (nt // tblock) * tblock
    return DoEqParams(nx, na, tblock, nt, example, power, D, max_x, max_t, gpu_pi)


def doeq_solve(p):
    """(C [nx][nt] node-major, RMS error of the last level against the exact
    solution of example 0)."""
    C = np.zeros(p.nx * p.nt)
    err = lib().orc_doeq_solve(ctypes.byref(p), _p(C))
    return C.reshape(p.nx, p.nt), err


def do_cder(F, tau, na, example=-1, power=1.5, maxt=-1, t0=1, j0=0):
    F = np.ascontiguousarray(F, dtype=np.float64)
    DF = np.zeros(F.size)
    lib().orc_do_cder(F.size, tau, na, example, power, _p(F), maxt, t0, j0, _p(DF))
    return DF
