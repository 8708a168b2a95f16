"""Pins the FP64 CPU oracle (oracle/fpde_oracle.c) against what the
mathematics and independent libraries fix: known-answer vectors, closed forms,
invariants, brute force and library routines.  Each check is chosen so that
a dropped term, a wrong sign or an index shift in the oracle fails it."""
import math

import numpy as np
import pytest
from scipy import integrate
from scipy.special import gamma as G


# ---------------------------------------------------------------- MT19937-64
def test_mt19937_64_known_answer(oracle):
    # C++ standard [rand.predef]: the 10000th invocation of a default-
    # constructed std::mt19937_64 (seed 5489) produces 9981545732273789042.
    assert oracle.mt64_stream(5489, 10000)[-1] == 9981545732273789042


def test_uniform_draws_are_canonical_form(oracle):
    raw = oracle.mt64_stream(1001, 6)
    u = oracle.mt64_uniforms(1001, [(0.0, 0.05), (0.0, 2 * math.pi)] * 3)
    for x, (a, b), v in zip(raw, [(0.0, 0.05), (0.0, 2 * math.pi)] * 3, u):
        assert v == (float(x) / 2.0 ** 64) * (b - a) + a


# ---------------------------------------------------------------- L1 weights
@pytest.mark.parametrize("alpha", [0.1, 0.5, 0.7, 0.95])
def test_closed_weight_is_the_integral(oracle, alpha):
    for t0, t1, a in [(0.0, 0.1, 0.3), (0.2, 0.25, 0.25), (1.0, 1.5, 4.0)]:
        ref, _ = integrate.quad(lambda s: (a - s) ** (-alpha), t0, t1, limit=200)
        assert abs(oracle.kb_closed(alpha, t0, t1, a) - ref) <= 1e-9 * abs(ref)


def test_closed_weights_telescope(oracle):
    # sum_{t<=n} b(t,t+1;n+1) = T_{n+1}^(1-a)/(1-a)
    tau = 1e-3
    for alpha in (0.3, 0.8):
        for n in (1, 10, 500):
            T = oracle.time_grid(n + 1, tau)
            kkk0, kb3 = oracle.l1_weights(alpha, n, T, weight_mode=0)
            lhs = kb3.sum() + kkk0
            rhs = T[n + 1] ** (1 - alpha) / (1 - alpha)
            assert abs(lhs - rhs) <= 1e-11 * rhs


def test_alpha_one_limit(oracle):
    T = oracle.time_grid(5, 1e-3)
    kkk0, kb3 = oracle.l1_weights(1.0, 4, T)
    assert kkk0 == 1.0 and np.all(kb3 == 0.0)


@pytest.mark.parametrize("alpha", [0.3, 0.7, 0.9])
def test_series_weight_converges_to_closed_form(oracle, alpha):
    tau = 1e-3
    for n in (1, 2, 7, 300):
        for t0 in (0, 1):
            if t0 >= n:
                continue
            T = np.arange(n + 2) * tau
            v, _ = oracle.kb_series(alpha, T[t0], T[t0 + 1], T[n + 1], eps=1e-16)
            c = oracle.kb_closed(alpha, T[t0], T[t0 + 1], T[n + 1])
            assert abs(v - c) <= 1e-12 * c


def _series_terms(alpha, p0, p1, pa, count):
    # term_i = (alpha)_i / i! * pa^-alpha * (p1^(i+1) - p0^(i+1)) / ((i+1) pa^i)
    out = []
    for i in range(count):
        poch = math.gamma(alpha + i) / math.gamma(alpha) / math.factorial(i)
        out.append(poch * pa ** (-alpha) * (p1 ** (i + 1) - p0 ** (i + 1)) / ((i + 1) * pa ** i))
    return out


@pytest.mark.parametrize("alpha", [0.5, 0.8])
def test_series_stopping_rule_matches_direct_terms(oracle, alpha):
    """_kb_row2 adds terms until the first one with |term| <= eps (absolute)."""
    tau, eps = 1e-3, 1e-5
    for n, t0 in [(1, 0), (2, 0), (2, 1), (9, 1), (50, 0)]:
        p0, p1, pa = t0 * tau, (t0 + 1) * tau, (n + 1) * tau
        terms = _series_terms(alpha, p0, p1, pa, 80)
        k = next(i for i, t in enumerate(terms) if abs(t) <= eps)
        expect = math.fsum(terms[: k + 1])
        v, _ = oracle.kb_series(alpha, p0, p1, pa, eps=eps)
        assert abs(v - expect) <= 1e-12 * abs(expect)


def test_reference_dispatch_uses_series_for_two_oldest(oracle):
    T = oracle.time_grid(12, 1e-3)
    n = 10
    for t0 in range(n):
        w = oracle.lib().orc_kb(0.7, t0, n + 1, T.ctypes.data_as(oracle._dp), 1, 1e-5)
        if t0 <= 1:
            s, _ = oracle.kb_series(0.7, T[t0], T[t0 + 1], T[n + 1])
            assert w == s
        else:
            assert w == oracle.kb_closed(0.7, T[t0], T[t0 + 1], T[n + 1])


# ------------------------------------------- generalized kernel (FK 1/2/3)
_GP = {1: 0.5, 2: 2.0, 3: 0.9}          # g(t) = t^gp (fpde_hpc_1d.cpp:1889-1898), FP 0.9 for FK 3


@pytest.mark.parametrize("fk", [1, 2, 3])
@pytest.mark.parametrize("alpha", [0.3, 0.85, 0.93, 0.98])
def test_generalized_weight_is_the_integral(oracle, fk, alpha):
    """b(t0,t1;a) = int (g(T_a) - g(s))^-alpha ds of every kind of interval
    (oldest, middle, next to the newest, the newest itself) against quad."""
    gp = _GP[fk]
    tau = 1e-3
    for n in (1, 2, 9, 200):
        T = oracle.time_grid(n + 1, tau)
        Ga = T[n + 1] ** gp
        for t in sorted({0, 1, n // 2, n - 1, n}):
            if t > n:
                continue
            p0, p1, pa = T[t], T[t + 1], T[n + 1]
            if t < n:
                ref, _ = integrate.quad(lambda s: (Ga - s ** gp) ** -alpha, p0, p1, epsabs=0, epsrel=1e-13, limit=200)
            else:   # (pa - s)^-alpha times a smooth factor: quad's algebraic weight;
                # g(pa) - g(s) = -G expm1(p log1p(-(pa - s)/pa)) has no cancellation near s = pa
                def phi(s):
                    d = pa - s
                    if d <= 0.0:
                        return (gp * pa ** (gp - 1)) ** -alpha
                    return (-Ga * math.expm1(gp * math.log1p(-d / pa)) / d) ** -alpha
                ref, _ = integrate.quad(phi, p0, pa, weight="alg", wvar=(0, -alpha), epsabs=0, epsrel=1e-13,
                                        limit=200)
            v = oracle.kb_general(alpha, gp, p0, p1, pa)
            assert abs(v - ref) <= 2e-12 * ref, (n, t)


@pytest.mark.parametrize("gp", [0.5, 2.0, 0.9])
def test_generalized_first_step_weight_is_a_beta_function(oracle, gp):
    """The first step's only interval, [0, pa]: s = pa x^(1/gp) turns the
    integral into pa^(1 - alpha gp) B(1/gp, 1 - alpha) / gp, across the order
    range a distributed-order rule samples (end points included)."""
    for alpha in (0.02, 0.1, 0.3, 0.5, 0.7, 0.9, 0.98):
        for pa in (1e-3, 0.37, 2.0):
            v = oracle.kb_general(alpha, gp, 0.0, pa, pa)
            ref = pa ** (1 - alpha * gp) * math.exp(math.lgamma(1 + 1 / gp) + math.lgamma(1 - alpha)
                                                    - math.lgamma(1 + 1 / gp - alpha))
            assert abs(v - ref) <= 1e-12 * ref, (alpha, pa)


@pytest.mark.parametrize("fk", [1, 2, 3])
@pytest.mark.parametrize("alpha", [0.25, 0.7])
def test_generalized_l1_exact_for_linear_in_time(oracle, fk, alpha):
    """For u = c t every interval is linear, so the history plus the local
    term is the generalized Caputo derivative itself:
    c/Gamma(1-a) int_0^T (T^p - s^p)^-a ds = c T^(1-a p) Gamma(1+1/p)/Gamma(1+1/p-a),
    on a uniform and on a random grid."""
    gp = _GP[fk]
    rng = np.random.default_rng(31)
    n, c = 24, 1.7
    for taus in (np.full(n + 1, 2e-2), 2e-2 * (0.3 + rng.random(n + 1))):
        T = np.concatenate([[0.0], np.cumsum(taus)])
        kkk0, kb3 = oracle.l1_weights_g(alpha, fk, gp if fk == 3 else 0.6, n, T)
        old = (c * T[: n + 1])[:, None] * np.ones((1, 2))
        hist = oracle.time_summing(old, kb3, taus)[0]
        approx = (hist + kkk0 * c) / G(1 - alpha)
        exact = c * T[n + 1] ** (1 - alpha * gp) * G(1 + 1 / gp) / G(1 + 1 / gp - alpha)
        assert abs(approx - exact) <= 1e-11 * exact


@pytest.mark.parametrize("fk", [1, 2, 3])
def test_generalized_l1_of_t_squared_tends_to_reference_analytic_form(oracle, fk):
    """The reference's analytic tests use D_g^a t^2 = Gamma(2/p+1)/Gamma(2/p+1-a)
    t^(2-a p) for g = t^p (g2fp/g2fpa, fpde_hpc_1d.cpp:2716-2720, :2843): the L1
    scheme converges to it as tau -> 0."""
    gp, alpha = _GP[fk], 0.6
    errs = []
    for n in (40, 80, 160):
        tau = 1.0 / (n + 1)
        T = oracle.time_grid(n + 1, tau)
        kkk0, kb3 = oracle.l1_weights_g(alpha, fk, 0.9, n, T)
        u = T ** 2
        hist = oracle.time_summing(u[: n + 1, None], kb3, np.full(n + 1, tau))[0]
        approx = (hist + kkk0 * (u[n + 1] - u[n]) / tau) / G(1 - alpha)
        exact = G(2 / gp + 1) / G(2 / gp + 1 - alpha) * T[n + 1] ** (2 - alpha * gp)
        errs.append(abs(approx - exact) / exact)
    assert errs[0] < 2e-2 and errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1]


@pytest.mark.parametrize("fk", [1, 2, 3])
def test_reference_taylor_series_converges_to_the_integral(oracle, fk):
    """_kb_row (fpde_hpc_1d.cpp:1933-1996) runs the Taylor series of f = g^-1
    through inv_g_der (:1901-1930); with eps -> 0 it reaches the integral on
    intervals away from 0.  For g = sqrt(t), f = t^2 and the series has two
    terms: it is exact at the reference's eps = 1e-5 already."""
    gp, alpha, tau = _GP[fk], 0.7, 1e-3
    for n, t in [(5, 2), (40, 20), (40, 38), (300, 150)]:
        T = oracle.time_grid(n + 1, tau)
        gtj = T[n + 1] ** gp
        exact = oracle.kb_general(alpha, gp, T[t], T[t + 1], T[n + 1])
        v, _ = oracle.kb_row_taylor(alpha, fk, 0.9, T[t], T[t + 1], gtj, eps=1e-18, max_niter=200000)
        assert abs(v - exact) <= 1e-9 * exact, (n, t)
        if fk == 1:
            v5, it = oracle.kb_row_taylor(alpha, fk, 0.9, T[t], T[t + 1], gtj, eps=1e-5)
            assert abs(v5 - exact) <= 1e-12 * exact and it <= 2       # f''' = 0


def test_generalized_linear_solve_converges_to_mittag_leffler(oracle):
    """K = 1/C = 1 with g = t^0.9: the generalized (psi-) Caputo derivative
    has the eigenfunction E_a(lambda g(t)^a), so u0 = u_top + A sin(pi z/2)
    gives u = u_top + A E_a(-(pi/2)^2 g(T)^a) sin(pi z/2)."""
    alpha, A, ut, Tend, gp = 0.6, 0.1, -1.0, 0.2, 0.9
    errs = []
    for nodes, steps in ((81, 100), (161, 200)):
        z = np.linspace(0, 1, nodes)
        u0 = ut + A * np.sin(np.pi * z / 2)
        p = oracle.make_params(nodes, steps, Tend / steps, alpha, coef_mode=1, weight_mode=0, fk=3, fp=gp)
        u = oracle.solve_1d(p, u0)
        exact = ut + A * _mittag_leffler(alpha, -(np.pi / 2) ** 2 * (Tend ** gp) ** alpha) * np.sin(np.pi * z / 2)
        errs.append(np.max(np.abs(u - exact)))
    assert errs[1] < 0.6 * errs[0] and errs[1] < 2e-3


def test_generalized_kernel_needs_exact_weights(oracle):
    p = oracle.make_params(10, 3, 1e-3, 0.5, coef_mode=1, weight_mode=1, fk=3, fp=0.9)
    with pytest.raises(RuntimeError):
        oracle.solve_1d(p, np.zeros(10))


# ---------------------------------------------------------------- history sum
def test_time_summing_brute_force(oracle):
    rng = np.random.default_rng(3)
    levels, nodes = 9, 7
    old = rng.standard_normal((levels, nodes))
    kb = rng.random(levels - 1)
    tau = rng.random(levels) + 0.5
    S = oracle.time_summing(old, kb, tau)
    ref = np.zeros(nodes)
    for i in range(nodes):
        for t in range(levels - 2, -1, -1):
            ref[i] += kb[t] * (old[t + 1, i] - old[t, i]) / tau[t]
    assert np.allclose(S, ref, rtol=1e-14, atol=0)


def test_time_summing_restricted_stops_at_first_small_weight(oracle):
    """sum_alg 1 (fpde_hpc_1d.cpp:4117-4131) walks newest first and stops at the
    first |kb| < sum_param, even when older weights are large again: the result
    is the full sum over the newer levels only."""
    rng = np.random.default_rng(5)
    levels, nodes = 9, 6
    old = rng.standard_normal((levels, nodes))
    tau = rng.random(levels) + 0.5
    kb = np.array([5.0, 4.0, 3.0, 1e-9, 2.0, 1.5, 1.2, 1.1])
    S = oracle.time_summing(old, kb, tau, sum_alg=1, sum_param=1e-5)
    newer = oracle.time_summing(old[4:], kb[4:], tau[4:])
    assert np.array_equal(S, newer)
    # a cut of 0 keeps everything; a cut above every weight keeps nothing
    assert np.array_equal(oracle.time_summing(old, kb, tau, sum_alg=1, sum_param=0.0),
                          oracle.time_summing(old, kb, tau))
    assert not np.any(oracle.time_summing(old, kb, tau, sum_alg=1, sum_param=10.0))


def test_restricted_solve_with_zero_cut_is_the_full_solve(oracle):
    u0 = oracle.init_1d(80, 9)
    full = oracle.solve_1d(oracle.make_params(80, 60, 1e-3, 0.6, u_top=u0[0]), u0)
    cut0 = oracle.solve_1d(oracle.make_params(80, 60, 1e-3, 0.6, u_top=u0[0], sum_alg=1, sum_param=0.0), u0)
    assert np.array_equal(full, cut0)
    cut = oracle.solve_1d(oracle.make_params(80, 60, 1e-3, 0.6, u_top=u0[0], sum_alg=1, sum_param=1e-2), u0)
    assert 0 < np.max(np.abs(cut - full)) < 1e-2 * np.max(np.abs(full))


# ------------------------------------------------- series history (sum_alg 2)
@pytest.mark.parametrize("alpha,n", [(0.3, 1), (0.7, 2), (0.8, 40), (0.5, 300)])
def test_series_row_terms_and_stopping_rule(oracle, alpha, n):
    """v3_i and the term count K against the binomial series written directly:
    term_i = (alpha)_i/i! a^-alpha p1 (p1/a)^i v3_i, v3_i = (p1^{i+1}-p0^{i+1})/((i+1) p1^{i+1});
    the series stops after the first |term| <= eps."""
    from scipy.special import gammaln
    tau, eps = 1e-3, 1e-5
    T = oracle.time_grid(n + 1, tau)
    p0, p1, pa = T[n - 1], T[n], T[n + 1]
    K, v3 = oracle.series_row(alpha, p0, p1, pa, eps, cap=100000)
    i = np.arange(K, dtype=np.float64)
    v3d = (1.0 - (p0 / p1) ** (i + 1)) / (i + 1)
    assert np.allclose(v3, v3d, rtol=1e-9, atol=1e-300)
    coef = np.exp(gammaln(alpha + i) - gammaln(alpha) - gammaln(i + 1))     # (alpha)_i / i!
    term = coef * pa ** -alpha * p1 * (p1 / pa) ** i * v3d
    assert abs(term[-1]) <= eps * (1 + 1e-9)
    assert np.all(np.abs(term[:-1]) > eps * (1 - 1e-9))


def _series_run(oracle, levels, alpha, tau, P, eps):
    T = oracle.time_grid(len(levels), tau)
    H = np.zeros((levels.shape[1], P))
    out = []
    for n in range(1, len(levels) - 1):
        K, v3 = oracle.series_row(alpha, T[n - 1], T[n], T[n + 1], eps, cap=P)
        out.append(oracle.series_step(H, v3, K, T[n], T[n + 1], alpha, tau, levels[n - 1], levels[n]))
    return T, out


def test_series_one_moment_telescopes(oracle):
    """P = 1 keeps the zeroth moment sum_t (T_{t+1}-T_t) D_t = tau (U^n - U^0):
    S_n = T_{n+1}^-alpha (U^n - U^0)."""
    rng = np.random.default_rng(8)
    lv = np.cumsum(rng.standard_normal((12, 5)), axis=0)
    T, out = _series_run(oracle, lv, 0.6, 1e-2, 1, 1e-5)
    for n, S in enumerate(out, start=1):
        assert np.allclose(S, T[n + 1] ** -0.6 * (lv[n] - lv[0]), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("alpha", [0.3, 0.75])
def test_l1_exact_for_linear_in_time_on_a_nonuniform_grid(oracle, alpha):
    """L1 integrates the kernel exactly against piecewise-linear u on any grid:
    for u = c t the history plus the local term is c T^(1-a)/Gamma(2-a)."""
    rng = np.random.default_rng(4)
    n, c = 30, 2.5
    taus = 1e-2 * (0.3 + rng.random(n + 1))
    T = np.concatenate([[0.0], np.cumsum(taus)])
    kkk0, kb3 = oracle.l1_weights(alpha, n, T, weight_mode=0)
    old = (c * T[: n + 1])[:, None] * np.ones((1, 3))
    hist = oracle.time_summing(old, kb3, taus)[0]
    approx = (hist + kkk0 * c * taus[n] / taus[n]) / G(1 - alpha)
    exact = c * T[n + 1] ** (1 - alpha) / G(2 - alpha)
    assert abs(approx - exact) <= 1e-11 * exact


def test_uniform_tau_array_is_the_uniform_solve(oracle):
    u0 = oracle.init_1d(60, 3)
    a = oracle.solve_1d(oracle.make_params(60, 40, 1e-3, 0.7, u_top=u0[0]), u0)
    b = oracle.solve_1d(oracle.make_params(60, 40, 1e-3, 0.7, u_top=u0[0], taus=np.full(40, 1e-3)), u0)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("alpha", [0.35, 0.8])
def test_series_converges_to_full_history_sum(oracle, alpha):
    """With many terms and a tiny eps the moment series is the binomial
    expansion of every closed-form weight: it reproduces the sum_alg 0 sum."""
    rng = np.random.default_rng(9)
    tau = 0.25
    lv = np.cumsum(rng.standard_normal((8, 4)), axis=0)
    T, out = _series_run(oracle, lv, alpha, tau, 400, 1e-17)
    for n, S in enumerate(out, start=1):
        _, kb3 = oracle.l1_weights(alpha, n, T, weight_mode=0)
        ref = oracle.time_summing(lv[: n + 1], kb3, np.full(n + 1, tau))
        assert np.allclose(S, ref, rtol=1e-10, atol=1e-12 * np.max(np.abs(ref))), n


@pytest.mark.parametrize("alpha", [0.2, 0.6, 0.9])
def test_l1_exact_for_linear_in_time(oracle, alpha):
    # D^a (c t) = c t^(1-a)/Gamma(2-a): the L1 scheme is exact for linear u
    tau, n, c = 1e-2, 40, 3.0
    T = oracle.time_grid(n + 1, tau)
    kkk0, kb3 = oracle.l1_weights(alpha, n, T, weight_mode=0)
    old = (c * T[: n + 1])[:, None] * np.ones((1, 2))
    hist = oracle.time_summing(old, kb3, np.full(n + 1, tau))[0]
    Da = 1.0 / G(1 - alpha)
    approx = Da * (hist + kkk0 * c * tau / tau)
    exact = c * T[n + 1] ** (1 - alpha) / G(2 - alpha)
    assert abs(approx - exact) <= 1e-11 * exact


@pytest.mark.parametrize("alpha", [0.3, 0.7])
def test_l1_converges_at_order_2_minus_alpha(oracle, alpha):
    errs = []
    for n in (50, 100, 200):
        tau = 1.0 / (n + 1)
        T = oracle.time_grid(n + 1, tau)
        kkk0, kb3 = oracle.l1_weights(alpha, n, T, weight_mode=0)
        u = T ** 2
        hist = oracle.time_summing(u[: n + 1, None], kb3, np.full(n + 1, tau))[0]
        Da = 1.0 / G(1 - alpha)
        approx = Da * (hist + kkk0 * (u[n + 1] - u[n]) / tau)
        exact = 2 * T[n + 1] ** (2 - alpha) / G(3 - alpha)
        errs.append(abs(approx - exact))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert abs(math.log2(r2) - (2 - alpha)) < 0.1 and abs(math.log2(r1) - (2 - alpha)) < 0.15


# ---------------------------------------------------------------- VGM soil
def test_inverse_capacity_is_derivative_of_wetness(oracle):
    for P in (0.05, 0.5, 1.0, 2.0, 7.0):
        d = 1e-6 * P
        dw = (oracle.wetness(P + d) - oracle.wetness(P - d)) / (2 * d)   # d theta / dP = -C
        assert abs(1.0 / oracle.inv_dw_dh(P) + dw) <= 1e-6 * abs(dw)


def test_vgm_limits(oracle):
    assert oracle.wetness(-0.1) == 0.55 and oracle.wetness(0.0) == 0.55
    assert oracle.inv_dw_dh(0.0) == 1e20
    assert abs(oracle.kr(1e-40) - 1.0) < 1e-6          # Kr -> 1 like x^(1-1/n) at saturation
    ks = [oracle.kr(P) for P in (0.1, 0.5, 1.0, 2.0, 5.0)]
    assert all(0 < k < 1 for k in ks) and all(a > b for a, b in zip(ks, ks[1:]))


# Independent van Genuchten-Mualem evaluation in numpy, from the published
# model (van Genuchten 1980, m = 1 - 1/n) with the reference's constants
# n = 1.25, s0 = 0.0736, s1 = 0.55, a = 0.066 1/cm (fpde_hpc_1d.cpp:2513-2518);
# P is the suction head in metres (a * 100 P in 1/cm * cm).
_VG_N, _VG_S0, _VG_S1, _VG_A = 1.25, 0.0736, 0.55, 0.066
_VG_M = 1.0 - 1.0 / _VG_N


def _vg_se(P):
    x = _VG_A * 100.0 * np.maximum(P, 0.0)
    return (1.0 + x ** _VG_N) ** -_VG_M


def _vg_capacity(P):
    """C = -d theta / dP of theta = s0 + (s1 - s0) Se(P)."""
    x = _VG_A * 100.0 * P
    return (_VG_S1 - _VG_S0) * _VG_M * _VG_N * _VG_A * 100.0 * x ** (_VG_N - 1) * (1 + x ** _VG_N) ** (-_VG_M - 1)


def _mualem_kr(se):
    """Mualem (1976): Kr(Se) = Se^(1/2) [int_0^Se dx/h(x) / int_0^1 dx/h(x)]^2,
    h(x) the suction of the retention curve inverted: (x^(-1/m) - 1)^(1/n) / a.
    Integrated numerically (the (1-x)^(-1/n) end point of the denominator
    through quad's algebraic weight)."""
    inv_h = lambda x: _VG_A / (x ** (-1.0 / _VG_M) - 1.0) ** (1.0 / _VG_N)
    smooth = lambda x: inv_h(x) * (1.0 - x) ** (1.0 / _VG_N) if x < 1.0 else _VG_A * _VG_M ** (1.0 / _VG_N)
    den, _ = integrate.quad(smooth, 0.0, 1.0, weight="alg", wvar=(0.0, -1.0 / _VG_N), epsabs=0, epsrel=1e-13,
                            limit=200)
    num, _ = integrate.quad(inv_h, 0.0, se, epsabs=0, epsrel=1e-13, limit=200)
    return se ** 0.5 * (num / den) ** 2


@pytest.mark.parametrize("P", [0.003, 0.05, 0.5, 1.0, 4.0, 20.0])
def test_kr_is_the_mualem_integral(oracle, P):
    """orc_kr (k_form 1, fpde_hpc_1d.cpp:2634-2635) against the Mualem integral
    over the retention curve, and the saturation orc_wetness gives against the
    van Genuchten curve (cv_form 4, :2513-2524)."""
    se = (oracle.wetness(P) - _VG_S0) / (_VG_S1 - _VG_S0)
    assert abs(se - _vg_se(P)) <= 1e-14
    ref = _mualem_kr(se)
    assert abs(oracle.kr(P) - ref) <= 1e-12 * ref
    assert abs(_mualem_kr_closed(P) - ref) <= 1e-12 * ref


def test_inverse_capacity_matches_van_genuchten_derivative(oracle):
    """1/C of inv_dw_dh (:2558-2571) against the analytic derivative of the
    van Genuchten curve written in numpy."""
    for P in (0.01, 0.3, 1.0, 5.0):
        assert abs(oracle.inv_dw_dh(P) * _vg_capacity(P) - 1.0) <= 1e-12


def test_picard_step_is_a_fixed_point_of_its_own_coefficients(oracle):
    """A Picard step run to tol 1e-14 satisfies the tridiagonal rows (A_/B/R,
    :3095-3142) assembled in numpy with K and 1/C at U^{n+1} itself, the
    closed-form history term of the two older intervals and the Dirichlet top
    / zero-flux bottom.  The lagged step (coefficients at U^n, picard_max 1,
    the reference scheme) does not: the check sees the coefficient update."""
    nodes, tau, alpha, fo = 60, 1e-3, 0.7, 10.0
    u0 = oracle.init_1d(nodes, 11)
    h = 1.0 / (nodes - 1)
    z = np.arange(nodes) * h
    Da = 1.0 / G(1 - alpha)
    L = Da * tau ** (1 - alpha) / (1 - alpha) / tau
    n = 2                                            # the step that makes U^3

    def b(t):                                        # b(t, t+1; n+1), g(t) = t
        return tau ** (1 - alpha) * ((n + 1 - t) ** (1 - alpha) - (n - t) ** (1 - alpha)) / (1 - alpha)

    def run(picard_max):
        out = [u0]
        for steps in (1, 2, 3):
            p = oracle.make_params(nodes, steps, tau, alpha, u_top=u0[0], fourier=fo, picard_max=picard_max,
                                   picard_tol=1e-14, weight_mode=0)
            out.append(oracle.solve_1d(p, u0))
        return out, p.ksat

    def residual(U, ksat):
        H = Da * sum(b(t) * (U[t + 1] - U[t]) / tau for t in range(n))
        un, unew = U[n], U[n + 1]
        P = z - unew
        K = ksat * _mualem_kr_closed(P)
        iC = 1.0 / _vg_capacity(P)
        i = np.arange(1, nodes - 1)
        lhs = (K[i - 1] * iC[i] * unew[i - 1] - iC[i] * (K[i] + K[i - 1]) * unew[i] + K[i] * iC[i] * unew[i + 1]) \
            / h ** 2 - L * unew[i]
        return np.max(np.abs(lhs - (-L * un[i] + H[i]))) / (L * np.max(np.abs(unew)))

    U, ksat = run(40)
    assert residual(U, ksat) <= 1e-11
    assert U[3][0] == u0[0] and abs(U[3][-1] - U[3][-2]) <= 1e-14
    Ul, ksat_l = run(1)
    assert residual(Ul, ksat_l) >= 1e-6


def _mualem_kr_closed(P):
    """Kr at suction P by the Mualem integral, vectorised over P (used for
    residual checks: tabulated once per distinct P would be slow, so the
    closed form that the integral test above pins is evaluated)."""
    se = _vg_se(P)
    return np.sqrt(se) * (1 - (1 - se ** (1 / _VG_M)) ** _VG_M) ** 2


@pytest.mark.parametrize("split3d", [0, 1])
def test_3d_sweeps_satisfy_their_rows(oracle, split3d):
    """Each LOD sweep of a VGM 3D step (step 2, so the history term H is not
    0) satisfies its own tridiagonal rows assembled in numpy: the y sweep
    (m3d_c=2, fpde_hpc_1d.cpp:5211-5325) with coefficients at U^n and
    Om = -L U^n + H; the x sweep (:5362-5476) at the field the y sweep left,
    H absent; the z sweep (:5513-5626) at the field the x sweep left, Dirichlet
    top.  With split3d odd (Om1 :4172-4173, :4258-4263) every sweep carries
    H/3.  Zero flux at both ends of the lateral lines, at the bottom of the
    depth lines."""
    n, tau, alpha = 9, 1e-3, 0.75
    u0 = oracle.init_3d(n, 21)
    h = 1.0 / (n - 1)

    def params(M):
        return oracle.make_params(n, M, tau, alpha, u_top=u0[0], fourier=3.0, weight_mode=0, split3d=split3d)

    U = [u0] + [oracle.solve_3d(params(k), u0) for k in (1, 2)]
    u3, sw = oracle.solve_3d_trace(params(3), u0)
    assert np.array_equal(u3, sw[2])
    ksat = params(1).ksat
    Da = 1.0 / G(1 - alpha)
    L = Da * tau ** (1 - alpha) / (1 - alpha) / tau
    step = 2
    b = lambda t: tau ** (1 - alpha) * ((step + 1 - t) ** (1 - alpha) - (step - t) ** (1 - alpha)) / (1 - alpha)
    H = Da * sum(b(t) * (U[t + 1] - U[t]) / tau for t in range(step))

    def lines(F, axis):                       # idx = z + n (x + n y): F3[y][x][z]
        F3 = F.reshape(n, n, n)
        if axis == "y":
            return F3.transpose(2, 1, 0).reshape(n * n, n)
        if axis == "x":
            return F3.transpose(2, 0, 1).reshape(n * n, n)
        return F3.reshape(n * n, n)

    def depth(axis):
        zz = np.arange(n) * h
        if axis == "z":
            return np.ones((n * n, 1)) * zz[None, :]
        return np.repeat(zz, n)[:, None] * np.ones((1, n))     # lateral lines lie at one depth

    def residual(axis, before, after, Hs):
        ub, ua, Hl, Z = lines(before, axis), lines(after, axis), lines(Hs, axis), depth(axis)
        P = Z - ub
        K = ksat * _mualem_kr_closed(P)
        iC = 1.0 / _vg_capacity(P)
        i = np.arange(1, n - 1)
        lhs = (K[:, i - 1] * iC[:, i] * ua[:, i - 1] - iC[:, i] * (K[:, i] + K[:, i - 1]) * ua[:, i]
               + K[:, i] * iC[:, i] * ua[:, i + 1]) / h ** 2 - L * ua[:, i]
        top = ua[:, 1] if axis != "z" else lines(u0, "z")[:, 0]
        assert np.max(np.abs(ua[:, 0] - top)) <= 1e-14
        assert np.max(np.abs(ua[:, -1] - ua[:, -2])) <= 1e-14
        return np.max(np.abs(lhs - (-L * ub[:, i] + Hl[:, i]))) / (L * np.max(np.abs(ua)))

    share = H / 3.0 if split3d else H
    rest = H / 3.0 if split3d else 0.0 * H
    assert residual("y", U[2], sw[0], share) <= 1e-12
    assert residual("x", sw[0], sw[1], rest) <= 1e-12
    assert residual("z", sw[1], sw[2], rest) <= 1e-12
    # the check has teeth: H is visible against the rows' scale
    assert np.max(np.abs(H)) >= 1e-4 * L * np.max(np.abs(u3))


# ---------------------------------------------------------------- linear algebra
def test_thomas_matches_dense_solve(oracle):
    rng = np.random.default_rng(5)
    for n in (2, 5, 64, 301):
        a = rng.random(n); c = rng.random(n); b = -(a + c + rng.random(n) + 0.1)
        a[0] = 0.0; c[-1] = 0.0
        d = rng.standard_normal(n)
        M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.allclose(oracle.thomas(a, b, c, d), np.linalg.solve(M, d), rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("n", [4, 16, 32])
def test_gauss_legendre_matches_numpy(oracle, n):
    x, w = oracle.gauss_legendre(n, 0.1, 1.0)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, 0.55 + 0.45 * xr, atol=1e-14)
    assert np.allclose(w, 0.45 * wr, atol=1e-14)


# ---------------------------------------------------------------- solvers
def _mittag_leffler(alpha, z, terms=120):
    return sum(z ** k / math.gamma(alpha * k + 1) for k in range(terms))


def test_1d_linear_solve_converges_to_mittag_leffler(oracle):
    """K=1, 1/C=1: D^a u = u_zz on [0,1], u(0)=u_top, u_z(1)=0,
    u0 = u_top + A sin(pi z/2)  ->  u = u_top + A E_a(-(pi/2)^2 t^a) sin(pi z/2)."""
    alpha, A, ut, T = 0.6, 0.1, -1.0, 0.2
    errs = []
    for nodes, steps in ((81, 100), (161, 200)):
        z = np.linspace(0, 1, nodes)
        u0 = ut + A * np.sin(np.pi * z / 2)
        p = oracle.make_params(nodes, steps, T / steps, alpha, coef_mode=1, weight_mode=0)
        u = oracle.solve_1d(p, u0)
        exact = ut + A * _mittag_leffler(alpha, -(np.pi / 2) ** 2 * T ** alpha) * np.sin(np.pi * z / 2)
        errs.append(np.max(np.abs(u - exact)))
    assert errs[1] < 0.6 * errs[0] and errs[1] < 2e-3


def test_vgm_step_satisfies_discrete_equation(oracle):
    """One reference step: the result satisfies the tridiagonal rows built
    from the lagged coefficients (independent residual check)."""
    nodes, tau, alpha = 50, 1e-3, 0.8
    u0 = oracle.init_1d(nodes, 3)
    p = oracle.make_params(nodes, 1, tau, alpha, u_top=u0[0])
    u1 = oracle.solve_1d(p, u0)
    h = 1.0 / (nodes - 1)
    z = np.arange(nodes) * h
    K = np.array([p.ksat * oracle.kr(zz - uu) for zz, uu in zip(z, u0)])
    iC = np.array([oracle.inv_dw_dh(zz - uu) for zz, uu in zip(z, u0)])
    L = (1 / G(1 - alpha)) * tau ** (1 - alpha) / (1 - alpha) / tau
    for i in range(1, nodes - 1):
        lhs = (K[i - 1] * iC[i] * u1[i - 1] - iC[i] * (K[i] + K[i - 1]) * u1[i] + K[i] * iC[i] * u1[i + 1]) / h ** 2 \
            - L * u1[i]
        assert abs(lhs - (-L * u0[i])) <= 1e-9 * L
    assert u1[0] == u0[0]
    assert abs(u1[-1] - u1[-2]) <= 1e-12


@pytest.mark.parametrize("gam", [0.0, 0.7])
def test_3d_with_laterally_uniform_field_equals_1d(oracle, gam):
    """With field-independent coefficients a laterally uniform 3D field
    follows the 1D scheme along depth (the y sweep adds the history term, the
    x sweep is an identity; the space-fractional sum of a lateral line of a
    uniform field is 0).  With VGM coefficients the two differ at O(H/L)
    because the depth sweep evaluates its coefficients after the y sweep; the
    space-fractional depth sweep likewise forms its explicit sum from the
    field after the y sweep (U^n + H/L), so with gamma < 1 the reduction is
    exact at the first step, where H = 0."""
    n, steps = 10, (12 if gam == 0.0 else 1)
    col = oracle.init_1d(n, 7)
    u3 = np.tile(col, n * n)            # z fastest
    p3 = oracle.make_params(n, steps, 1e-3, 0.85, coef_mode=1, gamma=gam)
    p1 = oracle.make_params(n, steps, 1e-3, 0.85, coef_mode=1, gamma=gam)
    r3 = oracle.solve_3d(p3, u3).reshape(n * n, n)
    r1 = oracle.solve_1d(p1, col)
    assert np.max(np.abs(r3 - r1[None, :])) <= 1e-12


def test_distributed_order_derivative_closed_form(oracle):
    """int_{0.1}^{1} Gamma(3-a)/2 * D^a t^2 da = (t^1.9 - t)/ln t (at t = 1: 0.9)."""
    M, tau = 400, 1.0 / 400
    for rule, q, tol in ((0, 32, 2e-3), (1, 65, 4e-3)):
        p = oracle.make_params(5, M, tau, do_q=q, do_rule=rule, do_a0=0.1, do_a1=1.0, weight_mode=0,
                               ksat=1.0)
        local, omega, corr = oracle.folded_weights(p)
        T = np.arange(M + 1) * tau
        u = T ** 2
        n = M - 1                        # derivative at T[M] = 1
        d = np.diff(u)
        approx = local * d[n] + sum(omega[n - t - 1] * d[t] for t in range(n))
        exact = 0.9                      # limit of (t^1.9 - t)/ln t at t = 1
        assert abs(approx - exact) <= tol * exact


def test_folded_weights_equal_per_order_sums(oracle):
    M, tau = 60, 1e-3
    p = oracle.make_params(5, M, tau, do_q=6, do_rule=0, weight_mode=1, ksat=1.0)
    local, omega, corr = oracle.folded_weights(p)
    x, w = oracle.gauss_legendre(6, 0.1, 1.0)
    T = np.arange(M + 2) * tau
    n = M - 1
    for t in range(n):
        ref = 0.0
        for a, wq in zip(x, w):
            kb = oracle.lib().orc_kb(a, t, n + 1, T.ctypes.data_as(oracle._dp), 1, 1e-5)
            ref += wq * G(3 - a) / 2 / G(1 - a) * kb / tau
        got = omega[n - t - 1] + (corr[n, t] if t < 2 else 0.0)
        assert abs(got - ref) <= 1e-12 * abs(ref)


# ------------------------------------------------- space-fractional operator
@pytest.mark.parametrize("gam", [0.3, 0.8])
def test_space_sum_telescopes_for_constant_flux(oracle, gam):
    """F = 1: the closed-form weights of row i sum to
    Dg * int_0^{Z_{i-1}} (Z_i - s)^-g ds = Dg (Z_i^(1-g) - h^(1-g)) / (1-g)."""
    n = 300
    h = 1.0 / (n - 1)
    F = np.zeros(n)
    F[1:n - 1] = 1.0
    V = oracle.space_v(F, h, gam, weight_mode=0)
    Z = np.arange(n) * h
    ref = np.zeros(n)
    ref[1:n - 1] = (Z[1:n - 1] ** (1 - gam) - h ** (1 - gam)) / (1 - gam) / G(1 - gam)
    assert np.max(np.abs(V - ref)) <= 1e-13 * np.max(np.abs(ref))
    assert V[0] == 0.0 and V[-1] == 0.0


@pytest.mark.parametrize("gam", [0.25, 0.7])
def test_space_sum_is_the_integral_of_the_piecewise_constant_flux(oracle, gam):
    """V_i = Dg int_0^{Z_{i-1}} (Z_i - s)^-g f(s) ds, f = F[j] on (Z_{j-1}, Z_j):
    the interval left of node j carries F[j] (an index shift fails this)."""
    n = 24
    h = 0.5 / (n - 1)
    rng = np.random.default_rng(17)
    F = np.zeros(n)
    F[1:n - 1] = rng.standard_normal(n - 2)
    Z = np.arange(n) * h
    ref = np.zeros(n)
    for i in range(1, n - 1):
        for j in range(1, i):
            w, _ = integrate.quad(lambda s: (Z[i] - s) ** -gam, Z[j - 1], Z[j], epsabs=1e-15, epsrel=1e-13)
            ref[i] += w * F[j]
    ref /= G(1 - gam)
    V0 = oracle.space_v(F, h, gam, weight_mode=0)
    assert np.max(np.abs(V0 - ref)) <= 1e-11 * np.max(np.abs(ref))
    # the reference series for the two intervals next to the surface (j <= 2)
    # agrees with the same integrals to the series tolerance
    V1 = oracle.space_v(F, h, gam, weight_mode=1)
    assert np.max(np.abs(V1 - ref)) <= 1e-4 * np.max(np.abs(ref))
    assert not np.array_equal(V0, V1)


def test_space_flux_of_quadratic_head_with_linear_conductivity(oracle):
    """u = x^2, K = x: F_j = (K_j (u_{j+1}-u_j) - K_{j-1} (u_j-u_{j-1})) / h^2
    = 4 x_j - h exactly (continuum: (x u')' = 4x)."""
    n = 50
    h = 1.0 / (n - 1)
    x = np.arange(n) * h
    F = oracle.space_flux(x, x ** 2, h)
    assert np.allclose(F[1:n - 1], 4 * x[1:n - 1] - h, rtol=1e-11, atol=1e-11)
    assert F[0] == 0.0 and F[-1] == 0.0


@pytest.mark.parametrize("gam", [0.4, 0.85])
def test_space_fractional_step_equals_dense_assembly(oracle, gam):
    """One step, linear coefficients (K = 1/C = 1): the oracle's Thomas solve
    equals the dense system assembled from quadrature weights and solved by
    numpy: L (U_i - U0_i) = kap (U_{i+1} - 2U_i + U_{i-1})/h^2 + V_i(F(U0)),
    kap = h^(1-g)/Gamma(2-g), L = tau^-a/Gamma(2-a), Dirichlet top, U_N = U_{N-1}."""
    n, tau, alpha = 20, 1e-3, 0.6
    h = 1.0 / (n - 1)
    rng = np.random.default_rng(23)
    u0 = np.sin(np.arange(n) * 0.4) + 0.1 * rng.standard_normal(n)
    p = oracle.make_params(n, 1, tau, alpha, coef_mode=1, weight_mode=0, gamma=gam)
    u1 = oracle.solve_1d(p, u0)
    Z = np.arange(n) * h
    F0 = np.zeros(n)
    F0[1:n - 1] = (u0[2:] - 2 * u0[1:n - 1] + u0[:n - 2]) / h ** 2
    V = np.zeros(n)
    for i in range(1, n - 1):
        for j in range(1, i):
            w, _ = integrate.quad(lambda s: (Z[i] - s) ** -gam, Z[j - 1], Z[j], epsabs=1e-15, epsrel=1e-13)
            V[i] += w * F0[j] / G(1 - gam)
    kap = h ** (1 - gam) / G(2 - gam)
    L = tau ** -alpha / G(2 - alpha)
    A = np.zeros((n, n))
    b = np.zeros(n)
    A[0, 0] = 1.0
    b[0] = u0[0]
    for i in range(1, n - 1):
        A[i, i - 1] = A[i, i + 1] = kap / h ** 2
        A[i, i] = -2 * kap / h ** 2 - L
        b[i] = -L * u0[i] - V[i]
    A[n - 1, n - 1], A[n - 1, n - 2] = 1.0, -1.0
    ref = np.linalg.solve(A, b)
    assert np.max(np.abs(u1 - ref)) <= 1e-11 * np.max(np.abs(ref))


def test_space_fractional_tends_to_classical_as_gamma_to_one(oracle):
    """Dg kb(g,i-1,i,i) = h^(1-g)/Gamma(2-g) -> 1 and Dg = 1/Gamma(1-g) -> 0:
    the scheme becomes the classical one as g -> 1."""
    n, M = 80, 30
    u0 = oracle.init_1d(n, 5)
    base = oracle.solve_1d(oracle.make_params(n, M, 1e-3, 0.7, u_top=u0[0]), u0)
    near = oracle.solve_1d(oracle.make_params(n, M, 1e-3, 0.7, u_top=u0[0], gamma=1 - 1e-7), u0)
    far = oracle.solve_1d(oracle.make_params(n, M, 1e-3, 0.7, u_top=u0[0], gamma=0.8), u0)
    d_near = np.max(np.abs(near - base)) / np.max(np.abs(base))
    d_far = np.max(np.abs(far - base)) / np.max(np.abs(base))
    assert d_near <= 1e-7
    assert d_far >= 1e3 * d_near


def test_3d_space_fractional_lateral_sweeps_are_symmetric(oracle):
    """One step, linear coefficients: a field varying along x only and the
    same profile along y only give transposed results (the x and y sweeps
    apply the same line operator; the history term is 0 at step 0), and the
    space-fractional sum changes both."""
    n = 12
    prof = np.cos(np.arange(n) * 0.5) + 0.2 * np.sin(np.arange(n) * 1.7)
    z, x, y = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    fx = prof[x].transpose(2, 1, 0).ravel()      # idx = z + n (x + n y)
    fy = prof[y].transpose(2, 1, 0).ravel()
    p = oracle.make_params(n, 1, 1e-3, 0.85, coef_mode=1, gamma=0.7)
    rx = oracle.solve_3d(p, fx).reshape(n, n, n)  # [y][x][z]
    ry = oracle.solve_3d(p, fy).reshape(n, n, n)
    assert np.max(np.abs(rx - ry.transpose(1, 0, 2))) <= 1e-13
    r0 = oracle.solve_3d(oracle.make_params(n, 1, 1e-3, 0.85, coef_mode=1), fx)
    assert np.max(np.abs(rx.ravel() - r0)) >= 1e-4


# ------------------------------------------- fr_do.cpp distributed-order equation
def test_do_equation_converges_to_manufactured_solution(oracle):
    """eq_solve's Crank-Nicolson scheme for int_0^1 Gamma(5-a) D^a u da =
    2 u_xx + f on [0, 0.5]^2 (fr_do.cpp:733-775) against its exact solution
    100 t^4 sin(2 pi x) (eq_anal_u, the source of :458-465): the RMS error of
    the last level falls with refinement."""
    errs = []
    for nx, nt in ((32, 160), (64, 320), (128, 640)):
        _, err = oracle.doeq_solve(oracle.doeq_params(nx=nx, na=64, tblock=32, nt=nt))
        errs.append(err)
    assert errs[0] < 0.012 and errs[1] < 0.7 * errs[0] and errs[2] < 0.7 * errs[1]


def test_do_equation_block_scheme_is_exact(oracle):
    """The block scheme (history before the block at its start, the in-block
    intervals added level by level, :739-761) is a reordering of the same
    sums: blocks of 16 and of 64 levels give the same solution."""
    for ex in (0, 1):
        a, _ = oracle.doeq_solve(oracle.doeq_params(nx=40, na=33, tblock=16, nt=192, example=ex))
        b, _ = oracle.doeq_solve(oracle.doeq_params(nx=40, na=33, tblock=64, nt=192, example=ex))
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))
        assert np.max(np.abs(a[:, -1])) > 0.0


def test_frdo_derivative_of_t4_is_the_closed_form(oracle):
    """DO_Cder with eq_w = Gamma(5-a): int_0^1 Gamma(5-a) D^a t^4 da =
    24 (t^4 - t^3)/ln t (example 0's source term, fr_do.cpp:458-465)."""
    n, T = 2000, 0.5
    tau = T / (n - 1)
    t = np.arange(n) * tau
    DF = oracle.do_cder(t ** 4, tau, 128, example=0)
    exact = 24 * (T ** 4 - T ** 3) / np.log(T)
    assert abs(DF[-1] - exact) <= 2e-3 * abs(exact)


def test_do_equation_gpu_text_source(oracle):
    """The GPU program text forms the source with pi as a float (eq_ff,
    :256-263): the solution moves by no more than that rounding."""
    a, _ = oracle.doeq_solve(oracle.doeq_params(nx=32, na=32, tblock=32, nt=160))
    b, _ = oracle.doeq_solve(oracle.doeq_params(nx=32, na=32, tblock=32, nt=160, gpu_pi=1))
    d = np.max(np.abs(a - b)) / np.max(np.abs(a))
    assert 0.0 < d < 1e-6
