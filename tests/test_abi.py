"""CPU checks of the C ABI: the library loads, exports every symbol the
public header declares, host-side logic agrees with the oracle, and invalid
specs are rejected before any device work."""
import ctypes
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "fpde_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fpde_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol(fpde):
    L = fpde.lib()
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_header(fpde):
    from fpde_hpc_amd._lib import SIGNATURES
    bound = {n for n, _, _ in SIGNATURES}
    assert set(declared_symbols()) == bound


def test_init_conditions_match_oracle(fpde, oracle):
    for n, seed in [(1000, 1000), (10000, 1001), (37, 5)]:
        assert np.array_equal(fpde.init_1d(n, seed), oracle.init_1d(n, seed))
    for n, seed in [(16, 1003), (9, 7)]:
        assert np.array_equal(fpde.init_3d(n, seed), oracle.init_3d(n, seed))


def test_ksat_fourier_matches_oracle(fpde, oracle):
    for ut, h, tau in [(-1.0, 1e-3, 1e-3), (-0.95, 1e-4, 5e-4), (-1.07, 2.5e-4, 1e-3)]:
        a = fpde.lib().fpde_ksat_for_fourier(ut, h, tau, 1.0)
        b = oracle.ksat_for_fourier(ut, h, tau, 1.0)
        assert abs(a - b) <= 1e-14 * abs(b)


def test_product_vgm_matches_closed_form(fpde):
    """The library's K/C at suction P (through fpde_ksat_for_fourier, which
    returns fo h^2 / (tau Kr(P) / C(P))) against the van Genuchten-Mualem
    closed form in 50-digit arithmetic: Kr = Se^(1/2) (1 - (x/(1+x))^m)^2,
    C = (n-1)(s1-s0) x Se / (P (1+x)), x = (100 a P)^n, Se = (1+x)^-m, over
    wet to very dry soil (the dry end is where a difference of logarithms
    would lose digits)."""
    import mpmath
    mpmath.mp.dps = 50
    n, s0, s1, a = mpmath.mpf(5) / 4, mpmath.mpf("0.0736"), mpmath.mpf("0.55"), mpmath.mpf("0.066")
    m = 1 - 1 / n
    for P in np.logspace(-6, 3, 400):
        Pm = mpmath.mpf(float(P))
        x = (100 * a * Pm) ** n
        se = (1 + x) ** (-m)
        kr = mpmath.sqrt(se) * (1 - (x / (1 + x)) ** m) ** 2
        C = (n - 1) * (s1 - s0) * x * se / (Pm * (1 + x))
        ratio = fpde.lib().fpde_ksat_for_fourier(-float(P), 1.0, 1.0, 1.0)     # = 1 / (Kr / C)
        exact = float(C / kr)
        assert abs(ratio - exact) <= 1e-14 * exact, P


@pytest.mark.parametrize("bad", [
    dict(dims=2), dict(nodes=3), dict(steps=0), dict(alpha=0.0), dict(alpha=1.5),
    dict(picard_max=0), dict(hist_mode=9), dict(block=48), dict(dims=3, nodes=300),
    dict(tau=-1.0), dict(nodes=70000), dict(sum_alg=3), dict(sum_alg=1, sum_param=-1.0),
    dict(sum_alg=2, sum_param=0.0, hist_mode=0), dict(sum_alg=2, sum_param=65.0, hist_mode=0),
    dict(sum_alg=2, sum_param=2.5, hist_mode=0), dict(sum_alg=2, sum_param=8.0, hist_mode=1),
    dict(sum_alg=2, sum_param=8.0, hist_mode=0, do_q=4),
    dict(gamma=1.5), dict(gamma=-0.1),
    dict(chunk_rows=16), dict(chunk_rows=300), dict(nodes=20000, chunk_rows=64),
])
def test_invalid_specs_rejected(fpde, bad):
    kw = dict(nodes=100, steps=10, tau=1e-3, alpha=0.5, hist_mode=1, block=64)
    kw.update(bad)
    dims = kw.pop("dims", 1)
    spec = fpde.make_spec(kw.pop("nodes"), kw.pop("steps"), kw.pop("tau"), dims=dims, ksat=1.0, **kw)
    h = ctypes.c_void_p()
    rc = fpde.lib().fpde_solver_create(ctypes.byref(spec), ctypes.byref(h))
    assert rc == 1
    assert fpde.lib().fpde_last_error()


def test_footprint_scales_with_history(fpde):
    a = fpde.footprint(fpde.make_spec(10000, 20000, 5e-4, hist_mode="plain", ksat=1.0))
    b = fpde.footprint(fpde.make_spec(10000, 20000, 5e-4, hist_mode="split_bf16", ksat=1.0))
    assert a >= 8 * 10000 * 20000
    assert b >= a + 4 * 10000 * 20000 - 1
    # series moments (sum_alg 2): P values per node instead of M levels
    c = fpde.footprint(fpde.make_spec(10000, 20000, 5e-4, hist_mode="plain", ksat=1.0, sum_alg=2, sum_param=32))
    assert c < a / 50


def test_unknown_kernel_name_is_reported(fpde):
    # program/kernel objects need no device
    L = fpde.lib()
    prog = ctypes.c_void_p()
    assert L.fpde_cl_program_create(ctypes.c_void_p(1), b"#define OS 64\n", ctypes.byref(prog)) == 0
    k = ctypes.c_void_p()
    assert L.fpde_cl_kernel_create(prog, b"Nope", ctypes.byref(k)) == 4
    assert b"Sum" in L.fpde_last_error()
    assert L.fpde_cl_kernel_create(prog, b"Sum", ctypes.byref(k)) == 0
    L.fpde_cl_kernel_destroy(k)
    L.fpde_cl_program_destroy(prog)


@pytest.mark.parametrize("name", ["fpde_spec", "fpde_stats", "fpde_result_info"])
def test_struct_layout_matches_binding(fpde, name):
    """The ctypes structures have the C header's size and field offsets (gcc)."""
    from fpde_hpc_amd import _lib
    cls = {"fpde_spec": _lib.Spec, "fpde_stats": _lib.Stats, "fpde_result_info": _lib.ResultInfo}[name]
    fields = [f for f, _ in cls._fields_]
    body = "".join(f'printf("%zu\\n", offsetof({name}, {f}));' for f in fields)
    src = ('#include <stdio.h>\n#include "fpde_hip.h"\nint main(void){printf("%zu\\n", sizeof('
           + name + '));' + body + 'return 0;}\n')
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), c, "-o", exe])
        out = [int(v) for v in subprocess.check_output([exe]).split()]
    assert out[0] == ctypes.sizeof(cls)
    assert out[1:] == [getattr(cls, f).offset for f in fields]


@pytest.mark.parametrize("args", [
    (0, 0, 100, 0.01, 0.5, 1, 1e-5), (8, 8, 3, 0.01, 0.5, 1, 1e-5), (8, 8, 100, 0.0, 0.5, 1, 1e-5),
    (8, 8, 100, 0.01, 1.0, 1, 1e-5), (8, 8, 100, 0.01, 0.0, 1, 1e-5), (8, 8, 100, 0.01, 0.5, 2, 1e-5),
])
def test_space_sum_rejects_bad_arguments(fpde, args):
    """Argument checks run before any device call (no GPU needed)."""
    rc = fpde.lib().fpde_space_sum(*args)
    assert rc == 1
    assert fpde.lib().fpde_last_error()


def test_autotune_block_argument_checks_and_plain_form(fpde):
    """Null arguments and invalid specs are refused before any device call;
    hist_mode 0 has no history block, so no trial runs and the block stays."""
    L = fpde.lib()
    spec = fpde.make_spec(64, 300, 1e-3, hist_mode="split_bf16", block=0)
    u0 = np.zeros(64)
    b, ms = ctypes.c_int(-1), ctypes.c_double(-1.0)
    assert L.fpde_autotune_block(None, fpde.solver._p(u0), 0, ctypes.byref(b), None) == 1
    assert L.fpde_autotune_block(ctypes.byref(spec), None, 0, ctypes.byref(b), None) == 1
    assert L.fpde_autotune_block(ctypes.byref(spec), fpde.solver._p(u0), 0, None, None) == 1
    bad = fpde.make_spec(64, 300, 1e-3, hist_mode="split_bf16", block=48)
    assert L.fpde_autotune_block(ctypes.byref(bad), fpde.solver._p(u0), 0, ctypes.byref(b), None) == 1
    plain = fpde.make_spec(64, 300, 1e-3, hist_mode="plain", block=0)
    assert L.fpde_autotune_block(ctypes.byref(plain), fpde.solver._p(u0), 0, ctypes.byref(b), ctypes.byref(ms)) == 0
    assert b.value == 0 and ms.value == 0.0


def test_advance_batch_argument_checks(fpde):
    """A null array, a count outside 1..32 and a null solver are refused
    before any device call."""
    L = fpde.lib()
    null1 = (ctypes.c_void_p * 1)(None)
    assert L.fpde_solver_advance_batch(None, 1, 0) == 1
    assert L.fpde_solver_advance_batch(null1, 0, 0) == 1
    many = (ctypes.c_void_p * 33)()
    assert L.fpde_solver_advance_batch(many, 33, 0) == 1
    assert L.fpde_solver_advance_batch(null1, 1, 0) == 1
    assert b"null solver" in L.fpde_last_error()
