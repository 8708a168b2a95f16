"""The allocation-free plain history sum (fpde_history_sum_f64_noalloc)
against the allocating entry point (fpde_history_sum_f64) and numpy, through
the C ABI: two calls in a row, the second smaller, bitwise equal to the
allocating call; an oversize level stride is refused before any launch."""
import numpy as np
import pytest

from test_gpu_parity import Dev

pytestmark = pytest.mark.gpu


def _case(L, rng, n, nodes):
    ld = ((nodes + 63) // 64) * 64
    D = np.zeros((max(n, 1), ld))
    D[:n, :nodes] = rng.standard_normal((n, nodes))
    omega = rng.random(n + 8)
    ref = np.zeros(nodes)
    for j in range(1, n + 1):
        ref += omega[j - 1] * D[n - j, :nodes]
    dv = Dev(L)
    try:
        bD, bw = dv.put(D), dv.put(omega)
        bS1, bS2 = dv.put(np.full(ld, 7.0)), dv.put(np.full(ld, -7.0))
        rc = L.fpde_history_sum_f64(dv.ptr(bS1), dv.ptr(bD), n, nodes, ld, dv.ptr(bw))
        assert rc == 0, L.fpde_last_error()
        rc = L.fpde_history_sum_f64_noalloc(dv.ptr(bS2), dv.ptr(bD), n, nodes, ld, dv.ptr(bw))
        assert rc == 0, L.fpde_last_error()
        return dv.get(bS1, nodes), dv.get(bS2, nodes), ref
    finally:
        dv.close()


def test_noalloc_history_sum_equals_allocating_call(fpde):
    L = fpde.lib()
    rng = np.random.default_rng(11)
    # a large call (64 partial rows, 40 node tiles), then a smaller one that
    # reads fewer rows of the same scratch array; then no levels at all
    for n, nodes in [(3000, 20000), (37, 900), (0, 130)]:
        print(f"noalloc history sum: n={n} nodes={nodes}", flush=True)
        S_alloc, S_free, ref = _case(L, rng, n, nodes)
        assert np.array_equal(S_free, S_alloc)
        if n:
            assert np.max(np.abs(S_free - ref)) <= 1e-12 * np.max(np.abs(ref))
        else:
            assert not np.any(S_free)


def test_noalloc_history_sum_refuses_oversize_levels(fpde):
    L = fpde.lib()
    rc = L.fpde_history_sum_f64_noalloc(8, 8, 4, 70000, 70016, 8)
    assert rc == 1 and b"65536" in L.fpde_last_error()
