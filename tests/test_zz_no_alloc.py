"""kdq_measure_noalloc (working memory in the library's static per-device array) against
kdq_measure (working memory allocated with the handle), through the C ABI: two calls in a
row on one batch, then a call on a smaller batch, and the size limit. Bar: rows within 1e-12
relative (the same kernels in the same order)."""
import ctypes as C

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def _batch(ms, seeds):
    from kapdqmc_amd.dqmc import DQMCBatch, Initial_s
    b = DQMCBatch(ms, seeds, list(range(len(ms))), delay=16)
    for c, m in enumerate(ms):
        b.set_fields(c, Initial_s(m, seeds[c], c))
    b.init_chain()
    return b


def tu(L, U=4.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[L, L], BatchSize=5)


def test_noalloc_measure_matches_measure_two_calls_then_smaller():
    print("case: noalloc measure, t-U L=6/4 batch then L=3", flush=True)
    big = _batch([tu(6), tu(4, U=3.0)], [3, 4])
    for lt in (7, 15):                                   # wraps forward, then backward, to Nt/2
        ref = big.measure(lt)
        got = big.measure_noalloc(lt)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-14), (lt, got, ref)
    big.phy_update(1, False)
    assert np.allclose(big.measure_noalloc(3), big.measure(3), rtol=1e-12, atol=1e-14)
    big.close()
    small = _batch([tu(3)], [5])                         # smaller input: reuses the start of the array
    for lt in (2, 10):
        ref = small.measure(lt)
        got = small.measure_noalloc(lt)
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-14), (lt, got, ref)
    small.close()


def test_noalloc_measure_tv():
    print("case: noalloc measure, t-V L=3", flush=True)
    m = tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                        site=[3, 3], BatchSize=5)
    b = _batch([m, m], [8, 9])
    for lt in (4, 13):
        assert np.allclose(b.measure_noalloc(lt), b.measure(lt), rtol=1e-12, atol=1e-14)
    b.close()


def test_noalloc_measure_rejects_batches_above_the_limit():
    print("case: noalloc measure size limit", flush=True)
    from kapdqmc_amd import _lib
    m = tu(12)                                           # 4 x 288² doubles = 2.65 MB per chain
    b = _batch([m] * 16, list(range(16)))                # 42 MB > 32 MiB
    out = np.zeros((16, 6))
    rc = b.L.kdq_measure_noalloc(b.h, 10, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == -1
    assert b"limit" in b.L.kdq_last_error(b.h)
    b.close()
