"""One batch whose chains differ in Nt (the uneven configuration: honeycomb L with 2Θ = 2L),
through the C ABI.

* each chain of a mixed-Nt batch (L = 2/3/4 with 2Θ = 2L, U = 2/4/6, chains interleaved so
  that the Nt groups are not contiguous in the batch) equals the oracle's sweep: fields bit
  for bit, G to 1e-10, rows to 1e-9;
* the same chains in separate one-Nt handles give identical fields, G, rows, counters, log
  W and health rows (the grouped batch only routes and orders);
* replayed draws and set/get of G reach the right chain;
* per-slice calls (wrap, update_layer, measure, measure_noalloc) are refused with a message.
The t-V kind and mixed precision go through the same routing (one check each)."""
import ctypes as C

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dq():
    from kapdqmc_amd import dqmc
    return dqmc


def tu(L, U):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=float(L) / 2, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=5)


# batch order: L = 3, 2, 4, 2, 3 (Nt = 30, 20, 40, 20, 30)
LS, US = [3, 2, 4, 2, 3], [4.0, 2.0, 6.0, 6.0, 2.0]
SEEDS = [11, 12, 13, 14, 15]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def test_mixed_nt_batch_matches_oracle(dq, oracle_mod):
    print("case: mixed-Nt batch L=3,2,4,2,3 vs oracle", flush=True)
    ms = [tu(L, U) for L, U in zip(LS, US)]
    assert len({m.Nt for m in ms}) == 3
    b = dq.DQMCBatch(ms, SEEDS, list(range(5)), delay=8)
    os_ = []
    for c, m in enumerate(ms):
        o = oracle_mod.OracleTU(m, seed=SEEDS[c], chain=c)
        o.initial_s()
        os_.append(o)
        b.set_fields(c, o.s)
    rb = b.phy_update(2, True)
    assert rb.shape == (5, 2, b.nobs)
    for c, o in enumerate(os_):
        ro = o.phy_update(2, True)
        assert np.array_equal(b.get_fields(c), o.s), c
        assert b.near_ties()[c] == o.near_ties
        assert rel(b.get_G(c), o.G) < 1e-10, c
        assert np.allclose(rb[c], ro, rtol=1e-9, atol=1e-11), c
    b.close()


def test_mixed_nt_batch_equals_separate_handles(dq):
    print("case: mixed-Nt batch vs one handle per Nt", flush=True)
    ms = [tu(L, U) for L, U in zip(LS, US)]
    b = dq.DQMCBatch(ms, SEEDS, list(range(5)), delay=8)
    b.initial_s()
    b.phy_update(0, True)          # zero sweeps (graph capture only) is a valid call
    rb = b.phy_update(3, True)
    sep = {}
    for L in sorted(set(LS)):
        idx = [c for c in range(5) if LS[c] == L]
        h = dq.DQMCBatch([ms[c] for c in idx], [SEEDS[c] for c in idx], idx, delay=8)
        h.initial_s()
        sep[L] = (idx, h, h.phy_update(3, True))
    lw = b.log_weight()
    cnt = b.counters()
    health = b.health()
    for L, (idx, h, rows) in sep.items():
        lw_h, cnt_h, health_h = h.log_weight(), h.counters(), h.health()
        for k, c in enumerate(idx):
            assert np.array_equal(b.get_fields(c), h.get_fields(k)), c
            assert np.array_equal(b.get_G(c), h.get_G(k)), c
            assert np.array_equal(rb[c], rows[k]), c
            assert lw[c] == lw_h[k] and np.array_equal(cnt[c], cnt_h[k]), c
            assert np.array_equal(health[c], health_h[k]), c
        h.close()
    # replayed draws reach the right chain: chain 2 (L = 4) rejects everything for one sweep
    c = 2
    s0 = b.get_fields(c)
    from kapdqmc_amd.dqmc import elements_per_slice
    m = ms[c]
    ch = np.zeros((1, 2, m.Nt, elements_per_slice(m)), dtype=np.uint8)
    b.set_replay(c, ch, np.full(ch.shape, np.inf))
    others = [b.get_fields(k) for k in range(5)]
    b.phy_update(1, False)
    assert np.array_equal(b.get_fields(c), s0)
    assert sum(not np.array_equal(b.get_fields(k), others[k]) for k in range(5) if k != c) == 4
    # set_G / get_G route to the chain
    G = b.get_G(4)
    b.set_G(4, 0.5 * G)
    assert np.array_equal(b.get_G(4), 0.5 * G)
    b.close()


def test_per_slice_calls_are_refused(dq):
    print("case: mixed-Nt batch per-slice calls", flush=True)
    from kapdqmc_amd import _lib as L
    ms = [tu(2, 4.0), tu(3, 4.0)]
    b = dq.DQMCBatch(ms, [1, 2], [0, 1], delay=8)
    out = np.zeros(64)
    p = out.ctypes.data_as(C.POINTER(C.c_double))
    for call, name in ((lambda: b.L.kdq_wrap(b.h, 1, 1), "kdq_wrap"),
                       (lambda: b.L.kdq_update_layer(b.h, 1, 0, 0), "kdq_update_layer"),
                       (lambda: b.L.kdq_measure(b.h, 1, p), "kdq_measure"),
                       (lambda: b.L.kdq_measure_noalloc(b.h, 1, p), "kdq_measure_noalloc")):
        assert call() == L.KDQ_EARG, name
        msg = b.L.kdq_last_error(b.h).decode()
        assert name in msg and "one nt" in msg, msg
    b.close()


def test_mixed_nt_tv_and_mixed_precision(dq, oracle_mod):
    print("case: mixed-Nt t-V batch vs oracle; mixed-precision t-U batch", flush=True)
    ms = [tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=th, Θquench=0.0,
                          Lattice="HoneyComb120", site=[L, L], BatchSize=5) for L, th in ((2, 1.0), (3, 1.5))]
    b = dq.DQMCBatch(ms, [3, 4], [0, 1], delay=8)
    os_ = []
    for c, m in enumerate(ms):
        o = oracle_mod.OracleTV(m, seed=3 + c, chain=c)
        o.initial_s()
        os_.append(o)
        b.set_fields(c, o.s)
    rb = b.phy_update(1, True)
    for c, o in enumerate(os_):
        ro = o.phy_update(1, True)
        assert np.array_equal(b.get_fields(c), o.s), c
        assert rel(b.get_G(c), o.G) < 1e-9, c
        assert np.allclose(rb[c], ro, rtol=1e-9, atol=1e-11), c
    b.close()
    mm = [tu(2, 4.0), tu(4, 4.0), tu(3, 4.0)]
    bm = dq.DQMCBatch(mm, [7, 8, 9], [0, 1, 2], delay=8, precision="mixed")
    bm.initial_s()
    rows = bm.phy_update(2, True)
    err = bm.precision_error()
    assert rows.shape == (3, 2, bm.nobs) and np.all(np.isfinite(rows))
    assert np.all(err > 0) and np.all(err < 1e-3), err
    bm.close()
