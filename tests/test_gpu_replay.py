"""The replay-buffer random stream (kdq_set_replay) on the GPU against the oracle's replay,
through the C ABI: arbitrary draws (not from Philox) drive the same decisions on both sides;
replaying a chain's own Philox draws reproduces its Philox run; calls beyond the replayed
sweeps and invalid choices are refused. Bars: fields bit-exact with equal near-tie counts, G
within 1e-10 relative, rows within 1e-9 relative."""
import ctypes as C

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def tu():
    return tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[3, 3], BatchSize=5)


def tv():
    return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[3, 3], BatchSize=5)


def tuv():
    return tUV_Hubbard_Para(Ht=1.0, Hu=4.0, Hv=0.5, Lattice="HoneyComb120", site=[3, 3], Δt=0.1, Θ=1.0, BatchSize=5)


MODELS = [("tU", tu, "OracleTU", 16), ("tV", tv, "OracleTV", 8), ("tUV", tuv, "OracleTUV", 8)]


def random_draws(m, nsweeps, seed):
    from kapdqmc_amd.dqmc import elements_per_slice
    rng = np.random.default_rng(seed)
    shape = (nsweeps, 2, m.Nt, elements_per_slice(m))
    return rng.integers(0, 3, size=shape).astype(np.uint8), rng.random(shape)


@pytest.mark.parametrize("name,mk,cls,delay", MODELS)
def test_replayed_draws_match_the_oracle(oracle_mod, name, mk, cls, delay):
    """Chain 0 replays arbitrary draws after one Philox sweep (so the captured sweep graph
    reads the replay table), chain 1 keeps its Philox stream; both against the oracle."""
    print(f"case: replay {name}", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch
    m = mk()
    seeds = [21, 22]
    os_ = []
    for c in range(2):
        o = getattr(oracle_mod, cls)(m, seed=seeds[c], chain=c)
        o.initial_s()
        os_.append(o)
    b = DQMCBatch([m, m], seeds, [0, 1], delay=delay)
    for c in range(2):
        b.set_fields(c, os_[c].s)
    b.phy_update(1, True)
    for o in os_:
        o.phy_update(1, True)
    ch, u = random_draws(m, 2, 5)
    b.set_replay(0, ch, u)
    os_[0].set_replay(ch, u)
    rb = b.phy_update(2, True)
    for c, o in enumerate(os_):
        ro = o.phy_update(2, True)
        assert np.array_equal(b.get_fields(c), o.s), f"chain {c}"
        assert b.near_ties()[c] == o.near_ties
        assert rel(b.get_G(c), o.G) < 1e-10
        assert np.allclose(rb[c], ro, rtol=1e-9, atol=1e-11)
    b.close()


def test_replaying_own_philox_draws_reproduces_the_run(oracle_mod):
    print("case: replay of Philox draws, 2 chains", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch, elements_per_slice
    m = tu()
    E, nt = elements_per_slice(m), m.Nt
    ch = np.zeros((2, 2, nt, E), dtype=np.uint8)
    u = np.zeros((2, 2, nt, E))
    for s in range(2):
        for d in (0, 1):
            for lt in range(1, nt + 1):
                for e in range(E):
                    ch[s, d, lt - 1, e], u[s, d, lt - 1, e] = oracle_mod.draw(31, 0, 0, s, lt | (d << 31), e)
    from kapdqmc_amd.dqmc import Initial_s
    s0 = Initial_s(m, 31, 0)
    a = DQMCBatch([m], [31], [0], delay=32)
    a.set_fields(0, s0)
    ra = a.phy_update(2, True)
    b = DQMCBatch([m], [77], [5], delay=32)
    b.set_fields(0, s0)
    b.set_replay(0, ch, u)
    rb = b.phy_update(2, True)
    assert np.array_equal(a.get_fields(0), b.get_fields(0))
    assert rel(a.get_G(0), b.get_G(0)) < 1e-12
    assert np.allclose(ra, rb, rtol=1e-12, atol=1e-14)
    # back to Philox: nsweeps = 0
    b.set_replay(0, np.zeros(0, np.uint8), np.zeros(0))
    b.phy_update(1, False)
    a.close()
    b.close()


def test_replay_limits_are_refused():
    print("case: replay limits", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch, Initial_s
    m = tu()
    b = DQMCBatch([m], [3], [0], delay=16)
    b.set_fields(0, Initial_s(m, 3, 0))
    ch, u = random_draws(m, 1, 9)
    b.set_replay(0, ch, u)
    rc = b.L.kdq_phy_update(b.h, 2, 0, None)           # two sweeps, one replayed
    assert rc == -1 and b"cover sweeps" in b.L.kdq_last_error(b.h)
    bad = ch.copy()
    bad[0, 0, 0, 0] = 3
    rc = b.L.kdq_set_replay(b.h, 0, bad.ctypes.data_as(C.POINTER(C.c_uint8)),
                            u.ctypes.data_as(C.POINTER(C.c_double)), 1)
    assert rc == -1
    b.phy_update(1, False)                             # the one replayed sweep still runs
    b.close()
