"""GPU parity of the t-U projector-DQMC path against the FP64 oracle, through the C ABI.

Bars: fields (accept/reject decisions with identical Philox draws) bit-exact, near-tie counts
(|u − p| < 1e-12) equal; G within 1e-10 relative (max |ΔG| / max |G|); measurement rows
within 1e-9 relative.
"""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

pytestmark = pytest.mark.gpu

G_TOL = 1e-10


def model(L=3, U=4.0, theta=1.0, bs=5, dt=0.1, lattice="HoneyComb120"):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=dt, Θrelax=theta, Θquench=0.0, Lattice=lattice,
                           site=[L, L], BatchSize=bs)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def dq():
    from kapdqmc_amd import dqmc
    return dqmc


def make_pair(dq, oracle_mod, m, seed=11, chain=0, delay=32):
    o = oracle_mod.OracleTU(m, seed=seed, chain=chain)
    o.initial_s()
    b = dq.DQMCBatch([m], [seed], [chain], delay=delay)
    b.set_fields(0, o.s)
    return o, b


def test_init_chain_get_G(dq, oracle_mod):
    print("case: init_chain L=3", flush=True)
    m = model()
    o, b = make_pair(dq, oracle_mod, m)
    o.init_chain()
    b.init_chain()
    G = b.get_G(0)
    assert rel(G, o.G) < G_TOL
    assert np.abs(G @ G - G).max() < 1e-10
    assert abs(np.trace(G).real - (m.Ns - m.Ns // 2)) < 1e-10


@pytest.mark.parametrize("forward", [True, False])
def test_wrap(dq, oracle_mod, forward):
    print(f"case: wrap forward={forward}", flush=True)
    m = model(L=6)
    o, b = make_pair(dq, oracle_mod, m)
    rs = np.random.default_rng(3)
    G = rs.standard_normal((m.Ns, m.Ns)) + 1j * rs.standard_normal((m.Ns, m.Ns))
    b.set_G(0, G)
    lt = 7
    sgn = 1 if forward else -1
    D = np.exp(1j * sgn * m.α[lt - 1] * m.η[o.s[:, lt - 1].astype(int) - 1])
    ref = o.wrap(G, D, forward, "B")
    b.wrap(lt, forward)
    assert rel(b.get_G(0), ref) < 1e-12


@pytest.mark.parametrize("delay", [1, 8, 32, 64])
def test_update_layer(dq, oracle_mod, delay):
    print(f"case: update_layer delay={delay}", flush=True)
    m = model(L=6)
    o, b = make_pair(dq, oracle_mod, m, delay=delay)
    o.init_chain()
    b.init_chain()
    for lt, d in [(1, 0), (2, 0)]:
        D = np.exp(1j * m.α[lt - 1] * m.η[o.s[:, lt - 1].astype(int) - 1])
        o.G = o.wrap(o.G, D, True, "B")
        b.wrap(lt, True)
        o.update_layer(lt, d)
        b.update_layer(lt, d, 0)
        assert np.array_equal(b.get_fields(0), o.s), f"decisions differ at lt={lt}"
        assert rel(b.get_G(0), o.G) < G_TOL
    cnt = b.counters()[0]
    assert cnt[0] == 2 * m.Ns and 0 < cnt[1] <= cnt[0]


def test_measure(dq, oracle_mod):
    print("case: measure", flush=True)
    m = model(L=3, theta=1.0)
    o, b = make_pair(dq, oracle_mod, m)
    o.init_chain()
    b.init_chain()
    for lt in (3, 10, 17):
        got = b.measure(lt)[0]
        ref = o.measure(lt, o.G)
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-11), (lt, got, ref)


@pytest.mark.parametrize("L,delay", [(3, 32), (3, 1), (6, 16)])
def test_phy_update_parity(dq, oracle_mod, L, delay):
    print(f"case: phy_update L={L} delay={delay}", flush=True)
    m = model(L=L, theta=1.0, bs=5)
    o, b = make_pair(dq, oracle_mod, m, seed=2024 + L, delay=delay)
    ro = o.phy_update(2, True)
    rb = b.phy_update(2, True)[0]
    assert np.array_equal(b.get_fields(0), o.s)
    assert b.near_ties()[0] == o.near_ties
    assert rel(b.get_G(0), o.G) < G_TOL
    assert np.allclose(rb, ro, rtol=1e-9, atol=1e-11)


def test_batched_uneven_chains(dq, oracle_mod):
    print("case: batched chains L=3,6,4 U=2,4,6", flush=True)
    ms = [model(L=3, U=2.0), model(L=6, U=4.0), model(L=4, U=6.0)]
    seeds = [5, 6, 7]
    os_ = []
    for c, m in enumerate(ms):
        o = oracle_mod.OracleTU(m, seed=seeds[c], chain=c)
        o.initial_s()
        os_.append(o)
    b = dq.DQMCBatch(ms, seeds, [0, 1, 2], delay=32)
    for c in range(3):
        b.set_fields(c, os_[c].s)
    rb = b.phy_update(1, True)
    for c, o in enumerate(os_):
        ro = o.phy_update(1, True)
        assert np.array_equal(b.get_fields(c), o.s)
        assert b.near_ties()[c] == o.near_ties
        assert rel(b.get_G(c), o.G) < G_TOL
        assert np.allclose(rb[c], ro, rtol=1e-9, atol=1e-11)


def test_full_size_projector_properties(dq):
    """N = 288 (12 x 12, Nt = 200): after a sweep G stays a rank-(N - Np) projector and
    agrees with G rebuilt from scratch out of the stabilised chain."""
    print("case: full size L=12", flush=True)
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=10.0, Θquench=0.0,
                        Lattice="HoneyComb120", site=[12, 12], BatchSize=10)
    from kapdqmc_amd.rng import initial_fields
    b = dq.DQMCBatch([m], [1], [0], delay=32)
    b.set_fields(0, initial_fields(1, 0, m.Ns, m.Nt))
    rows = b.phy_update(1, True)[0]
    G = b.get_G(0)
    n, npart = m.Ns, m.Ns // 2
    assert np.abs(G @ G - G).max() < 1e-9
    assert abs(np.trace(G).real - (n - npart)) < 1e-9
    b.init_chain()
    assert rel(b.get_G(0), G) < 1e-9
    assert np.all(np.isfinite(rows))
    cnt = b.counters()[0]
    assert cnt[0] == 2 * m.Nt * m.Ns
    assert 0.05 < cnt[1] / cnt[0] < 0.95
