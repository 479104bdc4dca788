"""GPU time-displaced Green's functions (kdq_G4, replacing G4 of src/tU/GreenMatrix.jl:126-226)
against the oracle's restatement, through the C ABI: G(τ1), G(τ2), G(τ1,τ2), G(τ2,τ1) within
1e-10 relative of the largest entry, for τ1 > τ2, τ1 < τ2, τ1 == τ2, the end slices, ranges
that the BatchSize blocks do not divide, the centred node layout, and a batch of uneven
lattices; the call leaves the sweep state alone (G and the next sweep bit for bit)."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para
from kapdqmc_amd.rng import initial_fields

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def model(L=3, theta=1.0, bs=5, U=4.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


@pytest.fixture(scope="module")
def dq():
    from kapdqmc_amd import dqmc
    return dqmc


@pytest.mark.parametrize("bs", [5, 3])
def test_g4_parity(dq, oracle_mod, bs):
    print(f"case: kdq_G4 L=3 BatchSize={bs}", flush=True)
    m = model(bs=bs)   # Nt = 20; BatchSize 3 gives the centred node layout
    b = dq.DQMCBatch([m], [3], [0], delay=8)
    b.initial_s()
    b.phy_update(1, False)
    s = b.get_fields(0)
    for t1, t2 in [(13, 2), (20, 0), (7, 6), (3, 17), (9, 9), (0, 20), (20, 20), (11, 10)]:
        got = b.G4(0, t1, t2)
        want = oracle_mod.tu_G4(m, s, t1, t2)
        for name, x, y in zip(("G1", "G2", "G12", "G21"), got, want):
            assert rel(x, y) < 1e-10, (t1, t2, name, rel(x, y))
    b.close()


def test_g4_batch_and_state(dq, oracle_mod):
    print("case: kdq_G4 uneven batch, sweep state untouched", flush=True)
    ms = [model(3), model(4, U=3.0)]

    def make():
        b = dq.DQMCBatch(ms, [4, 5], [0, 1], delay=8)
        for c, m in enumerate(ms):
            b.set_fields(c, initial_fields(4 + c, c, m.Ns, m.Nt))
        b.phy_update(1, False)
        return b

    b, b0 = make(), make()
    G_before = [b.get_G(c) for c in range(2)]
    for c, m in enumerate(ms):
        got = b.G4(c, 15, 4)
        want = oracle_mod.tu_G4(m, b.get_fields(c), 15, 4)
        for name, x, y in zip(("G1", "G2", "G12", "G21"), got, want):
            assert rel(x, y) < 1e-10, (c, name, rel(x, y))
    for c in range(2):
        assert np.array_equal(b.get_G(c), G_before[c])
    r1 = b.phy_update(1, True)
    r0 = b0.phy_update(1, True)
    assert np.array_equal(r1, r0)
    for c in range(2):
        assert np.array_equal(b.get_fields(c), b0.get_fields(c))
        assert np.array_equal(b.get_G(c), b0.get_G(c))
    b.close()
    b0.close()
