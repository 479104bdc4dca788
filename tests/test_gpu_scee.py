"""GPU parity of the t-U swap-replica Rényi-2 sweep (ctrl_SCEEicr) against the FP64 oracle,
through the C ABI.

Bars: both replicas' fields bit-exact (identical Philox streams 1 and 2), G(τ), G(Θ),
G(τ,Θ), G(Θ,τ) within 1e-10 relative, |det g|² and the per-sweep O within 1e-9 relative."""
import numpy as np
import pytest

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para
from kapdqmc_amd.rng import initial_fields

pytestmark = pytest.mark.gpu

G_TOL = 1e-10


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def dq():
    from kapdqmc_amd import dqmc
    return dqmc


def model(L=3, theta=1.0, bs=5, U=4.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


def regions(m, with_b=True):
    iA = area_index(m.Lattice, m.site, ([1, 1], [2, m.site[1]]))
    iB = area_index(m.Lattice, m.site, ([1, 1], [1, m.site[1]])) if with_b else np.zeros(0, dtype=np.int64)
    return np.asarray(iA, dtype=np.int64), np.asarray(iB, dtype=np.int64)


def pair(dq, oracle_mod, m, iA, iB, seed, delay, lam=0.5, nlam=2):
    s1, s2 = initial_fields(seed, 0, m.Ns, m.Nt), initial_fields(seed + 100, 0, m.Ns, m.Nt)
    o = oracle_mod.OracleSCEE(m, iA.astype(np.int32), iB.astype(np.int32), lam, nlam, s1, s2, seed=seed)
    b = dq.SCEEBatch([m], iA, iB, lam, nlam, [seed], [0], delay=delay)
    b.set_fields(0, 0, s1)
    b.set_fields(0, 1, s2)
    return o, b


def test_scee_setup_G4_and_grover(dq, oracle_mod):
    print("case: SCEE set-up L=3", flush=True)
    m = model()
    iA, iB = regions(m)
    o, b = pair(dq, oracle_mod, m, iA, iB, seed=3, delay=8)
    o.init()
    b.init()
    for r in (0, 1):
        for name, x, y in zip(("Gt", "G0", "Gt0", "G0t"), b.G4(0, r), o.G4(r)):
            assert rel(x, y) < G_TOL, (r, name, rel(x, y))
    d = b.detg()[0]
    assert abs(d[0] / o.detg(0) - 1) < 1e-9 and abs(d[1] / o.detg(1) - 1) < 1e-9


@pytest.mark.parametrize("delay", [1, 8, 16])
def test_scee_sweeps_parity(dq, oracle_mod, delay):
    print(f"case: SCEE sweeps L=3 delay={delay}", flush=True)
    m = model()
    iA, iB = regions(m)
    o, b = pair(dq, oracle_mod, m, iA, iB, seed=5, delay=delay)
    Oo = o.sweeps(2)
    Ob = b.sweeps(2)[0]
    for r in (0, 1):
        assert np.array_equal(b.get_fields(0, r), o.fields(r)), r
        for name, x, y in zip(("Gt", "G0", "Gt0", "G0t"), b.G4(0, r), o.G4(r)):
            assert rel(x, y) < G_TOL, (r, name, rel(x, y))
    assert np.allclose(Ob, Oo, rtol=1e-9, atol=0)
    d = b.detg()[0]
    assert abs(d[0] / o.detg(0) - 1) < 1e-9 and abs(d[1] / o.detg(1) - 1) < 1e-9


def test_scee_empty_region_b_and_two_chains(dq, oracle_mod):
    print("case: SCEE B empty, 2 chains", flush=True)
    ms = [model(U=4.0), model(U=3.0)]
    iA, iB = regions(ms[0], with_b=False)
    os_, s = [], []
    for c, m in enumerate(ms):
        s1, s2 = initial_fields(11 + c, c, m.Ns, m.Nt), initial_fields(21 + c, c, m.Ns, m.Nt)
        s.append((s1, s2))
        os_.append(oracle_mod.OracleSCEE(m, iA.astype(np.int32), iB.astype(np.int32), 0.3, 3, s1, s2,
                                         seed=11, chain=c))
    b = dq.SCEEBatch(ms, iA, iB, 0.3, 3, [11, 11], [0, 1], delay=8)
    for c in range(2):
        b.set_fields(c, 0, s[c][0])
        b.set_fields(c, 1, s[c][1])
    Ob = b.sweeps(1)
    for c, o in enumerate(os_):
        Oo = o.sweeps(1)
        assert np.allclose(Ob[c], Oo, rtol=1e-9, atol=0)
        for r in (0, 1):
            assert np.array_equal(b.get_fields(c, r), o.fields(r))
            assert rel(b.G4(c, r)[0], o.G4(r)[0]) < G_TOL
        assert b.detg()[c][1] == 1.0
