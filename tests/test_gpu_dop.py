"""GPU parity of the t-U disorder operator (ctrl_SCDOPicr, src/tU/disorder_operate.jl) against
the FP64 oracle, through the C ABI.

Bars: the field bit-exact (identical Philox stream 1), G(τ), G(Θ), G(τ,Θ), G(Θ,τ) within
1e-10 relative, |det X|² and the per-sweep O within 1e-9 relative; block sizes 1..16, an
empty region B, a batch of two lattices, and the reference's CSV row."""
import os

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


def check(b, c, o):
    assert np.array_equal(b.get_fields(c), o.fields(0))
    for name, x, y in zip(("Gt", "G0", "Gt0", "G0t"), b.G4(c), o.G4(0)):
        assert rel(x, y) < G_TOL, (name, rel(x, y))
    d = b.detg()[c]
    assert abs(d[0] / o.detg(0) - 1) < 1e-9 and abs(d[1] / o.detg(1) - 1) < 1e-9


@pytest.mark.parametrize("delay", [1, 8, 16])
def test_dop_sweeps_parity(dq, oracle_mod, delay):
    print(f"case: disorder operator sweeps L=3 delay={delay}", flush=True)
    m = model()
    iA, iB = regions(m)
    alpha, seed = 0.8, 5
    s = initial_fields(seed, 0, m.Ns, m.Nt)
    o = oracle_mod.OracleDOP(m, alpha, iA.astype(np.int32), iB.astype(np.int32), 0.5, 2, s, seed=seed)
    b = dq.DOPBatch([m], alpha, iA, iB, 0.5, 2, [seed], [0], delay=delay)
    b.set_fields(0, s)
    Oo = o.sweeps(2)
    Ob = b.sweeps(2)[0]
    check(b, 0, o)
    assert np.allclose(Ob, Oo, rtol=1e-9, atol=0)
    assert not np.array_equal(b.get_fields(0), s)
    b.close()


def test_dop_empty_b_and_two_chains(dq, oracle_mod):
    print("case: disorder operator, B empty, 2 chains", flush=True)
    ms = [model(3), model(4, U=3.0)]
    alpha, lam, nlam = 2.0, 0.3, 3
    regs = [regions(m, with_b=False) for m in ms]
    seeds = [11, 12]
    b = dq.DOPBatch(ms, alpha, [r[0] for r in regs], [r[1] for r in regs], lam, nlam, seeds, [0, 1], delay=8)
    os_ = []
    for c, (m, (iA, iB), sd) in enumerate(zip(ms, regs, seeds)):
        s = initial_fields(sd, c, m.Ns, m.Nt)
        b.set_fields(c, s)
        os_.append(oracle_mod.OracleDOP(m, alpha, iA.astype(np.int32), iB.astype(np.int32), lam, nlam, s,
                                        seed=sd, chain=c))
    Ob = b.sweeps(2)
    for c, o in enumerate(os_):
        Oo = o.sweeps(2)
        check(b, c, o)
        assert np.allclose(Ob[c], Oo, rtol=1e-9, atol=0)
        assert b.detg()[c][1] == 1.0
    b.close()


def test_ctrl_SCDOPicr_csv(dq, oracle_mod, tmp_path):
    print("case: ctrl_SCDOPicr CSV row", flush=True)
    m = model()
    iA, iB = regions(m)
    s = initial_fields(2, 0, m.Ns, m.Nt)
    o = oracle_mod.OracleDOP(m, 0.5, iA.astype(np.int32), iB.astype(np.int32), 0.5, 2, s.copy(), seed=2)
    Oo = o.sweeps(2)
    out = dq.ctrl_SCDOPicr(str(tmp_path), m, 0.5, iA, iB, 2, 0.5, 2, s, True, seed=2)
    assert out is s and np.array_equal(s, o.fields(0))
    fn = f"tUSCDOPHC120_t{m.Ht}U{m.Hu1}size{m.site}Δt{m.Δt}Θ{m.Θrelax}N2BS{m.BatchSize}.csv"
    row = np.array([float(x) for x in open(os.path.join(tmp_path, fn)).read().strip().split(",")])
    assert row[0] == 0.5 and np.allclose(row[1:], Oo, rtol=1e-9)
