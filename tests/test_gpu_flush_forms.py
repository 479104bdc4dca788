"""The one-launch flush (a workgroup per 32-column strip forms its part of W (E_A - G[A,:])
and rewrites the strip) against the two-launch form (the product and the gathered columns
into scratch, then the rewrite) on identical batches. Both forms add the same rank terms in
the same order and put the sign on different, exactly scaled operands, so fields, G and the
measurement rows must agree bit for bit: t-U FP64 (batches below and above 8 chains, which
take the two launch grids of the strip kernel, block sizes 1..64), t-U mixed precision
(FP32 G), t-V and t-U-V sweeps (rank-2 elements, up to 64 ranks), and the swap-replica
four-matrix flushes of t-U, t-V and t-U-V (modes 0 and 1, gathered columns and rows from
other matrices). The oracle comparisons of the default (two-launch) form are the sweep and
swap-replica parity tests; "flush" 1 selects the one-launch form."""
import numpy as np
import pytest

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def tu(L, U=4.0, theta=1.0, bs=5):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


def tv(L=3, V=1.0):
    return tV_Hubbard_Para(Ht=1.0, Hv1=V, Hv2=V, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=5)


def tuv(L=3, U=4.0):
    return tUV_Hubbard_Para(Ht=1.0, Hu=U, Hv=0.5, Lattice="HoneyComb120", site=[L, L], Δt=0.1, Θ=1.0,
                            BatchSize=5)


def sweep(models, delay, form, precision="fp64", sweeps=2):
    from kapdqmc_amd import dqmc
    b = dqmc.DQMCBatch(models, list(range(11, 11 + len(models))), list(range(len(models))), delay=delay,
                       precision=precision)
    b.set_option("flush", form)
    b.set_option("prewrap", 0)   # the one-launch form has no prewrapped block 0
    b.initial_s()
    rows = b.phy_update(sweeps, True)
    out = ([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows)
    b.close()
    return out


def same(a, b, what):
    fa, ga, ra = a
    fb, gb, rb = b
    for c in range(len(fa)):
        assert np.array_equal(fa[c], fb[c]), (what, "fields", c)
        assert np.array_equal(ga[c], gb[c]), (what, "G", c, np.abs(ga[c] - gb[c]).max())
    assert np.array_equal(ra, rb), (what, "rows")


@pytest.mark.parametrize("delay", [1, 16, 32, 64])
def test_tu_forms_agree_few_chains(delay):
    print(f"case: flush forms, t-U 3 chains (L = 3/4/6), delay {delay}", flush=True)
    models = [tu(3), tu(4, U=2.0), tu(6, U=6.0)]
    same(sweep(models, delay, 1), sweep(models, delay, 0), f"k={delay}")


def test_tu_forms_agree_many_chains():
    # 10 chains: the XCD-mapped grid (every strip of chain c on XCD c % 8), ragged last group
    print("case: flush forms, t-U 10 chains of L = 4/6, delay 32", flush=True)
    models = [tu(4 if c % 2 else 6, U=2.0 + c % 3) for c in range(10)]
    same(sweep(models, 32, 0), sweep(models, 32, 1), "10 chains")


def test_tu_forms_agree_mixed_precision():
    print("case: flush forms, t-U mixed precision", flush=True)
    models = [tu(4), tu(6, U=2.0)]
    same(sweep(models, 16, 0, precision="mixed"), sweep(models, 16, 1, precision="mixed"), "mixed")


@pytest.mark.parametrize("name", ["tV", "tUV"])
def test_r2_sweep_forms_agree(name):
    print(f"case: flush forms, {name} sweep", flush=True)
    models = [tv(3), tv(4, V=0.5)] if name == "tV" else [tuv(3), tuv(4, U=-4.0)]
    same(sweep(models, 16, 0), sweep(models, 16, 1), name)


def regions(m):
    iA = area_index(m.Lattice, m.site, ([1, 1], [2, m.site[1]]))
    iB = area_index(m.Lattice, m.site, ([1, 1], [1, m.site[1]]))
    return np.asarray(iA, dtype=np.int64), np.asarray(iB, dtype=np.int64)


@pytest.mark.parametrize("name,mk,delay", [("tU", lambda: tu(3, theta=1.0), 8), ("tV", tv, 8),
                                           ("tUV", tuv, 8)])
def test_scee_forms_agree(name, mk, delay):
    from kapdqmc_amd import dqmc
    print(f"case: flush forms, swap-replica {name}", flush=True)
    m = mk()
    iA, iB = regions(m)
    res = []
    for form in (0, 1):
        b = dqmc.SCEEBatch([m, m], iA, iB, 0.5, 2, [5, 6], [0, 1], delay=delay)
        b.set_option("flush", form)
        for c in range(2):
            b.set_fields(c, 0, dqmc.Initial_s(m, 20 + c))
            b.set_fields(c, 1, dqmc.Initial_s(m, 40 + c))
        O = b.sweeps(2)
        res.append((O, [b.get_fields(c, r) for c in range(2) for r in range(2)],
                    [g for c in range(2) for r in range(2) for g in b.G4(c, r)]))
        b.close()
    (O0, f0, g0), (O1, f1, g1) = res
    assert np.array_equal(O0, O1)
    for x, y in zip(f0, f1):
        assert np.array_equal(x, y)
    for x, y in zip(g0, g1):
        assert np.array_equal(x, y), np.abs(x - y).max()
