"""Sweeps replayed from a captured HIP graph give exactly what eager launches give.

Every replay is checked: several sweeps in one call (host reads of the measurement rows
between replays), then more calls with eager work (G read-back, chain re-initialisation)
between them; t-U, t-V and t-U-V, plus a batch of uneven chains. Bit for bit in the
one-stream block order; with the pipelined t-U blocks (graphs only) the fields bit for bit
and G / the rows to rounding."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para
from kapdqmc_amd.rng import initial_fields

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dq():
    from kapdqmc_amd import dqmc
    return dqmc


def tu(L=4, U=4.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[L, L], BatchSize=5)


def tv(L=3):
    return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=5)


def tuv(L=3):
    return tUV_Hubbard_Para(Ht=1.0, Hu=4.0, Hv=0.5, Lattice="HoneyComb120", site=[L, L], Δt=0.1, Θ=1.0,
                            BatchSize=5)


def run(dq, models, graph: bool, delay=8, pipe=0):
    b = dq.DQMCBatch(models, [9 + c for c in range(len(models))], list(range(len(models))), delay=delay)
    b.set_option("graph", int(graph))
    b.set_option("pipe", pipe)
    b.initial_s()
    out = [b.phy_update(3, True)]
    for _ in range(2):
        _ = b.get_G(0)                 # eager work between replays
        b.init_chain()
        out.append(b.phy_update(1, True))
    fields = [b.get_fields(c) for c in range(len(models))]
    Gs = [b.get_G(c) for c in range(len(models))]
    return np.concatenate(out, axis=1), fields, Gs


@pytest.mark.parametrize("kind", ["tU", "tV", "tUV", "tU_batch"])
def test_graph_replays_equal_eager(dq, kind):
    print(f"case: graph vs eager {kind}", flush=True)
    models = {"tU": [tu()], "tV": [tv()], "tUV": [tuv()], "tU_batch": [tu(3), tu(4, 3.0), tu(6, 5.0)]}[kind]
    r1, f1, g1 = run(dq, models, True)
    r0, f0, g0 = run(dq, models, False)
    assert np.array_equal(r1, r0)
    for a, b in zip(f1, f0):
        assert np.array_equal(a, b)
    for a, b in zip(g1, g0):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["tU", "tU_batch"])
def test_pipelined_graph_replays_against_eager(dq, kind):
    print(f"case: pipelined graph vs eager {kind}", flush=True)
    models = {"tU": [tu()], "tU_batch": [tu(3), tu(4, 3.0), tu(6, 5.0)]}[kind]
    r1, f1, g1 = run(dq, models, True, pipe=1)
    r0, f0, g0 = run(dq, models, False, pipe=1)
    assert np.allclose(r1, r0, rtol=1e-9, atol=1e-11)
    for a, b in zip(f1, f0):
        assert np.array_equal(a, b)
    for a, b in zip(g1, g0):
        assert np.abs(a - b).max() <= 1e-10 * np.abs(b).max()
