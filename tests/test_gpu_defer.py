"""Deferred flush of t-U slices (G0 and low-rank factors, one GEMM per slice) against the
per-block flush on the same batch: identical Philox draws, so the fields must agree bit for
bit and G / the measurement rows to rounding (1e-10 relative), for block sizes 1..64 and a
batch of uneven lattices; the oracle comparison of the default (deferred) path is in
test_gpu_tu.py."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

pytestmark = pytest.mark.gpu


def tu(L, U=4.0, theta=1.0, bs=5):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run(dqmc, models, delay, defer, sweeps=2, graph=1):
    b = dqmc.DQMCBatch(models, list(range(7, 7 + len(models))), list(range(len(models))), delay=delay)
    b.set_option("defer", defer)
    b.set_option("graph", graph)
    b.initial_s()
    rows = b.phy_update(sweeps, True)
    out = ([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows)
    b.close()
    return out


@pytest.mark.parametrize("delay", [1, 5, 16, 32, 64])
def test_defer_equals_per_block_flush(delay):
    from kapdqmc_amd import dqmc
    print(f"case: deferred vs per-block flush, delay={delay}", flush=True)
    models = [tu(3), tu(4, U=2.0), tu(6, U=6.0)]
    f1, g1, r1 = run(dqmc, models, delay, 1)
    f0, g0, r0 = run(dqmc, models, delay, 0)
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c]), c
        assert rel(g1[c], g0[c]) < 1e-10, (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-9, atol=1e-11)


def test_defer_graph_equals_eager():
    from kapdqmc_amd import dqmc
    print("case: deferred flush, graph vs eager", flush=True)
    models = [tu(4), tu(3)]
    f1, g1, r1 = run(dqmc, models, 16, 1, sweeps=3, graph=1)
    f0, g0, r0 = run(dqmc, models, 16, 1, sweeps=3, graph=0)
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c])
        assert np.array_equal(g1[c], g0[c])
    assert np.array_equal(r1, r0)
