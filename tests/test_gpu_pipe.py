"""Pipelined t-U blocks (the decisions of block b+1 on a second stream beside block b's
flush, reading their submatrix of the updated G from the G before that flush and block b's
accepted sites and W) against the one-stream order on the same batch: identical Philox
draws, so the fields must agree bit for bit and G / the measurement rows to rounding, for
block sizes 1..64, FP64 and mixed precision, a batch of uneven lattices, and graph replays
(where the pipelining applies) against host launches (which keep the one-stream order). The
oracle comparison of the default (pipelined) path is in test_gpu_tu.py."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

pytestmark = pytest.mark.gpu


def tu(L, U=4.0, theta=1.0, bs=5):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run(dqmc, models, delay, pipe, sweeps=2, graph=1, precision="fp64"):
    b = dqmc.DQMCBatch(models, list(range(7, 7 + len(models))), list(range(len(models))), delay=delay,
                       precision=precision)
    b.set_option("pipe", pipe)
    b.set_option("graph", graph)
    b.initial_s()
    rows = b.phy_update(sweeps, True)
    out = ([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows)
    b.close()
    return out


@pytest.mark.parametrize("delay", [1, 5, 16, 32, 64])
def test_pipe_equals_one_stream(delay):
    from kapdqmc_amd import dqmc
    print(f"case: pipelined vs one-stream blocks, delay={delay}", flush=True)
    models = [tu(3), tu(4, U=2.0), tu(6, U=6.0)]
    f1, g1, r1 = run(dqmc, models, delay, 1)
    f0, g0, r0 = run(dqmc, models, delay, 0)
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c]), c
        assert rel(g1[c], g0[c]) < 1e-10, (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-9, atol=1e-11)


def test_pipe_mixed_precision():
    from kapdqmc_amd import dqmc
    print("case: pipelined vs one-stream blocks, mixed precision", flush=True)
    models = [tu(4), tu(6, U=2.0)]
    f1, g1, r1 = run(dqmc, models, 16, 1, precision="mixed")
    f0, g0, r0 = run(dqmc, models, 16, 0, precision="mixed")
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c]), c
        assert rel(g1[c], g0[c]) < 1e-5, (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-4, atol=1e-6)


def test_pipe_graph_against_host_launches():
    # host-launched sweeps always take the one-stream order
    from kapdqmc_amd import dqmc
    print("case: pipelined graph vs host launches", flush=True)
    models = [tu(4), tu(3)]
    f1, g1, r1 = run(dqmc, models, 16, 1, sweeps=3, graph=1)
    f0, g0, r0 = run(dqmc, models, 16, 1, sweeps=3, graph=0)
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c])
        assert rel(g1[c], g0[c]) < 1e-10, (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-9, atol=1e-11)


def test_pipe_graph_replays_repeat():
    # two batches with the same seeds, captured and replayed: bit-identical
    from kapdqmc_amd import dqmc
    print("case: pipelined graph replays, two identical batches", flush=True)
    models = [tu(4), tu(6, U=2.0)]
    f1, g1, r1 = run(dqmc, models, 32, 1, sweeps=3)
    f0, g0, r0 = run(dqmc, models, 32, 1, sweeps=3)
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c])
        assert np.array_equal(g1[c], g0[c])
    assert np.array_equal(r1, r0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_chain_products_aside_equal_inline(precision):
    # the chain products of each slice issued on a third stream as soon as the slice's fields
    # are final (option "aside") are the node's own products in the same order: bit-identical
    from kapdqmc_amd import dqmc
    print(f"case: chain products beside the sweep vs at the node, {precision}", flush=True)
    models = [tu(4), tu(6, U=2.0), tu(3, U=6.0)]
    res = []
    for aside in (1, 0):
        b = dqmc.DQMCBatch(models, [5, 6, 7], [0, 1, 2], delay=16, precision=precision)
        b.set_option("aside", aside)
        b.initial_s()
        rows = b.phy_update(3, True)
        res.append(([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows))
        b.close()
    (f1, g1, r1), (f0, g0, r0) = res
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c])
        assert np.array_equal(g1[c], g0[c])
    assert np.array_equal(r1, r0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_node_second_pass_beside_get_G(precision):
    # option "qsplit": the node's G from the chain after the first Cholesky-QR pass (same span,
    # well conditioned) while the second pass runs on another stream; G agrees to rounding
    from kapdqmc_amd import dqmc
    print(f"case: second orthonormalisation pass beside get_G, {precision}", flush=True)
    models = [tu(4), tu(6, U=2.0), tu(3, U=6.0, bs=3)]
    res = []
    for qs in (1, 0):
        b = dqmc.DQMCBatch(models, [5, 6, 7], [0, 1, 2], delay=16, precision=precision)
        b.set_option("qsplit", qs)
        b.initial_s()
        rows = b.phy_update(3, True)
        res.append(([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows))
        b.close()
    (f1, g1, r1), (f0, g0, r0) = res
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c])
        assert rel(g1[c], g0[c]) < (1e-11 if precision == "fp64" else 1e-5), (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-9, atol=1e-11) if precision == "fp64" else np.allclose(r1, r0, rtol=1e-4, atol=1e-6)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_measurements_beside_the_sweep(precision):
    # option "maside": a measurement's products and tiles on another stream; same kernels on
    # the same operands, so the rows and everything else are bit-identical
    from kapdqmc_amd import dqmc
    print(f"case: measurements beside the sweep, {precision}", flush=True)
    models = [tu(4, theta=2.0), tu(6, U=2.0, theta=2.0), tu(3, U=6.0, theta=2.0, bs=3)]
    res = []
    for ms in (1, 0):
        b = dqmc.DQMCBatch(models, [5, 6, 7], [0, 1, 2], delay=16, precision=precision)
        b.set_option("maside", ms)
        b.initial_s()
        rows = b.phy_update(3, True)
        res.append(([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows))
        b.close()
    (f1, g1, r1), (f0, g0, r0) = res
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c])
        assert np.array_equal(g1[c], g0[c])
    assert np.array_equal(r1, r0)


@pytest.mark.parametrize("delay", [16, 32])
def test_first_block_beside_the_wrap(delay):
    # option "prewrap": each layer's block 0 decided between the wrap's two products from the
    # first product and the dense kinetic factor, its accepted fields written back by its flush;
    # the block's G[S,S] rounds differently, so G agrees to rounding and fields exactly
    from kapdqmc_amd import dqmc
    print(f"case: block 0 beside the wrap's second product, delay {delay}", flush=True)
    models = [tu(4, theta=2.0), tu(6, U=2.0, theta=2.0), tu(3, U=6.0, theta=2.0, bs=3)]
    res = []
    for pw in (1, 0):
        b = dqmc.DQMCBatch(models, [5, 6, 7], [0, 1, 2], delay=delay)
        b.set_option("prewrap", pw)
        b.initial_s()
        rows = b.phy_update(3, True)
        res.append(([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows,
                    b.near_ties()))
        b.close()
    (f1, g1, r1, n1), (f0, g0, r0, n0) = res
    assert np.array_equal(n1, n0)
    for c in range(len(models)):
        assert np.array_equal(f1[c], f0[c]), c
        assert rel(g1[c], g0[c]) < 1e-10, (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-9, atol=1e-11)
