"""Kinetic products through cell momenta (block-circulant eK, eK⁻¹: length-L DFTs over the
periodic L x L cells and a 2 x 2 block per momentum) against the dense GEMM path
(kdq_set_option "circ" 0) and the oracle.

Bars: one wrap within 1e-12 relative of the oracle's dense product for every cell edge with a
kernel instantiation and both directions; sweeps with the two paths: fields bit-exact, G
within 1e-11 relative, rows within 1e-9 (the paths round differently; the oracle parity bars);
sweeps of t-U on every lattice, complex hopping (π flux on SQUARE90); t-V and t-U-V, a batch
whose eK is not block-circulant and a cell edge without an instantiation take the dense path
(bit-identical results with the option on and off). The sweep measurements as two chained
cell-momentum passes (option "mchain") against the slot-matrix products, to the same bars, and
a model whose HalfeK is not block-circulant keeps the slot products (bit-identical)."""
import copy

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def tu(L, U=4.0, lattice="HoneyComb120", flux=0.0, theta=1.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=theta, Θquench=0.0, Lattice=lattice,
                           site=[L, L], BatchSize=5, flux=flux)


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("L", [2, 3, 4, 6, 8, 9, 12, 15])
@pytest.mark.parametrize("forward", [True, False])
def test_wrap_against_oracle(oracle_mod, L, forward):
    from kapdqmc_amd.dqmc import DQMCBatch
    print(f"case: circulant wrap L={L} forward={forward}", flush=True)
    m = tu(L)
    o = oracle_mod.OracleTU(m, seed=5)
    o.initial_s()
    b = DQMCBatch([m], [5], [0], delay=16)
    b.set_fields(0, o.s)
    rs = np.random.default_rng(L)
    G = rs.standard_normal((m.Ns, m.Ns)) + 1j * rs.standard_normal((m.Ns, m.Ns))
    lt = 3
    sgn = 1 if forward else -1
    D = np.exp(1j * sgn * m.α[lt - 1] * m.η[o.s[:, lt - 1].astype(int) - 1])
    ref = o.wrap(G, D, forward, "B")
    out = []
    for circ in (1, 0):
        b.set_option("circ", circ)
        b.set_G(0, G)
        b.wrap(lt, forward)
        out.append(b.get_G(0))
    b.close()
    assert rel(out[0], ref) < 1e-12, rel(out[0], ref)
    assert rel(out[1], ref) < 1e-12


def sweep(models, circ, sweeps=2, delay=16, precision="fp64", extra=()):
    from kapdqmc_amd.dqmc import DQMCBatch
    b = DQMCBatch(models, list(range(3, 3 + len(models))), list(range(len(models))), delay=delay,
                  precision=precision)
    b.set_option("circ", circ)
    for k, v in extra:
        b.set_option(k, v)
    b.initial_s()
    rows = b.phy_update(sweeps, True)
    out = ([b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows,
           b.near_ties())
    b.close()
    return out


def agree(a, b, gtol=1e-11, exact=False):
    fa, ga, ra, na = a
    fb, gb, rb, nb = b
    assert np.array_equal(na, nb)
    for c in range(len(fa)):
        assert np.array_equal(fa[c], fb[c]), c
        if exact:
            assert np.array_equal(ga[c], gb[c]), c
        else:
            assert rel(ga[c], gb[c]) < gtol, (c, rel(ga[c], gb[c]))
    if exact:
        assert np.array_equal(ra, rb)
    else:
        assert np.allclose(ra, rb, rtol=1e-9, atol=1e-11), np.abs(ra - rb).max(axis=(0, 1))


def test_tu_uneven_batch_sweep():
    # cell edges 3, 4 and 6 in one batch (each chain its own transform length)
    print("case: circulant vs dense sweeps, t-U L = 3/4/6", flush=True)
    models = [tu(3), tu(4, U=2.0), tu(6, U=6.0)]
    agree(sweep(models, 1), sweep(models, 0))


def test_tu_many_chains_and_mixed_precision():
    print("case: circulant vs dense sweeps, 9 chains L = 6, FP64 and mixed", flush=True)
    models = [tu(6, U=2.0 + c % 3) for c in range(9)]
    agree(sweep(models, 1), sweep(models, 0))
    a, b = sweep(models[:2], 1, precision="mixed"), sweep(models[:2], 0, precision="mixed")
    for c in range(2):
        assert np.array_equal(a[0][c], b[0][c])


@pytest.mark.parametrize("lattice", ["HoneyComb60", "SQUARE90", "SQUARE45"])
def test_other_lattices(lattice):
    print(f"case: circulant vs dense sweeps, t-U {lattice} 4x4", flush=True)
    models = [tu(4, lattice=lattice)]
    agree(sweep(models, 1), sweep(models, 0))


def test_complex_hopping():
    print("case: circulant vs dense sweeps, pi-flux SQUARE90 6x6", flush=True)
    models = [tu(6, lattice="SQUARE90", flux=np.pi)]
    agree(sweep(models, 1), sweep(models, 0))


def test_tv_and_tuv_keep_dense_products():
    # their measurements wrap G over a block of slices unstabilised, which amplifies the
    # last-bit difference between the exactly circulant and the eigen-decomposed eK to ~1e-10
    # on some observables: these models keep the dense products (bit-identical either way)
    print("case: t-V and t-U-V take the dense products", flush=True)
    tv = [tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                          Lattice="HoneyComb120", site=[L, L], BatchSize=5) for L in (3, 4)]
    agree(sweep(tv, 1, delay=8), sweep(tv, 0, delay=8), exact=True)
    tuv = [tUV_Hubbard_Para(Ht=1.0, Hu=U, Hv=0.5, Lattice="HoneyComb120", site=[4, 4], Δt=0.1, Θ=1.0,
                            BatchSize=5) for U in (4.0, -4.0)]
    agree(sweep(tuv, 1, delay=8), sweep(tuv, 0, delay=8), exact=True)


def test_dense_fallbacks():
    print("case: dense fallbacks (no instantiation for L = 5; perturbed eK)", flush=True)
    five = [tu(5)]
    agree(sweep(five, 1), sweep(five, 0), exact=True)
    m = copy.deepcopy(tu(4))
    rs = np.random.default_rng(0)
    R = rs.standard_normal(m.eK.shape)
    m.eK = m.eK + 1e-6 * (R + R.T)
    m.eKinv = np.linalg.inv(m.eK)
    agree(sweep([m], 1), sweep([m], 0), exact=True)


@pytest.mark.parametrize("case", ["uneven", "flux", "mixed"])
def test_chained_measurements(case):
    print(f"case: chained vs slot-product measurements, {case}", flush=True)
    prec = "mixed" if case == "mixed" else "fp64"
    if case == "flux":
        models = [tu(6, lattice="SQUARE90", flux=np.pi)]
    else:
        models = [tu(3), tu(4, U=2.0), tu(6, U=6.0)]
    a = sweep(models, 1, precision=prec, extra=(("mchain", 1),))
    b = sweep(models, 1, precision=prec, extra=(("mchain", 0),))
    if case == "mixed":
        for c in range(len(models)):
            assert np.array_equal(a[0][c], b[0][c]), c
            assert np.array_equal(a[1][c], b[1][c]), c
        assert np.allclose(a[2], b[2], rtol=1e-9, atol=1e-11), np.abs(a[2] - b[2]).max()
    else:
        agree(a, b)


def test_chained_measurements_need_circulant_halfek():
    print("case: perturbed HalfeK keeps the slot-product measurements", flush=True)
    m = copy.deepcopy(tu(4))
    rs = np.random.default_rng(1)
    R = rs.standard_normal(m.HalfeK.shape)
    m.HalfeK = m.HalfeK + 1e-6 * (R + R.T)
    m.HalfeKinv = np.linalg.inv(m.HalfeK)
    agree(sweep([m], 1, extra=(("mchain", 1),)), sweep([m], 1, extra=(("mchain", 0),)), exact=True)
