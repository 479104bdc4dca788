"""The reference's own test workloads, and the CPU-runnable headline configuration, through the
GPU path against the FP64 oracle (C ABI).

* t-U quench (`test/tU/HCtU_runtests.jl:12-36`): honeycomb 3x3, Hu 0 → 3.8 (α = 0 on the
  relax slices), Θrelax = Θquench = 0.4, BatchSize 10 (nodes 0, 8, 16); `phy_update` 10 sweeps
  with record = true, then `ctrl_SCEEicr` 2 sweeps, λ = 0.5, Nλ = 2, from ss = [s, s], regions
  area_index(([1,1],[1,3])) and area_index(([1,1],[1,2])).
* t-V quench and static swap-replica runs (`test/tV/tV_runtests.jl:12-52`): Hv 0.1 → 1.0,
  Θrelax = 0, Θquench = 1, BatchSize 5; and Hv = 1, Δτ = 0.05, Θrelax = 1, BatchSize 5; 2 SCEE
  sweeps each. The t-V quench also runs through `phy_update` (commented out in the reference
  test, `:16`), which exercises the per-slice α of the bond kernels.
* t-U-V (`test/tUV/tUV_runtests.jl:12-31`): Hu = 4, Hv = 0, Θ = 1, BatchSize 10; `phy_update`
  10 sweeps with record = true, then 2 SCEE sweeps.
* `Initial_s` (`src/tU/GreenMatrix.jl:1-8`, `src/tV/GreenMatrix.jl:1-10`,
  `src/tUV/GreenMatrix.jl:1-8`): the GPU kernel behind kdq_initial_s against the oracle's,
  bit for bit, for the three field layouts and several chains of one batch.
* The CPU-runnable headline case: honeycomb 3x3, U = 4, Δτ = 0.1, 2Θ = 10 (100 slices),
  BatchSize 10, 200 warm-up and 1000 measured sweeps: the whole trajectory (every field after
  the run, every one of the 1000 rows) against the oracle's.

Bars as in the other parity tests: fields bit-exact (sweeps: with equal near-tie counts); G within 1e-10
relative (t-V / t-U-V swap replicas 1e-9); rows and O within 1e-9 relative.
"""
import numpy as np
import pytest

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def dq():
    from kapdqmc_amd import dqmc
    return dqmc


def ref_regions(m):
    L = m.site[1]
    iA = area_index(m.Lattice, m.site, ([1, 1], [L // 3, L]))
    iB = area_index(m.Lattice, m.site, ([1, 1], [L // 3, 2 * L // 3]))
    return np.asarray(iA, dtype=np.int64), np.asarray(iB, dtype=np.int64)


def tu_quench():
    return tU_Hubbard_Para(Ht=1.0, Hu1=0.0, Hu2=3.8, Θrelax=0.4, Θquench=0.4, Lattice="HoneyComb120",
                           site=[3, 3], Δt=0.1, BatchSize=10, Initial="H0")


def tv_quench():
    return tV_Hubbard_Para(Ht=1.0, Hv1=0.1, Hv2=1.0, Θrelax=0.0, Θquench=1.0, Lattice="HoneyComb120",
                           site=[3, 3], Δt=0.1, BatchSize=5, Initial="H0")


def tv_static():
    return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[3, 3], Δt=0.05, BatchSize=5, Initial="H0")


def tuv_ref():
    return tUV_Hubbard_Para(Ht=1.0, Hu=4.0, Hv=0.0, Θ=1.0, Lattice="HoneyComb120", site=[3, 3], Δt=0.1,
                            BatchSize=10, Initial="H0")


ORACLES = {"tU": "OracleTU", "tV": "OracleTV", "tUV": "OracleTUV"}


def kind(m):
    return {"tU_Hubbard_Para_": "tU", "tV_Hubbard_Para_": "tV", "tUV_Hubbard_Para_": "tUV"}[type(m).__name__]


def sweep_pair(dq, oracle_mod, m, seed, delay):
    o = getattr(oracle_mod, ORACLES[kind(m)])(m, seed=seed)
    o.initial_s()
    b = dq.DQMCBatch([m], [seed], [0], delay=delay)
    b.set_fields(0, o.s)
    return o, b


def scee_pair(dq, oracle_mod, m, s, seed, delay=8):
    iA, iB = ref_regions(m)
    cls = oracle_mod.OracleSCEE if kind(m) == "tU" else oracle_mod.OracleSCEE2
    o = cls(m, iA.astype(np.int32), iB.astype(np.int32), 0.5, 2, s.copy(), s.copy(), seed=seed)
    b = dq.SCEEBatch([m], iA, iB, 0.5, 2, [seed], [0], delay=delay)
    b.set_fields(0, 0, s)
    b.set_fields(0, 1, s)
    return o, b


def check_scee(o, b, gtol, otol, sweeps=2):
    Oo = o.sweeps(sweeps)
    Ob = b.sweeps(sweeps)[0]
    for r in (0, 1):
        assert np.array_equal(b.get_fields(0, r), o.fields(r)), f"replica {r} fields differ"
        for name, x, y in zip(("Gt", "G0", "Gt0", "G0t"), b.G4(0, r), o.G4(r)):
            assert rel(x, y) < gtol, (r, name, rel(x, y))
    assert np.allclose(Ob, Oo, rtol=otol, atol=0), (Ob, Oo)


def test_quench_profiles_are_the_reference_ones():
    """α on the relax slices is 0 for Hu1 = 0; the profile is symmetric; Nt and the nodes follow
    the reference formulas (src/tU/model.jl:40-52,81-85)."""
    m = tu_quench()
    assert m.Nt == 16 and list(m.nodes) == [0, 8, 16]
    assert np.all(m.α[:4] == 0) and np.all(m.α[-4:] == 0)
    assert np.allclose(m.α, m.α[::-1]) and m.α[7] == np.sqrt(0.1 * 3.8 / 2)
    v = tv_quench()
    assert v.Nt == 20 and np.isclose(v.α[0], np.sqrt(0.1 * 0.19 / 2)) and np.isclose(v.α[9], np.sqrt(0.1 / 2))


@pytest.mark.parametrize("delay", [32, 1])
def test_tu_quench_phy_update_then_scee(dq, oracle_mod, delay):
    """HCtU_runtests.jl:12-36 with Philox draws: 10 recorded sweeps, then 2 SCEE sweeps."""
    print(f"case: t-U quench workload delay={delay}", flush=True)
    m = tu_quench()
    o, b = sweep_pair(dq, oracle_mod, m, seed=1234, delay=delay)
    ro = o.phy_update(10, True)
    rb = b.phy_update(10, True)[0]
    assert np.array_equal(b.get_fields(0), o.s)
    assert b.near_ties()[0] == o.near_ties
    assert rel(b.get_G(0), o.G) < 1e-10
    assert np.allclose(rb, ro, rtol=1e-9, atol=1e-11)
    b.close()
    s = o.s.copy()
    print("case: t-U quench SCEE", flush=True)
    so, sb = scee_pair(dq, oracle_mod, m, s, seed=1234, delay=min(delay, 16))
    check_scee(so, sb, 1e-10, 1e-9)
    sb.close()


@pytest.mark.parametrize("mk", [tv_quench, tv_static], ids=["quench", "static"])
def test_tv_reference_scee(dq, oracle_mod, mk):
    """tV_runtests.jl:12-30 (quench) and :34-52 (Δτ = 0.05): 2 SCEE sweeps from ss = [s, s]."""
    print(f"case: t-V reference SCEE {mk.__name__}", flush=True)
    m = mk()
    s = dq.Initial_s(m, 1234)
    so, sb = scee_pair(dq, oracle_mod, m, s, seed=1234)
    check_scee(so, sb, 1e-9, 1e-8)
    sb.close()


@pytest.mark.parametrize("delay", [8, 1])
def test_tv_quench_phy_update(dq, oracle_mod, delay):
    print(f"case: t-V quench phy_update delay={delay}", flush=True)
    m = tv_quench()
    o, b = sweep_pair(dq, oracle_mod, m, seed=77, delay=delay)
    ro = o.phy_update(5, True)
    rb = b.phy_update(5, True)[0]
    assert np.array_equal(b.get_fields(0), o.s)
    assert b.near_ties()[0] == o.near_ties
    assert rel(b.get_G(0), o.G) < 1e-10
    assert np.allclose(rb, ro, rtol=1e-9, atol=1e-11)


def test_tuv_reference_workload(dq, oracle_mod):
    """tUV_runtests.jl:12-31: 10 recorded sweeps (phy_measure with the documented xy_i(site, x, y)),
    then 2 SCEE sweeps from ss = [s, s]."""
    print("case: t-U-V reference workload", flush=True)
    m = tuv_ref()
    o, b = sweep_pair(dq, oracle_mod, m, seed=1234, delay=8)
    ro = o.phy_update(10, True)
    rb = b.phy_update(10, True)[0]
    assert np.array_equal(b.get_fields(0), o.s)
    assert b.near_ties()[0] == o.near_ties
    assert rel(b.get_G(0), o.G) < 1e-10
    assert np.allclose(rb, ro, rtol=1e-9, atol=1e-11)
    b.close()
    so, sb = scee_pair(dq, oracle_mod, m, o.s.copy(), seed=1234)
    check_scee(so, sb, 1e-9, 1e-8)
    sb.close()


@pytest.mark.parametrize("mk", [tu_quench, tv_quench, tuv_ref], ids=["tU", "tV", "tUV"])
def test_initial_s_kernel_matches_oracle(dq, oracle_mod, mk):
    """kdq_initial_s (GPU kernel) = orc_*_initial_s, bit for bit, for chains 0..3 of one batch
    (each chain's own Philox stream; seeds differ per chain)."""
    print(f"case: Initial_s kernel {mk.__name__}", flush=True)
    m = mk()
    seeds = [1234, 7, 7, 99]
    b = dq.DQMCBatch([m] * 4, seeds, [0, 1, 2, 3], delay=8)
    b.initial_s()
    cls = getattr(oracle_mod, ORACLES[kind(m)])
    for c in range(4):
        o = cls(m, seed=seeds[c], chain=c)
        o.initial_s()
        got = b.get_fields(c)
        assert got.shape == o.s.shape
        assert np.array_equal(got, o.s), f"chain {c}"
        vals, counts = np.unique(got, return_counts=True)
        assert list(vals) == [1, 2, 3, 4] and counts.min() > 0.15 * got.size
    assert not np.array_equal(b.get_fields(1), b.get_fields(2))       # same seed, other chain


def test_headline_cpu_case_full_trajectory(dq, oracle_mod):
    """Honeycomb 3x3, U = 4, Δτ = 0.1, 2Θ = 10, BatchSize 10: 200 warm-up + 1000 recorded sweeps
    on the GPU equal the oracle's run decision for decision, row for row."""
    print("case: 3x3 2Theta=10 BS=10, 200 + 1000 sweeps", flush=True)
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=5.0, Θquench=0.0, Lattice="HoneyComb120",
                        site=[3, 3], BatchSize=10, Initial="H0")
    assert m.Nt == 100 and len(m.nodes) == 11
    o, b = sweep_pair(dq, oracle_mod, m, seed=1234, delay=16)
    o.phy_update(200, False)
    b.phy_update(200, False)
    assert np.array_equal(b.get_fields(0), o.s)
    ro = o.phy_update(1000, True)
    rb = b.phy_update(1000, True)[0]
    assert np.array_equal(b.get_fields(0), o.s)
    assert b.near_ties()[0] == o.near_ties
    assert rel(b.get_G(0), o.G) < 1e-10
    assert np.allclose(rb, ro, rtol=1e-9, atol=1e-11)
    print(f"means: {rb.mean(axis=0)}", flush=True)
