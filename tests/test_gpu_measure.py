"""The GPU measurement (tile reduction over G + per-chain finish) against the oracle's
phy_measure loops and the separately written vectorised formulas of test_measure_pins, on
random G (no symmetry hides a slip) for lattices with one tile (N ≤ 32), several tiles and
L1 != L2, through the C ABI (kdq_set_G, then kdq_measure at slice Nt/2 where no wrap
applies). Bar: 1e-9 relative (the sums run in a different order)."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

import test_measure_pins as P

pytestmark = pytest.mark.gpu

SITES = [[3, 4], [8, 5], [12, 12]]


def make(kind, site, U=4.0):
    if kind == "tU":
        return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=0.2, Θquench=0.0, Lattice="HoneyComb120",
                               site=site, BatchSize=1)
    if kind == "tV":
        return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=0.2, Θquench=0.0, Lattice="HoneyComb120",
                               site=site, BatchSize=1)
    return tUV_Hubbard_Para(Ht=1.0, Hu=U, Hv=0.5, Lattice="HoneyComb120", site=site, Δt=0.1, Θ=0.2, BatchSize=1)


@pytest.mark.parametrize("site", SITES)
@pytest.mark.parametrize("kind,U", [("tU", 4.0), ("tV", 0.0), ("tUV", 4.0), ("tUV", -4.0)])
def test_measure_random_G(oracle_mod, kind, U, site):
    print(f"case: measure {kind} U={U} site={site}", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch, Initial_s
    m = make(kind, site, U)
    cls = {"tU": "OracleTU", "tV": "OracleTV", "tUV": "OracleTUV"}[kind]
    o = getattr(oracle_mod, cls)(m, seed=0)
    o.initial_s()
    b = DQMCBatch([m, m], [0, 1], [0, 1], delay=8)
    rng = np.random.default_rng(sum(site) + 7)
    cplx = kind == "tU" or (kind == "tUV" and U > 0)
    Gs = []
    for c in range(2):
        b.set_fields(c, Initial_s(m, c, c))
        G = P._random_G(m.Ns, rng, complex_=cplx)
        b.set_G(c, G)
        Gs.append(G)
    got = b.measure(m.Nt // 2)
    vec = {"tU": P.tu_observables, "tV": P.tv_observables, "tUV": P.tuv_observables}[kind]
    for c in range(2):
        ref = o.measure(m.Nt // 2, Gs[c])
        assert np.allclose(got[c], ref, rtol=1e-9, atol=1e-12), (c, got[c], ref)
        assert np.allclose(got[c], vec(m, P._sandwich(m, Gs[c])), rtol=1e-9, atol=1e-12)
    b.close()
