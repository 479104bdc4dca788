"""The GPU t-U sweep's observables against the exact value of the distribution it samples
(2x2 honeycomb, N = 8, U = 4, Δτ = 0.1, 2Θ = 8; `tests/fock_exact.py`), through the C ABI.

512 chains batched in one handle, 600 sweeps each, give a chain-to-chain error of ≈ 0.0043 on
Eu (256 x 300 gave 0.011: Eu decorrelates slowly on this lattice), so a dropped HalfeK
sandwich in the measurement (Eu − 0.037), a wrong HS constant (≥ 0.19), a q-phase slip in the
structure factors (CDW1, SDW1 change sign) or the discarded `+adjoint(` halves kept (the
structure factors double) fails. Bar: all six observables of the row jointly at the 95 % level
(Hotelling T² over the chains, `stat_bars.py`) and each within 3σ, in FP64 and in mixed
precision (FP32 G between stabilisations); S2 within 2σ.
"""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

import fock_exact as F
from stat_bars import assert_agrees

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_samples_exact_distribution(precision):
    from kapdqmc_amd.dqmc import DQMCBatch
    from kapdqmc_amd.rng import initial_fields
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=4.0, Θquench=0.0,
                        Lattice="HoneyComb120", site=[2, 2], BatchSize=10, Initial="H0")
    exact = np.array(F.expectation(m, structure_factors=True))
    nch, nsw = 512, 600
    seeds = [9000 + c for c in range(nch)]
    print(f"case: 2x2 honeycomb exact distribution, {nch} chains, {precision}", flush=True)
    b = DQMCBatch([m] * nch, seeds, list(range(nch)), delay=8, precision=precision)
    for c in range(nch):
        b.set_fields(c, initial_fields(seeds[c], c, m.Ns, m.Nt))
    b.phy_update(20, False)
    rows = b.phy_update(nsw, True)                      # [chains, sweeps, 6]
    per = rows.mean(axis=1)
    mean = per.mean(axis=0)
    err = per.std(axis=0, ddof=1) / np.sqrt(nch)
    for k, name in enumerate(["Ek", "Eu", "CDW0", "CDW1", "SDW0", "SDW1"]):
        print(f"{name} {mean[k]:.5f} ± {err[k]:.5f} (exact {exact[k]:.5f})", flush=True)
    assert err[1] < 0.009
    assert_agrees(per, exact, ["Ek", "Eu", "CDW0", "CDW1", "SDW0", "SDW1"])


def _grover_abs2(G1, G2, A):
    """|det(G1_A G2_A + (I − G1_A)(I − G2_A))|², the Grover determinant of get_ABGM!
    (src/tU/SCEE.jl:264-276, GroverMatrix! src/public/GF.jl:101-109), from scratch."""
    a, b = G1[np.ix_(A, A)], G2[np.ix_(A, A)]
    I = np.eye(len(A))
    return abs(np.linalg.det(a @ b + (I - a) @ (I - b))) ** 2


def test_gpu_swap_replica_renyi2():
    """S2(A) = −log⟨O⟩_{λ=0} − log⟨O⟩_{λ=1/2} (Nλ = 2, B empty) of the GPU swap-replica sweep
    on 2 sites of the 2x2 honeycomb at U = 4 (Δτ = 0.05, 2Θ = 8), 256 chains per λ, against
    the exact value of the sampled ensemble (1.5219; ground state 1.5248); bar 2σ.

    Equilibration diagnostics, so that a slow warm-up cannot hide a bias: after the warm-up
    (400 sweeps at λ = 1/2, 50 at λ = 0) the 200 measured sweeps are cut into four blocks of 50;
    the slope of the block means of O against block index (pooled over chains, error from the
    chain-to-chain spread of each chain's slope) must be within 3σ of zero, and τ_int of O
    (pooled chains) is reported. The running det g the sweep carries (updated per accepted flip
    by GMupdate!) must equal a Grover determinant computed from scratch from both replicas'
    G(Θ) (kdq_scee_get_G4), to 1e-8 relative, for every chain."""
    from kapdqmc_amd.stats import tau_int_chains
    from kapdqmc_amd.dqmc import SCEEBatch
    from kapdqmc_amd.geometry import area_index
    from kapdqmc_amd.rng import initial_fields
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.05, Θrelax=4.0, Θquench=0.0,
                        Lattice="HoneyComb120", site=[2, 2], BatchSize=10, Initial="H0")
    iA = np.asarray(area_index(m.Lattice, m.site, ([1, 1], [1, 1])), dtype=np.int64)
    iB = np.zeros(0, dtype=np.int64)
    exact = F.renyi2_sampled(m, iA - 1)
    nch = 256
    est = {}
    for lam in (0.0, 0.5):
        print(f"case: 2x2 swap-replica S2, {nch} chains, lambda={lam}", flush=True)
        b = SCEEBatch([m] * nch, iA, iB, lam, 2, [100 + c for c in range(nch)], list(range(nch)))
        for c in range(nch):
            b.set_fields(c, 0, initial_fields(100 + c, c, m.Ns, m.Nt))
            b.set_fields(c, 1, initial_fields(200 + c, c, m.Ns, m.Nt))
        b.sweeps(400 if lam > 0 else 50)
        O = b.sweeps(200)                                   # [chains, sweeps]
        A = iA - 1
        tracked = b.detg()[:, 0]
        fresh = np.array([_grover_abs2(b.G4(c, 0)[1], b.G4(c, 1)[1], A) for c in range(nch)])
        b.close()
        rel = np.abs(tracked - fresh) / np.abs(fresh)
        blocks = O.reshape(nch, 4, 50).mean(axis=2)         # [chains, block]
        x = np.arange(4) - 1.5
        slopes = blocks @ x / (x @ x)
        slope, slope_err = slopes.mean(), slopes.std(ddof=1) / np.sqrt(nch)
        tau = tau_int_chains(O)
        print(f"lambda={lam}: block means {blocks.mean(axis=0)}, slope {slope:.2e} ± {slope_err:.2e}, "
              f"tau_int(O) {tau:.2f}, max|det g tracked / fresh - 1| {rel.max():.2e}", flush=True)
        assert rel.max() < 1e-8
        assert abs(slope) < 3 * slope_err
        v = O.mean(axis=1)
        est[lam] = (v.mean(), v.std(ddof=1) / np.sqrt(nch))
    S2 = -np.log(est[0.0][0]) - np.log(est[0.5][0])
    err = np.hypot(est[0.0][1] / est[0.0][0], est[0.5][1] / est[0.5][0])
    print(f"S2 {S2:.4f} ± {err:.4f} (exact {exact:.4f})", flush=True)
    assert err < 0.01
    assert abs(S2 - exact) < 2 * err
