"""The GPU t-U sweep's observables against the exact value of the distribution it samples
(2x2 honeycomb, N = 8, U = 4, Δτ = 0.1, 2Θ = 8; `tests/fock_exact.py`), through the C ABI.

512 chains batched in one handle, 600 sweeps each, give a chain-to-chain error of ≈ 0.0043 on
Eu (256 x 300 gave 0.011: Eu decorrelates slowly on this lattice), so a dropped HalfeK
sandwich in the measurement (Eu − 0.037) or a wrong HS constant (≥ 0.19) fails; the bar is 3σ
for Ek and Eu in FP64 and in mixed precision (FP32 G between stabilisations).
"""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

import fock_exact as F

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_samples_exact_distribution(precision):
    from kapdqmc_amd.dqmc import DQMCBatch
    from kapdqmc_amd.rng import initial_fields
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=4.0, Θquench=0.0,
                        Lattice="HoneyComb120", site=[2, 2], BatchSize=10, Initial="H0")
    Ek_x, Eu_x = F.expectation(m)
    nch, nsw = 512, 600
    seeds = [9000 + c for c in range(nch)]
    print(f"case: 2x2 honeycomb exact distribution, {nch} chains, {precision}", flush=True)
    b = DQMCBatch([m] * nch, seeds, list(range(nch)), delay=8, precision=precision)
    for c in range(nch):
        b.set_fields(c, initial_fields(seeds[c], c, m.Ns, m.Nt))
    b.phy_update(20, False)
    rows = b.phy_update(nsw, True)                      # [chains, sweeps, 6]
    per = rows[:, :, :2].mean(axis=1)
    mean = per.mean(axis=0)
    err = per.std(axis=0, ddof=1) / np.sqrt(nch)
    print(f"Ek {mean[0]:.4f} ± {err[0]:.4f} (exact {Ek_x:.4f}); "
          f"Eu {mean[1]:.4f} ± {err[1]:.4f} (exact {Eu_x:.4f})", flush=True)
    assert err[1] < 0.009
    assert abs(mean[0] - Ek_x) < 3 * err[0]
    assert abs(mean[1] - Eu_x) < 3 * err[1]


def test_gpu_swap_replica_renyi2():
    """S2(A) = −log⟨O⟩_{λ=0} − log⟨O⟩_{λ=1/2} (Nλ = 2, B empty) of the GPU swap-replica sweep
    on 2 sites of the 2x2 honeycomb at U = 4 (Δτ = 0.05, 2Θ = 8), 256 chains per λ, against
    the exact value of the sampled ensemble (1.5219; ground state 1.5248); bar 3σ. The λ = 1/2
    chains equilibrate slowly (after 10 or 100 warm-up sweeps ⟨O⟩ is still falling and S2 reads
    1.510-1.512): 400 warm-up sweeps there, 50 at λ = 0."""
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
        v = b.sweeps(200).mean(axis=1)
        b.close()
        est[lam] = (v.mean(), v.std(ddof=1) / np.sqrt(nch))
    S2 = -np.log(est[0.0][0]) - np.log(est[0.5][0])
    err = np.hypot(est[0.0][1] / est[0.0][0], est[0.5][1] / est[0.5][0])
    print(f"S2 {S2:.4f} ± {err:.4f} (exact {exact:.4f})", flush=True)
    assert err < 0.01
    assert abs(S2 - exact) < 3 * err
