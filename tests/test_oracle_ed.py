"""The oracle's t-U Markov chain against exact results on the 2x2 honeycomb (N = 8, 4↑4↓):
the Fock-space value of the distribution the sweep samples (`tests/fock_exact.py`, exact at
finite Δτ) and, as Δτ → 0, exact diagonalisation of the Hubbard Hamiltonian.

Bars: the Fock machinery is pinned by closed forms (U = 0 rows, |det|² of one field, Richardson
limit = ED ground state); the sampled averages of all six observables of the row (Ek, Eu and the
structure factors CDW0, CDW1, SDW0, SDW1 of `src/tU/phy_update.jl:194-235`) agree with the exact
values jointly at the 95 % level (Hotelling T², `stat_bars.py`) and each within 3σ. A dropped
HalfeK sandwich moves Eu by 0.037 (≈ 4σ here), a swapped γ or a 10 % error in α by ≥ 0.19, a
q-phase off by one cell flips the sign of CDW1 and SDW1 (≥ 180σ), keeping the discarded
`+adjoint(` halves doubles the structure factors (≥ 40σ).
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

import fock_exact as F
from stat_bars import assert_agrees

ROW = ["Ek", "Eu", "CDW0", "CDW1", "SDW0", "SDW1"]


def tiny(U=4.0, dt=0.1, theta=4.0, bs=10):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=dt, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[2, 2], BatchSize=bs, Initial="H0")


def test_fock_free_limit_equals_oracle_rows(oracle_mod):
    """U = 0: every field gives the same G, so one measured row is exact, structure factors
    included (their Fock operators are written from the staggered densities, not from G)."""
    m = tiny(U=0.0, theta=2.0)
    o = oracle_mod.OracleTU(m, seed=3)
    o.initial_s()
    row = o.phy_update(1, True)[0]
    ex = F.expectation(m, structure_factors=True)
    assert np.allclose(row, ex, rtol=1e-9, atol=1e-11), (row, ex)
    Ek, Eu = ex[:2]
    eps = np.sort(np.linalg.eigvalsh(m.Ht * m.K))[: m.Ns // 2]
    assert abs(Ek - 2 * eps.sum()) < 1e-6 and abs(Eu - m.Ns / 4) < 1e-6


def test_fock_weight_of_one_field_is_abs_det_squared():
    """⟨ψ_T| Π_l D↑_l E ⊗ D↓_l E |ψ_T⟩ = |det(Pᵀ B_Nt ⋯ B_1 P)|², B_l = diag(e^{iαη_s}) eK."""
    m = tiny(theta=0.5)
    rng = np.random.default_rng(7)
    for _ in range(3):
        s = rng.integers(1, 5, size=(m.Ns, m.Nt)).astype(np.uint8)
        B = np.eye(m.Ns, dtype=np.complex128)
        for lt in range(m.Nt):
            B = np.diag(np.exp(1j * m.α[lt] * m.η[s[:, lt] - 1])) @ m.eK @ B
        w = abs(np.linalg.det(m.Pt.T @ B @ m.Pt)) ** 2
        assert abs(F.field_weight(m, s) - w) < 1e-10 * w


def test_fock_dtau_limit_is_ground_state():
    """Richardson extrapolation of the exact finite-Δτ values (error O(Δτ²)) = ED."""
    E0, Ek0, Eu0 = F.ground_state(tiny().K, 1.0, 4.0)
    a, b = F.expectation(tiny(dt=0.1)), F.expectation(tiny(dt=0.05))
    Ek = (4 * b[0] - a[0]) / 3
    Eu = (4 * b[1] - a[1]) / 3
    assert abs(Ek - Ek0) < 2e-3 and abs(Eu - Eu0) < 5e-4
    # the finite-Δτ values themselves differ from the ground state by O(Δτ²)
    assert abs(a[0] - Ek0) > 0.03


def test_oracle_samples_exact_distribution(oracle_mod):
    m = tiny()
    exact = np.array(F.expectation(m, structure_factors=True))
    nch, nsw = 32, 800

    def chain(c):
        o = oracle_mod.OracleTU(m, seed=500 + c, chain=c)
        o.initial_s()
        o.phy_update(20, False)
        return o.phy_update(nsw, True).mean(axis=0)

    with ThreadPoolExecutor(8) as ex:
        per = np.array(list(ex.map(chain, range(nch))))
    err = per.std(axis=0, ddof=1) / np.sqrt(nch)
    assert_agrees(per, exact, ROW)
    # the slips named above would be far outside these bars
    assert abs(2 * exact[3]) > 50 * err[3] and abs(2 * exact[5]) > 50 * err[5]
    assert abs(exact[2]) > 20 * err[2] and abs(exact[4]) > 20 * err[4]


# ----------------------------------------------------------------- Rényi-2 entropy
def _region(m):
    from kapdqmc_amd.geometry import area_index
    return np.asarray(area_index(m.Lattice, m.site, ([1, 1], [1, 1])), dtype=np.int32)   # cell (1,1): 2 sites


def test_renyi2_exact_pins():
    """The Fock-space Rényi-2: a product state has S2 = 0 on any region; the U = 0 ground
    state agrees with the free-fermion formula S2 = −2 log det(G_A² + (1 − G_A)²) per spin
    (the Grover determinant the swap-replica estimator averages, `GroverMatrix!`
    src/public/GF.jl:101-109); the finite-Δτ value of the sampled ensemble tends to the
    ground-state value as O(Δτ²) (Richardson extrapolation of Δτ = 0.1, 0.05)."""
    m = tiny(U=0.0)
    A = _region(m) - 1
    confs, psi = F.ground_state_vector(m.K, 1.0, 0.0)
    w, V = np.linalg.eigh(m.K)
    P = V[:, : m.Ns // 2]
    GA = (np.eye(m.Ns) - P @ P.T)[np.ix_(A, A)]            # G = <c c†> of the free ground state
    free = -2 * np.log(np.linalg.det(GA @ GA + (np.eye(len(A)) - GA) @ (np.eye(len(A)) - GA)))
    assert abs(F.renyi2(confs, psi, A) - free) < 1e-8
    prod = np.zeros_like(psi)
    prod[0, -1] = 1.0                                       # one Fock state: no entanglement
    assert abs(F.renyi2(confs, prod, A)) < 1e-12
    confs, psi = F.ground_state_vector(m.K, 1.0, 4.0)
    gs = F.renyi2(confs, psi, A)
    a, b = F.renyi2_sampled(tiny(dt=0.1), A), F.renyi2_sampled(tiny(dt=0.05), A)
    assert abs((4 * b - a) / 3 - gs) < 1e-3 and abs(a - gs) > 5e-3


def test_oracle_swap_replica_renyi2():
    """S2(A) = −log⟨O⟩_{λ=0} − log⟨O⟩_{λ=1/2} (Nλ = 2, B empty) of the oracle's
    ctrl_SCEEicr (src/tU/SCEE.jl:1-262) on 2 sites of the 2x2 honeycomb at U = 4 (Δτ = 0.05,
    2Θ = 8) against the exact value of the sampled ensemble (1.5219; ground state 1.5248);
    bar 3σ. The λ = 1/2 chains equilibrate slowly (after 100 warm-up sweeps ⟨O⟩ is still
    falling and S2 reads ~0.01 low): 300 warm-up sweeps there, 50 at λ = 0."""
    from oracle import oracle as O
    from kapdqmc_amd.rng import initial_fields
    m = tiny(dt=0.05)
    iA, iB = _region(m), np.zeros(0, dtype=np.int32)
    exact = F.renyi2_sampled(m, iA - 1)

    def chain(args):
        lam, c = args
        s1, s2 = initial_fields(100 + c, c, m.Ns, m.Nt), initial_fields(200 + c, c, m.Ns, m.Nt)
        e = O.OracleSCEE(m, iA, iB, lam, 2, s1, s2, seed=100 + c, chain=c)
        e.init()
        e.sweeps(300 if lam > 0 else 50)
        return e.sweeps(150).mean()

    est = {}
    for lam in (0.0, 0.5):
        with ThreadPoolExecutor(8) as ex:
            v = np.array(list(ex.map(chain, [(lam, c) for c in range(32)])))
        est[lam] = (v.mean(), v.std(ddof=1) / np.sqrt(len(v)))
    S2 = -np.log(est[0.0][0]) - np.log(est[0.5][0])
    err = np.hypot(est[0.0][1] / est[0.0][0], est[0.5][1] / est[0.5][0])
    print(f"S2 {S2:.4f} ± {err:.4f} (exact {exact:.4f})")
    assert abs(S2 - exact) < 3 * err
