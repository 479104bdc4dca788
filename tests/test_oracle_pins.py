"""Pins the oracle (CPU FP64 restatement) against independent facts: numpy/LAPACK
factorizations, closed forms of the free-fermion limit, brute-force determinants and
recomputed Green's functions. None of these re-types the oracle's own formulas."""
import numpy as np
import pytest

from kapdqmc_amd import geometry as geo
from kapdqmc_amd.models import tU_Hubbard_Para


def tu(L=3, U=4.0, theta=1.0, bs=5, dt=0.1, lattice="HoneyComb120", initial="H0"):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=dt, Θrelax=theta, Θquench=0.0, Lattice=lattice,
                           site=[L, L], BatchSize=bs, Initial=initial)


# ----------------------------------------------------------------------- geometry
@pytest.mark.parametrize("lattice", geo.LATTICES)
def test_neighbours_oracle_equal_host(oracle_mod, lattice):
    site = [4, 3]
    for i in range(1, 2 * 4 * 3 + 1):
        assert oracle_mod.nn2idx(lattice, site, i) == geo.nn2idx(lattice, site, i)


def test_honeycomb120_table_by_hand():
    # 3x3 cells, A sites odd (1-based). Site 1 = A of cell (1,1): its B partner 2,
    # the B site of cell (2,1) = 4 and the B site of cell (1,3) = 2*(1+2*3) = 14.
    assert geo.nn2idx("HoneyComb120", [3, 3], 1) == [2, 4, 14]
    # Site 2 = B of cell (1,1): A partner 1, A of cell (1,2) = 7, A of cell (3,1) = 5.
    assert geo.nn2idx("HoneyComb120", [3, 3], 2) == [1, 7, 5]


@pytest.mark.parametrize("lattice", ["HoneyComb120", "HoneyComb60"])
def test_honeycomb_K_structure(oracle_mod, lattice):
    site = [6, 6]
    K = geo.nnK_Matrix(lattice, site)
    assert np.array_equal(K, oracle_mod.nnK(lattice, site))
    assert np.array_equal(K, K.T)
    assert np.all(K.sum(1) == 3)
    A = np.arange(K.shape[0]) % 2 == 0
    assert not K[np.ix_(A, A)].any() and not K[np.ix_(~A, ~A)].any()  # bipartite
    # spectrum = ±|1 + e^{ik1} + e^{ik2}| over the 6x6 Brillouin zone (any of the two bases)
    k = 2 * np.pi * np.arange(6) / 6
    f = np.abs(1 + np.exp(1j * k)[:, None] + np.exp(1j * k)[None, :]).ravel()
    ref = np.sort(np.concatenate([f, -f]))
    assert np.allclose(np.sort(np.linalg.eigvalsh(K)), ref, atol=1e-12)


def test_area_index_rectangle():
    idx = geo.area_index("HoneyComb120", [3, 3], ([1, 1], [1, 3]))
    # cells (1,1),(1,2),(1,3): A/B pairs (1,2), (7,8), (13,14)
    assert list(idx) == [1, 2, 7, 8, 13, 14]


# ---------------------------------------------------------------- linear algebra
def test_jacobi_matches_numpy(oracle_mod):
    rs = np.random.default_rng(0)
    A = rs.standard_normal((30, 30))
    A = A + A.T
    w, V = oracle_mod.eigh(A)
    assert np.allclose(w, np.linalg.eigvalsh(A), atol=1e-11)
    assert np.allclose(V.T @ V, np.eye(30), atol=1e-12)
    assert np.allclose(V @ np.diag(w) @ V.T, A, atol=1e-11)


def test_householder_qr(oracle_mod):
    rs = np.random.default_rng(1)
    A = rs.standard_normal((40, 17)) + 1j * rs.standard_normal((40, 17))
    Q, logabsr = oracle_mod.qr_q(A)
    assert np.allclose(Q.conj().T @ Q, np.eye(17), atol=1e-13)
    R = Q.conj().T @ A
    assert np.allclose(np.tril(R, -1), 0, atol=1e-12)           # upper triangular
    assert np.allclose(Q @ R, A, atol=1e-12)
    _, Rn = np.linalg.qr(A)
    assert np.isclose(logabsr, np.log(np.abs(np.diag(Rn))).sum(), atol=1e-10)


def test_lu_solve_inverse_det(oracle_mod):
    rs = np.random.default_rng(2)
    A = rs.standard_normal((25, 25)) + 1j * rs.standard_normal((25, 25))
    B = rs.standard_normal((25, 7)) + 1j * rs.standard_normal((25, 7))
    X, info = oracle_mod.lu_solve(A, B)
    assert info == 0 and np.allclose(X, np.linalg.solve(A, B), atol=1e-11)
    Ai, info = oracle_mod.inv(A)
    assert info == 0 and np.allclose(Ai, np.linalg.inv(A), atol=1e-11)


def test_get_G_is_projector(oracle_mod):
    rs = np.random.default_rng(3)
    BL = rs.standard_normal((6, 14)) + 1j * rs.standard_normal((6, 14))
    BR = rs.standard_normal((14, 6)) + 1j * rs.standard_normal((14, 6))
    G, logdet = oracle_mod.get_G(BL, BR)
    assert np.allclose(G, np.eye(14) - BR @ np.linalg.inv(BL @ BR) @ BL, atol=1e-12)
    assert np.isclose(logdet, np.log(abs(np.linalg.det(BL @ BR))), atol=1e-10)
    assert np.allclose(G @ G, G, atol=1e-12)


# ------------------------------------------------------------------ model constants
def test_hs_identity():
    """¼ Σ_l γ_l e^{i α η_l x} = e^{-α² x²} + O((α x)⁸) for x ∈ {0, ±1, ±2, ±3}: the
    4-valued SU(2) decomposition of src/tU/model.jl:54-56. Halving α must shrink the
    error by ~2⁸; x = 0 is exact (Σ γ_l = 4)."""
    m = tu()
    err = lambda a, x: abs(0.25 * np.sum(m.γ * np.exp(1j * a * m.η * x)) - np.exp(-a * a * x * x))
    assert err(0.2, 0) < 1e-15
    for x in (1, -1, 2, -2, 3, -3):
        assert err(0.2, x) / err(0.1, x) > 200
    assert err(0.2, 1) < 3e-8


def test_ek_inverse_and_symmetry():
    m = tu(L=4)
    n = m.Ns
    assert np.allclose(m.eK @ m.eKinv, np.eye(n), atol=1e-12)
    assert np.allclose(m.HalfeK @ m.HalfeK, m.eK, atol=1e-12)
    assert np.allclose(m.eK, m.eK.T, atol=1e-14)


def test_trial_state_projector_oracle_eigensolver(oracle_mod):
    m = tu(L=4)
    KK = m.K + 1e-5 * np.diag(np.tile([-1.0, 1.0], m.Ns // 2))
    w, V = oracle_mod.eigh(KK)
    P1 = V[:, : m.Ns // 2]
    assert np.allclose(P1 @ P1.T, m.Pt @ m.Pt.T, atol=1e-8)


# --------------------------------------------------------------------- t-U algebra
def brute_weight_det(m, s):
    """det(Pᴴ B(Nt) ... B(1) P) with B(l) = diag(cis(α η_s)) eK, no stabilisation (tiny sizes)."""
    X = m.Pt.astype(complex)
    for lt in range(1, m.Nt + 1):
        D = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])
        X = D[:, None] * (m.eK @ X)
    return np.linalg.det(m.Pt.T @ X)


def brute_G(m, s, tau):
    R = m.Pt.astype(complex)
    for lt in range(1, tau + 1):
        D = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])
        R = D[:, None] * (m.eK @ R)
    Lm = m.Pt.T.astype(complex)
    for lt in range(m.Nt, tau, -1):
        D = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])
        Lm = (Lm * D[None, :]) @ m.eK
    return np.eye(m.Ns) - R @ np.linalg.inv(Lm @ R) @ Lm


def test_Gtau_matches_unstabilised_product(oracle_mod):
    m = tu(L=2, theta=0.5, bs=2)
    o = oracle_mod.OracleTU(m, seed=3)
    o.initial_s()
    for tau in (0, 3, 5, 10):
        assert np.allclose(o.Gtau(tau), brute_G(m, o.s, tau), atol=1e-11)


def test_wrap_is_similarity_transform(oracle_mod):
    m = tu(L=2, theta=0.5, bs=2)
    o = oracle_mod.OracleTU(m, seed=4)
    o.initial_s()
    s = o.s
    for lt in (1, 4, 7):
        G0 = brute_G(m, s, lt - 1)
        D = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])
        assert np.allclose(o.wrap(G0, D, True, "B"), brute_G(m, s, lt), atol=1e-11)
        assert np.allclose(o.wrap(brute_G(m, s, lt), np.conj(D), False, "B"), G0, atol=1e-11)


def test_ratio_and_update_against_brute_force(oracle_mod):
    """The acceptance weight of get_r! is |det_new/det_old|² γ'/γ, and Gupdate! reproduces
    the recomputed G after the flip (Sherman-Morrison vs from-scratch)."""
    m = tu(L=2, theta=0.5, bs=2)
    o = oracle_mod.OracleTU(m, seed=8)
    o.initial_s()
    s0 = o.s
    lt = 4
    # G at slice lt with the oracle's state machinery: set G = G(lt) and update layer lt
    o.G = brute_G(m, s0, lt)
    from kapdqmc_amd.rng import proposal_draws
    choice, u = proposal_draws(8, 0, 0, 0, lt, 0, m.Ns)
    o.update_layer(lt, 0)
    s1 = o.s
    assert not np.array_equal(s0, s1)  # some flips accepted
    # replay the decisions with brute-force determinants
    s = s0.copy()
    for i in range(m.Ns):
        si = int(s[i, lt - 1])
        allowed = [v for v in (1, 2, 3, 4) if v != si]
        sx = allowed[choice[i]]
        t = s.copy()
        t[i, lt - 1] = sx
        ratio = abs(brute_weight_det(m, t) / brute_weight_det(m, s)) ** 2 * m.γ[sx - 1] / m.γ[si - 1]
        if u[i] < ratio:
            s = t
    assert np.array_equal(s, s1)
    assert np.allclose(o.G, brute_G(m, s1, lt), atol=1e-10)


def test_free_limit_closed_form(oracle_mod):
    """U = 0: every field gives G(τ) = I - e^{-τH}P (Pᴴ e^{-2ΘH} P)⁻¹ Pᴴ e^{-(2Θ-τ)H}."""
    m = tu(L=3, U=0.0, theta=2.0, bs=5)
    o = oracle_mod.OracleTU(m, seed=1)
    o.initial_s()
    w, V = np.linalg.eigh(m.Ht * m.K)
    E = lambda t: (V * np.exp(-t * w)) @ V.T
    P = m.Pt
    for tau in (0, 7, 20, 33):
        ref = np.eye(m.Ns) - E(tau * m.Δt) @ P @ np.linalg.inv(P.T @ E(m.Nt * m.Δt) @ P) @ P.T @ E((m.Nt - tau) * m.Δt)
        assert np.allclose(o.Gtau(tau), ref, atol=1e-10)


def test_free_limit_observables(oracle_mod):
    """U = 0 ground state: Ek = 2 Σ_{occupied} ε (two spin species), Eu = Σ n(1-n) = N/4."""
    m = tu(L=4, U=0.0, theta=3.0, bs=10)
    o = oracle_mod.OracleTU(m, seed=1)
    o.initial_s()
    rows = o.phy_update(1, True)
    eps = np.sort(np.linalg.eigvalsh(m.Ht * m.K))[: m.Ns // 2]
    assert np.isclose(rows[0, 0], 2 * eps.sum(), rtol=1e-6)
    assert np.isclose(rows[0, 1], m.Ns / 4, rtol=1e-6)


def test_sweep_returns_consistent_G(oracle_mod):
    m = tu(L=3, theta=1.0, bs=5)
    o = oracle_mod.OracleTU(m, seed=5)
    o.initial_s()
    o.phy_update(2, False)
    assert o.idx == 1
    assert np.allclose(o.G, brute_G(m, o.s, 0), atol=1e-9)
    assert 0.05 < o.acceptance < 0.95


def test_measure_hand_values(oracle_mod):
    """phy_measure on the smallest honeycomb (site [1,1], N = 2) for a fixed G at slice
    Nt/2 (no wraps). G = a I + b K commutes with K and hence with the HalfeK sandwich.

    For G = [[a, b], [b, a]] (commutes with K = [[0,1],[1,0]] and hence with HalfeK),
    cell pair (a = b = site 1): δ = 1 and the reference accumulates (first lines only)
      tmp1 = (1-a)^2 + (1-a)a + (1-a)^2 + (1-a)a - [(1-a)^2 - b^2] - [(1-a)^2 - b^2]
           = 2(1-a) - 2(1-a)^2 + 2 b^2
      tmp2 = (1-a)a + (1-a)a - (1-a)a - (1-a)a = 0
    Ek = -2 Σ K∘G = -4 b, Eu = 2 (1-a) a, and CDW0 = CDW1 = SDW0 = SDW1 = tmp1 / 4 (cos(2π+2π) = 1).
    With a = 0.3, b = 0.2: tmp1 = 1.4 - 0.98 + 0.08 = 0.5.
    """
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=0.2, Θquench=0.0,
                        Lattice="HoneyComb120", site=[1, 1], BatchSize=1)
    o = oracle_mod.OracleTU(m, seed=0)
    o.initial_s()
    G = np.array([[0.3, 0.2], [0.2, 0.3]], dtype=complex)
    out = o.measure(m.Nt // 2, G)
    assert np.allclose(out, [-0.8, 0.42, 0.125, 0.125, 0.125, 0.125], atol=1e-13)
