"""Pins the t-U-V oracle (src/tUV) against brute-force products and determinant ratios."""
import numpy as np
import pytest

from kapdqmc_amd.models import tUV_Hubbard_Para


def tuv(U=4.0, V=0.5, L=2, theta=0.5, bs=2):
    return tUV_Hubbard_Para(Ht=1.0, Hu=U, Hv=V, Lattice="HoneyComb120", site=[L, L], Δt=0.1, Θ=theta,
                            BatchSize=bs)


def diag(m, s, lt):
    acc = np.zeros(m.Ns)
    bonds = np.transpose(m.nnidx, (1, 0, 2)).reshape(-1, 2) - 1
    for i, (x, y) in enumerate(bonds):
        e = m.η[s[i, lt - 1] - 1]
        acc[x] += e
        acc[y] += m.a * e
    return np.exp(1j * acc) if m.Type is complex else np.exp(acc)


def det_weight(m, s):
    X = m.Pt.astype(complex)
    for lt in range(1, m.Nt + 1):
        X = diag(m, s, lt)[:, None] * (m.eK @ X)
    return np.linalg.det(m.Pt.T @ X)


def brute_G(m, s, tau):
    R = m.Pt.astype(complex)
    for lt in range(1, tau + 1):
        R = diag(m, s, lt)[:, None] * (m.eK @ R)
    Lm = m.Pt.T.astype(complex)
    for lt in range(m.Nt, tau, -1):
        Lm = (Lm * diag(m, s, lt)[None, :]) @ m.eK
    return np.eye(m.Ns) - R @ np.linalg.inv(Lm @ R) @ Lm


@pytest.mark.parametrize("U", [4.0, -4.0])
def test_sweep_G(oracle_mod, U):
    m = tuv(U=U)
    o = oracle_mod.OracleTUV(m, seed=4)
    o.initial_s()
    o.phy_update(2, False)
    assert o.idx == 1
    assert np.allclose(o.G, brute_G(m, o.s, 0), atol=1e-9)


@pytest.mark.parametrize("U", [4.0, -4.0])
def test_rank2_ratio_and_update(oracle_mod, U):
    """Decisions replayed with |det ratio|² γ'/γ (× e^{(η-η')(a+1)} in the real case,
    src/tUV/phy_update.jl:195-197); G after the layer vs recomputed."""
    m = tuv(U=U)
    o = oracle_mod.OracleTUV(m, seed=6)
    o.initial_s()
    s0 = o.s
    lt = 3
    o.G = brute_G(m, s0, lt)
    from kapdqmc_amd.rng import proposal_draws
    nbd = s0.shape[0]
    choice, u = proposal_draws(6, 0, 0, 0, lt, 0, nbd)
    o.update_layer(lt, 0)
    s1 = o.s
    s = s0.copy()
    for i in range(nbd):
        si = int(s[i, lt - 1])
        sx = [v for v in (1, 2, 3, 4) if v != si][choice[i]]
        t = s.copy()
        t[i, lt - 1] = sx
        ratio = abs(det_weight(m, t) / det_weight(m, s)) ** 2 * m.γ[sx - 1] / m.γ[si - 1]
        if m.Type is float:
            ratio *= np.exp((m.η[si - 1] - m.η[sx - 1]) * (m.a + 1))
        if u[i] < ratio:
            s = t
    assert np.array_equal(s, s1)
    assert not np.array_equal(s0, s1)
    assert np.allclose(o.G, brute_G(m, s1, lt), atol=1e-9)


def test_weak_coupling_kinetic_energy(oracle_mod):
    """U → 0, V = 0: Ek = 2 Σ_{occupied} ε of -Ht K (two spin species)."""
    m = tuv(U=1e-8, V=0.0, L=4, theta=3.0, bs=10)   # L = 4: no Dirac zero modes
    o = oracle_mod.OracleTUV(m, seed=1)
    o.initial_s()
    rows = o.phy_update(1, True)
    eps = np.linalg.eigvalsh(-m.Ht * m.K)[: m.Ns // 2].sum()
    assert np.isclose(rows[0, 0], 2 * eps, rtol=1e-5)
    assert np.isclose(rows[0, 1] / m.U, m.Ns / 4, rtol=1e-5)   # Σ n(1-n) at half filling
