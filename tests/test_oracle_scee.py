"""Pins the t-U swap-replica Rényi-2 oracle (src/tU/SCEE.jl) against brute force:
time-displaced Green's functions from unstabilised products, Grover determinants from
scratch, and every accept/reject decision of a layer replayed with the full weight
|det W1|² |det W2|² Πγ |det g_A|^{2λ} |det g_B|^{2(1-λ)}."""
import numpy as np

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para
from kapdqmc_amd.rng import initial_fields, proposal_draws


def model():
    return tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=0.5, Θquench=0.0,
                           Lattice="HoneyComb120", site=[2, 2], BatchSize=5)


def Bs(m, s, lt):
    D = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])
    return D[:, None] * m.eK


def Bprod(m, s, t2, t1):
    X = np.eye(m.Ns, dtype=complex)
    for lt in range(t1 + 1, t2 + 1):
        X = Bs(m, s, lt) @ X
    return X


def G_eq(m, s, tau):
    R = Bprod(m, s, tau, 0) @ m.Pt
    L = m.Pt.T @ Bprod(m, s, m.Nt, tau)
    return np.eye(m.Ns) - R @ np.linalg.inv(L @ R) @ L


def wdet(m, s):
    return np.linalg.det(m.Pt.T @ Bprod(m, s, m.Nt, 0) @ m.Pt)


def grover_det(G1, G2, idx):
    a, b = G1[np.ix_(idx, idx)], G2[np.ix_(idx, idx)]
    return abs(np.linalg.det(2 * a @ b - a - b + np.eye(len(idx)))) ** 2


def test_initial_G4_and_grover(oracle_mod):
    m = model()
    iA = area_index(m.Lattice, m.site, ([1, 1], [1, 2]))
    iB = area_index(m.Lattice, m.site, ([1, 1], [1, 1]))
    s1, s2 = initial_fields(3, 0, m.Ns, m.Nt), initial_fields(4, 0, m.Ns, m.Nt)
    e = oracle_mod.OracleSCEE(m, iA, iB, 0.5, 2, s1, s2, seed=3)
    e.init()
    th = m.Nt // 2
    for r, s in ((0, s1), (1, s2)):
        Gt, G0, Gt0, G0t = e.G4(r)
        g0, gt = G_eq(m, s, 0), G_eq(m, s, th)
        Bt = Bprod(m, s, th, 0)
        assert np.allclose(Gt, g0, atol=1e-10)
        assert np.allclose(G0, gt, atol=1e-10)
        assert np.allclose(G0t, Bt @ g0, atol=1e-9)                          # G(Θ, 0)
        assert np.allclose(Gt0, -(np.eye(m.Ns) - g0) @ np.linalg.inv(Bt), atol=1e-9)  # G(0, Θ)
    a, b = iA - 1, iB - 1
    assert np.isclose(e.detg(0), grover_det(G_eq(m, s1, th), G_eq(m, s2, th), a), rtol=1e-9)
    assert np.isclose(e.detg(1), grover_det(G_eq(m, s1, th), G_eq(m, s2, th), b), rtol=1e-9)


def test_layer_decisions_replayed(oracle_mod):
    m = model()
    lam = 0.5
    iA = area_index(m.Lattice, m.site, ([1, 1], [1, 2]))
    iB = area_index(m.Lattice, m.site, ([1, 1], [1, 1]))
    a, b = iA - 1, iB - 1
    s1, s2 = initial_fields(7, 0, m.Ns, m.Nt), initial_fields(8, 0, m.Ns, m.Nt)
    e = oracle_mod.OracleSCEE(m, iA, iB, lam, 2, s1, s2, seed=9)
    e.init()
    lt, th = 1, m.Nt // 2
    e.forward_slice(lt, 0)
    f1, f2 = e.fields(0), e.fields(1)
    c1, u1 = proposal_draws(9, 0, 1, 0, lt, 0, m.Ns)
    c2, u2 = proposal_draws(9, 0, 2, 0, lt, 0, m.Ns)

    def weight(x1, x2):
        GA, GB = G_eq(m, x1, th), G_eq(m, x2, th)
        return (abs(wdet(m, x1)) ** 2 * abs(wdet(m, x2)) ** 2 * grover_det(GA, GB, a) ** lam
                * grover_det(GA, GB, b) ** (1 - lam))

    x1, x2 = s1.copy(), s2.copy()
    for i in range(m.Ns):
        for r, (ch, u) in enumerate(((c1, u1), (c2, u2))):
            cur = x1 if r == 0 else x2
            si = int(cur[i, lt - 1])
            sx = [v for v in (1, 2, 3, 4) if v != si][ch[i]]
            t = cur.copy()
            t[i, lt - 1] = sx
            new = weight(t, x2) if r == 0 else weight(x1, t)
            ratio = new / weight(x1, x2) * m.γ[sx - 1] / m.γ[si - 1]
            if u[i] < ratio:
                if r == 0:
                    x1 = t
                else:
                    x2 = t
    assert np.array_equal(x1, f1) and np.array_equal(x2, f2)
    assert not (np.array_equal(x1, s1) and np.array_equal(x2, s2))
    # G(τ=1), G(Θ) of both replicas and the Grover determinants after the updates
    for r, s in ((0, x1), (1, x2)):
        Gt, G0, _, _ = e.G4(r)
        assert np.allclose(Gt, G_eq(m, s, lt), atol=1e-9)
        assert np.allclose(G0, G_eq(m, s, th), atol=1e-9)
    assert np.isclose(e.detg(0), grover_det(G_eq(m, x1, th), G_eq(m, x2, th), a), rtol=1e-8)
    assert np.isclose(e.detg(1), grover_det(G_eq(m, x1, th), G_eq(m, x2, th), b), rtol=1e-8)
