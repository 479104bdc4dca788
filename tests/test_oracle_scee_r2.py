"""Pins the t-V and t-U-V swap-replica Rényi-2 oracle (src/tV/SCEE.jl, src/tUV/SCEE.jl)
against brute force: time-displaced Green's functions of both replicas from unstabilised
products, Grover determinants from scratch, and every accept/reject decision of a forward
slice replayed with the full two-replica weight W(s1) W(s2) g_A^λ g_B^{1-λ} Π γ
(t-V: W = det(Pᵀ B P), g = det(Grover); t-U-V: W = |det|², g = |det(Grover)|²)."""
import numpy as np
import pytest

from kapdqmc_amd.dqmc import Initial_s
from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tUV_Hubbard_Para, tV_Hubbard_Para
from kapdqmc_amd.rng import proposal_draws


def tv():
    return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=0.5, Θquench=0.0,
                           Lattice="HoneyComb120", site=[2, 2], BatchSize=5)


def tuv(U=4.0):
    return tUV_Hubbard_Para(Ht=1.0, Hu=U, Hv=0.5, Lattice="HoneyComb120", site=[2, 2], Δt=0.1, Θ=0.5,
                            BatchSize=5)


def B_slice(m, s, lt):
    n = m.Ns
    if m.kind == "tV":
        UV = m.UV()
        X = m.eK.astype(complex)
        for j in (2, 1, 0):
            D = np.ones(n)
            for i in range(m.nnidx.shape[0]):
                x, y = m.nnidx[i, j] - 1
                v = m.α[lt - 1] * m.η[int(s[i, j, lt - 1]) - 1]
                D[x], D[y] = np.exp(v), np.exp(-v)
            X = UV[:, :, j] @ np.diag(D) @ UV[:, :, j] @ X
        return X
    acc = np.zeros(n)
    z = m.nnidx.shape[1]
    for j in range(z):
        for i in range(n):
            x, y = m.nnidx[i, j] - 1
            v = m.η[int(s[i + j * n, lt - 1]) - 1]
            acc[x] += v
            acc[y] += m.a * v
    D = np.exp(1j * acc) if m.Type is complex else np.exp(acc)
    return D[:, None] * m.eK


def Bprod(m, s, t2, t1):
    X = np.eye(m.Ns, dtype=complex)
    for lt in range(t1 + 1, t2 + 1):
        X = B_slice(m, s, lt) @ X
    return X


def G_eq(m, s, tau):
    R = Bprod(m, s, tau, 0) @ m.Pt
    L = m.Pt.conj().T @ Bprod(m, s, m.Nt, tau)
    return np.eye(m.Ns) - R @ np.linalg.inv(L @ R) @ L


def W(m, s):
    d = np.linalg.det(m.Pt.conj().T @ Bprod(m, s, m.Nt, 0) @ m.Pt)
    return d.real if m.kind == "tV" else abs(d) ** 2


def grover(m, G1, G2, idx):
    a, b = G1[np.ix_(idx, idx)], G2[np.ix_(idx, idx)]
    d = np.linalg.det(2 * a @ b - a - b + np.eye(len(idx)))
    return d.real if m.kind == "tV" else abs(d) ** 2


def regions(m):
    return (area_index(m.Lattice, m.site, ([1, 1], [1, 2])), area_index(m.Lattice, m.site, ([1, 1], [1, 1])))


CASES = [("tV", tv), ("tUV+", lambda: tuv(4.0)), ("tUV-", lambda: tuv(-4.0))]


@pytest.mark.parametrize("name,mk", CASES)
def test_initial_G4_and_grover(oracle_mod, name, mk):
    m = mk()
    iA, iB = regions(m)
    s1, s2 = Initial_s(m, 3), Initial_s(m, 4)
    e = oracle_mod.OracleSCEE2(m, iA, iB, 0.5, 2, s1, s2, seed=3)
    e.init()
    th = m.Nt // 2
    for r, s in ((0, s1), (1, s2)):
        Gt, G0, Gt0, G0t = e.G4(r)
        g0, gt = G_eq(m, s, 0), G_eq(m, s, th)
        Bt = Bprod(m, s, th, 0)
        assert np.allclose(Gt, g0, atol=1e-10)
        assert np.allclose(G0, gt, atol=1e-10)
        assert np.allclose(G0t, Bt @ g0, atol=1e-9)
        assert np.allclose(Gt0, -(np.eye(m.Ns) - g0) @ np.linalg.inv(Bt), atol=1e-9)
    G1, G2 = G_eq(m, s1, th), G_eq(m, s2, th)
    assert np.isclose(e.detg(0), grover(m, G1, G2, iA - 1), rtol=1e-9)
    assert np.isclose(e.detg(1), grover(m, G1, G2, iB - 1), rtol=1e-9)


@pytest.mark.parametrize("name,mk", CASES)
def test_slice_decisions_replayed(oracle_mod, name, mk):
    m = mk()
    lam = 0.5
    iA, iB = regions(m)
    a, b = iA - 1, iB - 1
    s1, s2 = Initial_s(m, 7), Initial_s(m, 8)
    e = oracle_mod.OracleSCEE2(m, iA, iB, lam, 2, s1, s2, seed=9)
    e.init()
    lt, th = 1, m.Nt // 2
    e.forward_slice(lt, 0)
    f1, f2 = e.fields(0), e.fields(1)
    nel = int(np.prod(s1.shape[:-1]))
    c1, u1 = proposal_draws(9, 0, 1, 0, lt, 0, nel)
    c2, u2 = proposal_draws(9, 0, 2, 0, lt, 0, nel)

    def weight(x1, x2):
        GA, GB = G_eq(m, x1, th), G_eq(m, x2, th)
        return W(m, x1) * W(m, x2) * grover(m, GA, GB, a) ** lam * grover(m, GA, GB, b) ** (1 - lam)

    if m.kind == "tV":
        nb = m.Ns // 2
        order = [i + j * nb for j in (2, 1, 0) for i in range(nb)]
    else:
        order = list(range(nel))
    x1 = s1.reshape(nel, m.Nt, order="F").copy()
    x2 = s2.reshape(nel, m.Nt, order="F").copy()
    shp = lambda x: x.reshape(s1.shape, order="F")
    accepted = 0
    for el in order:
        for r, (ch, u) in enumerate(((c1, u1), (c2, u2))):
            cur = x1 if r == 0 else x2
            si = int(cur[el, lt - 1])
            sx = [v for v in (1, 2, 3, 4) if v != si][ch[el]]
            t = cur.copy()
            t[el, lt - 1] = sx
            new = weight(shp(t), shp(x2)) if r == 0 else weight(shp(x1), shp(t))
            with np.errstate(invalid="ignore"):
                ratio = new / weight(shp(x1), shp(x2)) * m.γ[sx - 1] / m.γ[si - 1]
            if u[el] < ratio:
                accepted += 1
                if r == 0:
                    x1 = t
                else:
                    x2 = t
    assert accepted > 0
    assert np.array_equal(shp(x1), f1) and np.array_equal(shp(x2), f2)
    for r, s in ((0, shp(x1)), (1, shp(x2))):
        Gt, G0, _, _ = e.G4(r)
        assert np.allclose(Gt, G_eq(m, s, lt), atol=1e-9)
        assert np.allclose(G0, G_eq(m, s, th), atol=1e-9)
    GA, GB = G_eq(m, shp(x1), th), G_eq(m, shp(x2), th)
    assert np.isclose(e.detg(0), grover(m, GA, GB, a), rtol=1e-8)
    assert np.isclose(e.detg(1), grover(m, GA, GB, b), rtol=1e-8)


@pytest.mark.parametrize("name,mk", CASES[:2])
def test_sweeps_run_and_G4_stays_consistent(oracle_mod, name, mk):
    m = mk()
    iA, iB = regions(m)
    s1, s2 = Initial_s(m, 11), Initial_s(m, 12)
    e = oracle_mod.OracleSCEE2(m, iA, iB, 0.5, 2, s1, s2, seed=13)
    O = e.sweeps(2)
    assert np.all(np.isfinite(O))
    # back at τ = 0 after the backward half: G4 equals the from-scratch values
    th = m.Nt // 2
    for r in range(2):
        s = e.fields(r)
        Gt, G0, _, _ = e.G4(r)
        assert np.allclose(Gt, G_eq(m, s, 0), atol=1e-8)
        assert np.allclose(G0, G_eq(m, s, th), atol=1e-8)
