"""Pins the t-U disorder-operator oracle (ctrl_SCDOPicr, src/tU/disorder_operate.jl) against
brute force: G(τ), G(Θ) from unstabilised products, |det X|² with X = e^{iα}I + (1-e^{iα})G(Θ)[A]
from scratch, every accept/reject decision of a layer replayed with the full weight
|det W|² Πγ |det X_A|^{2λ} |det X_B|^{2(1-λ)}, the α = 0 limit (X = I, O = 1) and the
free-fermion limit U = 0, where every field gives the same G and O is a closed form."""
import numpy as np

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para
from kapdqmc_amd.rng import initial_fields, proposal_draws


def model(U=4.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=0.5, Θquench=0.0,
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


def x_det(G, idx, alpha):
    if len(idx) == 0:
        return 1.0
    e = np.exp(1j * alpha)
    X = e * np.eye(len(idx)) + (1 - e) * G[np.ix_(idx, idx)]
    return abs(np.linalg.det(X)) ** 2


def regions(m):
    iA = area_index(m.Lattice, m.site, ([1, 1], [1, 2]))
    iB = area_index(m.Lattice, m.site, ([1, 1], [1, 1]))
    return iA, iB


def test_dop_initial_G4_and_x(oracle_mod):
    m = model()
    iA, iB = regions(m)
    alpha = 0.7
    s = initial_fields(3, 0, m.Ns, m.Nt)
    e = oracle_mod.OracleDOP(m, alpha, iA, iB, 0.5, 2, s, seed=3)
    e.init()
    th = m.Nt // 2
    Gt, G0, Gt0, G0t = e.G4(0)
    g0, gt = G_eq(m, s, 0), G_eq(m, s, th)
    Bt = Bprod(m, s, th, 0)
    assert np.allclose(Gt, g0, atol=1e-10)
    assert np.allclose(G0, gt, atol=1e-10)
    assert np.allclose(G0t, Bt @ g0, atol=1e-9)
    assert np.allclose(Gt0, -(np.eye(m.Ns) - g0) @ np.linalg.inv(Bt), atol=1e-9)
    assert np.isclose(e.detg(0), x_det(gt, iA - 1, alpha), rtol=1e-10)
    assert np.isclose(e.detg(1), x_det(gt, iB - 1, alpha), rtol=1e-10)
    # a plausible slip, X = e^{iα}I + (1 - e^{iα})(I - G), is far off
    assert not np.isclose(e.detg(0), x_det(np.eye(m.Ns) - gt, iA - 1, alpha), rtol=1e-3)


def test_dop_layer_decisions_replayed(oracle_mod):
    m = model()
    lam, alpha = 0.5, 1.1
    iA, iB = regions(m)
    a, b = iA - 1, iB - 1
    s0 = initial_fields(7, 0, m.Ns, m.Nt)
    e = oracle_mod.OracleDOP(m, alpha, iA, iB, lam, 2, s0, seed=9)
    e.init()
    lt, th = 1, m.Nt // 2
    e.forward_slice(lt, 0)
    f = e.fields(0)
    ch, u = proposal_draws(9, 0, 1, 0, lt, 0, m.Ns)   # Philox stream 1

    def weight(x):
        G = G_eq(m, x, th)
        return abs(wdet(m, x)) ** 2 * x_det(G, a, alpha) ** lam * x_det(G, b, alpha) ** (1 - lam)

    x = s0.copy()
    for i in range(m.Ns):
        si = int(x[i, lt - 1])
        sx = [v for v in (1, 2, 3, 4) if v != si][ch[i]]
        t = x.copy()
        t[i, lt - 1] = sx
        ratio = weight(t) / weight(x) * m.γ[sx - 1] / m.γ[si - 1]
        if u[i] < ratio:
            x = t
    assert np.array_equal(x, f)
    assert not np.array_equal(x, s0)
    Gt, G0, _, _ = e.G4(0)
    assert np.allclose(Gt, G_eq(m, x, lt), atol=1e-9)
    assert np.allclose(G0, G_eq(m, x, th), atol=1e-9)
    # the tracked |det X|² (products of |Tau|²) against the determinant from scratch
    assert np.isclose(e.detg(0), x_det(G_eq(m, x, th), a, alpha), rtol=1e-8)
    assert np.isclose(e.detg(1), x_det(G_eq(m, x, th), b, alpha), rtol=1e-8)


def test_dop_alpha_zero_is_identity(oracle_mod):
    m = model()
    iA, iB = regions(m)
    s = initial_fields(4, 0, m.Ns, m.Nt)
    e = oracle_mod.OracleDOP(m, 0.0, iA, iB, 0.5, 2, s, seed=4)
    O = e.sweeps(2)
    assert np.allclose(O, 1.0, rtol=0, atol=1e-12)
    assert abs(e.detg(0) - 1.0) < 1e-12 and abs(e.detg(1) - 1.0) < 1e-12


def test_dop_free_limit_closed_form(oracle_mod):
    # U = 0: α_l = 0, every field gives the free projector G(Θ), so every slice adds
    # (|det X_A|² / |det X_B|²)^(1/Nλ) of that G and O equals it at every sweep
    m = model(U=0.0)
    iA, iB = regions(m)
    alpha, nlam = 0.9, 3
    s = initial_fields(5, 0, m.Ns, m.Nt)
    e = oracle_mod.OracleDOP(m, alpha, iA, iB, 0.5, nlam, s, seed=5)
    O = e.sweeps(2)
    th = m.Nt // 2
    Bh = np.linalg.matrix_power(m.eK, th)
    R = Bh @ m.Pt
    L = m.Pt.T @ np.linalg.matrix_power(m.eK, m.Nt - th)
    G = np.eye(m.Ns) - R @ np.linalg.inv(L @ R) @ L
    want = (x_det(G, iA - 1, alpha) / x_det(G, iB - 1, alpha)) ** (1.0 / nlam)
    assert np.allclose(O, want, rtol=1e-10)
    assert abs(want - 1.0) > 1e-3   # the closed form is not trivially 1
