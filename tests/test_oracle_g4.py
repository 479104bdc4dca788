"""Pins the oracle's time-displaced Green's functions (Gτ, G4: src/tU/GreenMatrix.jl:90-226)
against unstabilised products on a tiny lattice: G(τ) = I - R (L R)⁻¹ L with R = B(τ,0)P and
L = Pᴴ B(Nt,τ); for τ1 > τ2, G(τ1,τ2) = B(τ1,τ2) G(τ2) and G(τ2,τ1) = -(I - G(τ2)) B(τ1,τ2)⁻¹;
the swapped and equal-time cases; a lattice size where the blocks do not divide the range."""
import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para
from kapdqmc_amd.rng import initial_fields


def model(L=2, theta=1.0, bs=5):
    return tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=theta, Θquench=0.0,
                           Lattice="HoneyComb120", site=[L, L], BatchSize=bs)


def B(m, s, lt):
    D = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])
    return D[:, None] * m.eK


def Bprod(m, s, t2, t1):
    X = np.eye(m.Ns, dtype=complex)
    for lt in range(t1 + 1, t2 + 1):
        X = B(m, s, lt) @ X
    return X


def G_eq(m, s, tau):
    R = Bprod(m, s, tau, 0) @ m.Pt
    L = m.Pt.T @ Bprod(m, s, m.Nt, tau)
    return np.eye(m.Ns) - R @ np.linalg.inv(L @ R) @ L


@pytest.mark.parametrize("t1,t2", [(13, 2), (20, 0), (7, 6), (3, 17), (9, 9), (0, 20)])
def test_g4_against_products(oracle_mod, t1, t2):
    m = model(bs=4)   # Nt = 20 slices: blocks of 4 that do not always divide τ1 - τ2
    s = initial_fields(5, 0, m.Ns, m.Nt)
    G1, G2, G12, G21 = oracle_mod.tu_G4(m, s, t1, t2)
    g1, g2 = G_eq(m, s, t1), G_eq(m, s, t2)
    assert np.allclose(G1, g1, atol=1e-11) and np.allclose(G2, g2, atol=1e-11)
    I = np.eye(m.Ns)
    if t1 > t2:
        Bt = Bprod(m, s, t1, t2)
        want12, want21 = Bt @ g2, -(I - g2) @ np.linalg.inv(Bt)
    elif t1 < t2:
        Bt = Bprod(m, s, t2, t1)
        want12, want21 = -(I - g1) @ np.linalg.inv(Bt), Bt @ g1
    else:
        want12, want21 = -(I - g1), g1
    assert np.allclose(G12, want12, atol=1e-10)
    assert np.allclose(G21, want21, atol=1e-10)
    if t1 > t2:   # a plausible slip, the hole propagator B(τ1,τ2)(I - G(τ2)), is far off
        assert not np.allclose(G12, Bprod(m, s, t1, t2) @ (I - g2), atol=1e-3)


def test_gtau_matches_products_at_every_slice(oracle_mod):
    m = model(bs=3)
    s = initial_fields(6, 0, m.Ns, m.Nt)
    for tau in range(0, m.Nt + 1, 3):
        assert np.allclose(oracle_mod.tu_Gtau(m, s, tau), G_eq(m, s, tau), atol=1e-11)
