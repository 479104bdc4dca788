"""Exact finite-Δτ expectation values of the t-U projector DQMC on a tiny lattice, computed
in Fock space (test infrastructure: an independent check of the sampled distribution, not a
restatement of the reference's algorithm).

The t-U sweep (`src/tU/phy_update.jl:3-152`) samples HS fields s with weight
|det(Pᵀ B(2Θ,0) P)|² Π γ_s, B_l = D_l eK, D_l = diag(e^{iα η_s}) (`src/tU/GreenMatrix.jl:10-27`,
`get_r!` `:49-54`). |det|² is the weight of the particle-hole transformed problem: spin up with
phases e^{+iαη n↑}, spin down (c↓ → ±c↓† on the two sublattices, hopping unchanged on the
bipartite honeycomb) with e^{-iαη ñ↓}, both with the trial Slater determinant of P. Summing the
fields site by site turns the sampled ensemble into the Fock-space transfer operator

    T = V_HS · e^{-Δτ H_kin},   V_HS = Π_i ¼ Σ_l γ_l cos(α η_l (ñ↑_i − ñ↓_i)),

with no Trotter or HS approximation left: the sampled average of any per-field estimator
equals ⟨ψ_T| T^{Nt−M} e^{+ΔτK/2} Ô e^{−ΔτK/2} T^{M} |ψ_T⟩ / ⟨ψ_T| T^{Nt} |ψ_T⟩ with M = Nt/2
and the HalfeK sandwich of `phy_measure` (`src/tU/phy_update.jl:181-182`). The observables are
the kinetic energy Ek = ⟨Ht Σ_σ Σ_ij K_ij c†_iσ c_jσ⟩ (`:184`) and Eu = Σ_i ⟨n_i↑ n_i↓⟩
(`:185-190`), in the original basis n_i↓ = 1 − ñ_i↓.

The many-body state is a matrix ψ[a, b] over (spin-up configuration a) x (spin-down
configuration b), each a set of Np occupied sites in ascending (Jordan-Wigner) order; a
one-species operator X acts as X ψ on the up side and ψ Xᵀ on the down side.

The Rényi-2 entropy S2(A) = −log tr ρ_A² of the exact ground state checks the swap-replica
estimator (`ctrl_SCEEicr`, `src/tU/SCEE.jl:1-262`); the particle-hole transformation is a
product of single-site unitaries and leaves it unchanged.
"""
import itertools

import numpy as np
import scipy.linalg as la


def _configurations(N: int, Np: int):
    confs = list(itertools.combinations(range(N), Np))
    occ = np.zeros((len(confs), N))
    for a, c in enumerate(confs):
        occ[a, list(c)] = 1.0
    return confs, occ


def hopping_fock(K: np.ndarray, Ht: float, Np: int) -> np.ndarray:
    """Ht Σ_ij K_ij c†_i c_j on the Np-particle space of one species (JW signs)."""
    N = K.shape[0]
    confs, _ = _configurations(N, Np)
    index = {c: a for a, c in enumerate(confs)}
    T = np.zeros((len(confs), len(confs)))
    for a, c in enumerate(confs):
        occ = set(c)
        for i in range(N):
            for j in range(N):
                if i == j or K[i, j] == 0 or j not in occ or i in occ:
                    continue
                rest = sorted(occ - {j})
                sign = (-1) ** (sum(x < j for x in rest) + sum(x < i for x in rest))
                T[index[tuple(sorted(rest + [i]))], a] += Ht * K[i, j] * sign
    return T


def slater(P: np.ndarray) -> np.ndarray:
    """Amplitudes ⟨a|ψ_P⟩ = det P[a, :] of the Slater determinant of the columns of P."""
    confs, _ = _configurations(P.shape[0], P.shape[1])
    return np.array([np.linalg.det(P[list(c), :]) for c in confs])


def ground_state(K: np.ndarray, Ht: float, U: float):
    """Lowest eigenpair of Ht K (both spins) + U/2 Σ_i (n_i − 1)² at half filling (exact
    diagonalisation); returns (E0, Ek, Eu) with Eu = Σ_i ⟨n_i↑ n_i↓⟩."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    N = K.shape[0]
    Np = N // 2
    _, occ = _configurations(N, Np)
    T = sp.csr_matrix(hopping_fock(K, Ht, Np))
    one = sp.identity(T.shape[0], format="csr")
    Hk = sp.kron(T, one) + sp.kron(one, T)
    nup = np.repeat(occ, len(occ), axis=0)
    ndn = np.tile(occ, (len(occ), 1))
    H = Hk + sp.diags(0.5 * U * ((nup + ndn - 1.0) ** 2).sum(1))
    w, v = sla.eigsh(H, k=1, which="SA")
    g = v[:, 0]
    return float(w[0]), float(g @ (Hk @ g)), float((g * g) @ (nup * ndn).sum(1))


def transfer_diagonal(occ: np.ndarray, alpha: float, eta, gamma) -> np.ndarray:
    """V_HS[a, b] = Π_i ¼ Σ_l γ_l cos(α η_l (ñ↑_i − ñ↓_i)) over up/down configurations."""
    x = occ[:, None, :] - occ[None, :, :]
    return sum(0.25 * g * np.cos(alpha * e * x) for g, e in zip(gamma, eta)).prod(axis=2)


def field_weight(m, s: np.ndarray) -> float:
    """⟨ψ_T| Π_l D↑_l(s) E ⊗ D↓_l(s) E |ψ_T⟩ for one field s (no sum over fields): equals
    |det(Pᵀ B(2Θ,0) P)|² of the DQMC weight, which pins the particle-hole picture above."""
    N, Np = m.Ns, m.Ns // 2
    _, occ = _configurations(N, Np)
    E = la.expm(-m.Δt * hopping_fock(m.K, m.Ht, Np))
    amp = slater(m.Pt)
    psi = np.outer(amp, amp).astype(np.complex128)
    for lt in range(m.Nt):
        ph = m.α[lt] * np.asarray(m.η)[s[:, lt] - 1]           # per-site phase α η_s
        up = np.exp(1j * occ @ ph)
        dn = np.exp(-1j * occ @ ph)
        psi = up[:, None] * (E @ psi @ E.T) * dn[None, :]
    return float(np.real((np.outer(amp, amp) * psi).sum()))


def staggered_correlations(m, occ: np.ndarray):
    """Diagonal operators (over up configuration a, down configuration b of the
    particle-hole picture) of the four structure factors of `phy_measure`
    (`src/tU/phy_update.jl:194-235`) as they accumulate there: with the staggered density
    m_σ(c) = n_σ(c, A) − n_σ(c, B) of cell c (site 2c of the cell is A, 2c+1 is B,
    c = x + L1 y), the per-field Wick estimators sum to

        tmp1(r) = Σ_c m↑(c) m↑(c + r),   tmp2(r) = Σ_c m↑(c) m↓(c + r)

    (the `+adjoint(` continuation lines that would add the ↓↓ and ↓↑ halves are separate,
    discarded statements); CDW0 = Σ_r (tmp1 + tmp2) / N², CDW1 = Σ_r cos(q·r)(tmp1 + tmp2) / N²
    with q·r = 2π rx / L1 + 2π ry / L2, SDW0 and SDW1 the same with tmp1 − tmp2.
    Returns an array [4, n_up, n_down]."""
    L1, L2 = m.site
    N = m.Ns
    cells = [(x, y) for y in range(L2) for x in range(L1)]
    mu = np.stack([occ[:, 2 * k] - occ[:, 2 * k + 1] for k in range(len(cells))], axis=1)
    md = -mu                                              # n↓ = 1 − ñ↓
    out = np.zeros((4,) + (len(occ), len(occ)))
    for rx in range(1, L1 + 1):
        for ry in range(1, L2 + 1):
            t1 = np.zeros(len(occ))
            t2 = np.zeros((len(occ), len(occ)))
            for k, (x, y) in enumerate(cells):
                kk = ((x + rx) % L1) + L1 * ((y + ry) % L2)
                t1 += mu[:, k] * mu[:, kk]
                t2 += np.outer(mu[:, k], md[:, kk])
            cq = np.cos(2 * np.pi / L1 * rx + 2 * np.pi / L2 * ry)
            out[0] += t1[:, None] + t2
            out[1] += cq * (t1[:, None] + t2)
            out[2] += t1[:, None] - t2
            out[3] += cq * (t1[:, None] - t2)
    return out / float(N) ** 2


def expectation(m, structure_factors: bool = False):
    """Exact (Ek, Eu) of the distribution the t-U sweep samples, measured at slice Nt/2;
    with structure_factors, (Ek, Eu, CDW0, CDW1, SDW0, SDW1)."""
    N, Np = m.Ns, m.Ns // 2
    if not np.allclose(m.α, m.α[0]):
        raise ValueError("constant coupling profiles only")
    _, occ = _configurations(N, Np)
    T = hopping_fock(m.K, m.Ht, Np)
    E = la.expm(-m.Δt * T)
    Eh, Ehi = la.expm(-0.5 * m.Δt * T), la.expm(0.5 * m.Δt * T)
    V = transfer_diagonal(occ, m.α[0], m.η, m.γ)
    amp = slater(m.Pt)
    R = np.outer(amp, amp)
    Lv = R.copy()
    M = m.Nt // 2
    for _ in range(M):
        R = V * (E @ R @ E.T)
    for _ in range(m.Nt - M):
        Lv = E.T @ (V * Lv) @ E
    R = Eh @ R @ Eh.T
    Lv = Ehi.T @ Lv @ Ehi
    norm = (Lv * R).sum()
    Ek = (Lv * (T @ R + R @ T.T)).sum() / norm
    dop = (occ[:, None, :] * (1.0 - occ[None, :, :])).sum(2)       # n↑ (1 − ñ↓) summed over sites
    Eu = (Lv * R * dop).sum() / norm
    if not structure_factors:
        return float(Ek), float(Eu)
    sf = [(Lv * R * op).sum() / norm for op in staggered_correlations(m, occ)]
    return (float(Ek), float(Eu)) + tuple(float(x) for x in sf)


def ground_state_vector(K: np.ndarray, Ht: float, U: float):
    """(configurations, ψ[up configuration, down configuration]) of the half-filled ground
    state of Ht K + U/2 Σ_i (n_i − 1)² (exact diagonalisation)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as sla
    N = K.shape[0]
    Np = N // 2
    confs, occ = _configurations(N, Np)
    T = sp.csr_matrix(hopping_fock(K, Ht, Np))
    one = sp.identity(T.shape[0], format="csr")
    nup = np.repeat(occ, len(occ), axis=0)
    ndn = np.tile(occ, (len(occ), 1))
    H = sp.kron(T, one) + sp.kron(one, T) + sp.diags(0.5 * U * ((nup + ndn - 1.0) ** 2).sum(1))
    _, v = sla.eigsh(H, k=1, which="SA")
    return confs, v[:, 0].reshape(len(confs), len(confs))


def _bipartite(confs, psi: np.ndarray, A: set) -> dict:
    """ψ as a map ((A↑, A↓) part, (B↑, B↓) part) -> amplitude. Modes are ordered (all ↑
    sites, all ↓ sites); bringing A's modes to the front, (A↑, A↓, B↑, B↓), costs the parity
    of the occupied (B before A) pairs within each spin block and of moving the occupied A↓
    modes past the occupied B↑ modes."""
    M = {}
    for a, cu in enumerate(confs):
        for b, cd in enumerate(confs):
            v = psi[a, b]
            if v == 0.0:
                continue
            parts = []
            for c in (cu, cd):
                inA = tuple(x for x in c if x in A)
                inB = tuple(x for x in c if x not in A)
                parts.append((inA, inB, sum(1 for i in inB for j in inA if i < j)))
            (aU, bU, iu), (aD, bD, idn) = parts
            key = ((aU, aD), (bU, bD))
            M[key] = M.get(key, 0.0) + (-1) ** (iu + idn + len(bU) * len(aD)) * v
    return M


def renyi2(confs, psi: np.ndarray, sites, left: np.ndarray = None) -> float:
    """S2(A) = −log tr ρ_A² for the set A of (0-based) sites, both spins: ρ = |ψ⟩⟨ψ| (ψ
    normalised), or ρ = |ψ⟩⟨left| / ⟨left|ψ⟩ when a left vector is given."""
    A = set(int(x) for x in sites)
    MR = _bipartite(confs, psi, A)
    ML = MR if left is None else _bipartite(confs, left, A)
    keys = set(MR) | set(ML)
    rows = {k: i for i, k in enumerate(sorted({k[0] for k in keys}))}
    cols = {k: i for i, k in enumerate(sorted({k[1] for k in keys}))}
    XR, XL = np.zeros((len(rows), len(cols))), np.zeros((len(rows), len(cols)))
    for X, M in ((XR, MR), (XL, ML)):
        for (r, c), v in M.items():
            X[rows[r], cols[c]] = v
    norm = (psi * psi).sum() if left is None else (psi * left).sum()
    rho = XR @ XL.T / norm
    return float(-np.log(np.trace(rho @ rho)))


def renyi2_sampled(m, sites) -> float:
    """Exact S2(A) of what the swap-replica estimator averages at finite Δτ: ρ = |R⟩⟨L| /
    ⟨L|R⟩ with R = T^{Nt/2} |ψ_T⟩ and ⟨L| = ⟨ψ_T| T^{Nt/2}, T = V_HS e^{-Δτ H_kin}, at the
    node Nt/2 where G(Θ) enters the Grover matrices (no HalfeK sandwich)."""
    N, Np = m.Ns, m.Ns // 2
    if not np.allclose(m.α, m.α[0]):
        raise ValueError("constant coupling profiles only")
    confs, occ = _configurations(N, Np)
    E = la.expm(-m.Δt * hopping_fock(m.K, m.Ht, Np))
    V = transfer_diagonal(occ, m.α[0], m.η, m.γ)
    amp = slater(m.Pt)
    R = np.outer(amp, amp)
    L = R.copy()
    for _ in range(m.Nt // 2):
        R = V * (E @ R @ E.T)
    for _ in range(m.Nt - m.Nt // 2):
        L = E.T @ (V * L) @ E
    return renyi2(confs, R, sites, left=L)
