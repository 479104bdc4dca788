"""Complex hopping (flux != 0) through the GPU path against the FP64 oracle (C ABI): the
reference's π-flux SQUARE90 workload (`test/tU/□tU_runtests.jl:44-57`: 6x6 cells, U = 4,
Θrelax = 1, BatchSize 5, `ctrl_SCEEicr` 2 sweeps with λ = 0.5, Nλ = 2, regions
area_index(([1,1],[6,3])) and area_index(([1,1],[3,3])) from ss = [s, s]) and phy_update
on it, a non-π flux with Sherman-Morrison and delayed blocks, and the exact log-weight
against a brute-force determinant. Bars: fields bit-exact (sweeps: equal near-tie counts),
G within 1e-10 relative, rows and O within 1e-9 relative."""
import numpy as np
import pytest

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def pi_flux():
    return tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Θrelax=1.0, Θquench=0.0, Lattice="SQUARE90",
                           site=[6, 6], Δt=0.1, BatchSize=5, Initial="H0", flux=np.pi)


@pytest.mark.parametrize("delay", [16, 1])
def test_pi_flux_phy_update(oracle_mod, delay):
    print(f"case: pi-flux SQUARE90 6x6 phy_update delay={delay}", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch
    m = pi_flux()
    o = oracle_mod.OracleTU(m, seed=1234)
    o.initial_s()
    b = DQMCBatch([m], [1234], [0], delay=delay)
    b.set_fields(0, o.s)
    ro = o.phy_update(2, True)
    rb = b.phy_update(2, True)[0]
    assert np.array_equal(b.get_fields(0), o.s)
    assert b.near_ties()[0] == o.near_ties
    assert rel(b.get_G(0), o.G) < 1e-10
    assert np.allclose(rb, ro, rtol=1e-9, atol=1e-11)
    G = b.get_G(0)
    assert np.abs(G @ G - G).max() < 1e-10 and abs(np.trace(G) - (m.Ns - m.Ns // 2)) < 1e-10
    b.close()


def test_pi_flux_reference_scee(oracle_mod):
    print("case: pi-flux SQUARE90 6x6 SCEE", flush=True)
    from kapdqmc_amd.dqmc import Initial_s, SCEEBatch
    m = pi_flux()
    L = m.site[1]
    iA = np.asarray(area_index(m.Lattice, m.site, ([1, 1], [L, L // 2])), dtype=np.int64)
    iB = np.asarray(area_index(m.Lattice, m.site, ([1, 1], [L // 2, L // 2])), dtype=np.int64)
    s = Initial_s(m, 1234)
    o = oracle_mod.OracleSCEE(m, iA.astype(np.int32), iB.astype(np.int32), 0.5, 2, s.copy(), s.copy(), seed=1234)
    b = SCEEBatch([m], iA, iB, 0.5, 2, [1234], [0], delay=8)
    b.set_fields(0, 0, s)
    b.set_fields(0, 1, s)
    Oo = o.sweeps(2)
    Ob = b.sweeps(2)[0]
    for r in (0, 1):
        assert np.array_equal(b.get_fields(0, r), o.fields(r)), r
        for name, x, y in zip(("Gt", "G0", "Gt0", "G0t"), b.G4(0, r), o.G4(r)):
            assert rel(x, y) < 1e-10, (r, name, rel(x, y))
    assert np.allclose(Ob, Oo, rtol=1e-9, atol=0)
    b.close()


def test_flux_batch_and_log_weight(oracle_mod):
    """Two chains with flux 0.7 (complex K without the π-flux symmetry) on 3x2 cells, delayed
    blocks; then kdq_log_weight against 2 log|det(Pᴴ B(2Θ,0) P)| + Σ log γ by brute force."""
    print("case: flux 0.7 batch + log weight", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch
    ms = [tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Θrelax=0.5, Θquench=0.0, Lattice="SQUARE90", site=[3, 2],
                          Δt=0.1, BatchSize=5, Initial="H0", flux=0.7) for U in (3.0, 5.0)]
    os_ = []
    for c, m in enumerate(ms):
        o = oracle_mod.OracleTU(m, seed=40 + c, chain=c)
        o.initial_s()
        os_.append(o)
    b = DQMCBatch(ms, [40, 41], [0, 1], delay=8)
    for c in range(2):
        b.set_fields(c, os_[c].s)
    rb = b.phy_update(2, True)
    for c, o in enumerate(os_):
        ro = o.phy_update(2, True)
        assert np.array_equal(b.get_fields(c), o.s)
        assert rel(b.get_G(c), o.G) < 1e-10
        assert np.allclose(rb[c], ro, rtol=1e-9, atol=1e-11)
    lw = b.log_weight()
    for c, m in enumerate(ms):
        s = b.get_fields(c)
        X = m.Pt.astype(complex)
        for lt in range(1, m.Nt + 1):
            X = np.exp(1j * m.α[lt - 1] * m.η[s[:, lt - 1].astype(int) - 1])[:, None] * (m.eK @ X)
        ref = 2 * np.log(abs(np.linalg.det(m.Pt.conj().T @ X))) + np.log(m.γ[s.astype(int) - 1]).sum()
        assert abs(lw[c] - ref) < 1e-9 * max(1.0, abs(ref)), (lw[c], ref)
    b.close()
