"""Failure detection through the C ABI: a chain whose G turns non-finite is flagged and named
while the other chains of its batch finish unaffected (their fields and rows equal the
oracle's); a rank-deficient trial state gives the singular-overlap error; the wrap error at the
stabilisation nodes (the reference's commented check, src/tU/phy_update.jl:54-58,92-97) is
reported per chain; the projector check of G sees a planted perturbation of the right size.
The failures are forced from the host through the model constants and G; every kernel still
runs to its end."""
import ctypes as C

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para

pytestmark = pytest.mark.gpu


def tu(U=4.0):
    return tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[3, 3], BatchSize=5)


def test_poisoned_chain_is_isolated(oracle_mod):
    print("case: poisoned chain in a batch of 3", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch
    ms = [tu(3.0), tu(4.0), tu(5.0)]
    ms[1].α = ms[1].α.copy()
    ms[1].α[7] = np.nan                                 # a non-finite coupling on slice 8 of chain 1
    os_ = []
    for c, m in enumerate(ms):
        o = oracle_mod.OracleTU(m, seed=60 + c, chain=c)
        o.initial_s()
        os_.append(o)
    b = DQMCBatch(ms, [60, 61, 62], [0, 1, 2], delay=8)
    for c in range(3):
        b.set_fields(c, os_[c].s)
    rows = np.zeros((3, 2, 6))
    rc = b.L.kdq_phy_update(b.h, 2, 1, rows.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == -3
    assert b"chain 1" in b.L.kdq_last_error(b.h) and b"non-finite" in b.L.kdq_last_error(b.h)
    hl = b.health()
    assert hl[1, 1] == 1.0 and hl[0, 1] == 0.0 and hl[2, 1] == 0.0
    for c in (0, 2):
        ro = os_[c].phy_update(2, True)
        assert np.array_equal(b.get_fields(c), os_[c].s), c
        assert np.allclose(rows[c], ro, rtol=1e-9, atol=1e-11), c
        assert 0.0 < hl[c, 0] < 1e-8                     # wrap error of a healthy chain: rounding
    b.set_fields(1, os_[1].s)                           # new fields clear the flags
    assert b.health()[1, 1] == 0.0
    b.close()


def test_singular_overlap_is_reported():
    print("case: rank-deficient trial state", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch, Initial_s
    good, bad = tu(), tu()
    bad.Pt = bad.Pt.copy()
    bad.Pt[:, 2] = 0.0                                  # P loses rank: U_L U_R is singular
    b = DQMCBatch([good, bad], [1, 2], [0, 1], delay=8)
    for c, m in enumerate((good, bad)):
        b.set_fields(c, Initial_s(m, 1 + c, c))
    rc = b.L.kdq_init_chain(b.h)
    assert rc == -3
    msg = b.L.kdq_last_error(b.h)
    assert b"chain 1" in msg and (b"singular" in msg or b"rank" in msg)
    b.close()


def test_projector_check_sees_a_perturbation():
    print("case: projector check", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch, Initial_s
    m = tu()
    b = DQMCBatch([m, m], [5, 6], [0, 1], delay=8)
    for c in range(2):
        b.set_fields(c, Initial_s(m, 5 + c, c))
    b.init_chain()
    ok = b.check_G()
    assert np.all(ok[:, 0] < 1e-10) and np.all(ok[:, 1] < 1e-10)
    G = b.get_G(1)
    rng = np.random.default_rng(0)
    P = rng.normal(size=G.shape) * 1e-4
    b.set_G(1, G + P)
    chk = b.check_G()
    assert chk[0, 0] < 1e-10
    # (G + P)² - (G + P) = G P + P G - P + O(P²): of the order of |P|
    assert 1e-6 < chk[1, 0] < 1e-2
    assert chk[1, 1] == pytest.approx(abs(np.trace(P)), rel=1e-6, abs=1e-12)
    b.close()
