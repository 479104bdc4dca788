"""Pins the Philox4x32-10 stream shared by the oracle and the HIP library."""
import numpy as np

from kapdqmc_amd import rng


# Known-answer vectors published with Random123 (kat_vectors, philox4x32 10 rounds).
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def test_philox_kat_oracle(oracle_mod):
    for ctr, key, out in KAT:
        assert tuple(oracle_mod.philox(ctr, key)) == out


def test_philox_kat_host():
    for ctr, key, out in KAT:
        got = rng.philox4x32_10(*[np.uint32(x) for x in ctr], np.uint32(key[0]), np.uint32(key[1]))
        assert tuple(int(x) for x in got) == out


def test_draws_host_equal_oracle(oracle_mod):
    seed, chain, sweep = 0x1234_5678_9ABC, 5, 3
    for lt, d in [(1, 0), (17, 1), (200, 0)]:
        ch, u = rng.proposal_draws(seed, chain, 0, sweep, lt, d, 40)
        for i in range(40):
            oc, ou = oracle_mod.draw(seed, chain, 0, sweep, lt | (d << 31), i)
            assert oc == ch[i] and ou == u[i]


def test_draw_distribution():
    ch, u = rng.proposal_draws(7, 0, 0, 0, 1, 0, 200000)
    counts = np.bincount(ch, minlength=3)
    assert set(np.unique(ch)) == {0, 1, 2}
    chi2 = ((counts - 200000 / 3) ** 2 / (200000 / 3)).sum()
    assert chi2 < 20          # 2 dof, p ~ 5e-5
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.005


def test_initial_fields_host_equal_oracle(oracle_mod):
    from kapdqmc_amd.models import tU_Hubbard_Para
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                        Lattice="HoneyComb120", site=[3, 3], BatchSize=5)
    o = oracle_mod.OracleTU(m, seed=99, chain=4)
    o.initial_s()
    s = rng.initial_fields(99, 4, m.Ns, m.Nt)
    assert s.dtype == np.uint8 and s.shape == (m.Ns, m.Nt)
    assert np.array_equal(s, o.s)
    assert set(np.unique(s)) == {1, 2, 3, 4}
