"""Replayed proposal draws (the replay-buffer random stream of the C ABI, kdq_set_replay) in
the oracle: the layout [sweep, pass, slice, element], the meaning of a draw (choice = index
among the other three field values in increasing order, accept iff u < p, as
`rand(rng, model.samplers_dict[s[i]])` and `rand(rng) < p` in UpdatePhyLayer!,
src/tU/phy_update.jl:256-266, src/tV/phy_update.jl:190-205, src/tUV/phy_update.jl:188-203),
and the arrangement of draws taken in the reference's visiting order."""
import numpy as np
import pytest

from kapdqmc_amd.dqmc import elements_per_slice, field_shape, replay_from_visiting_order
from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para


def tu():
    return tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=0.5, Θquench=0.0, Lattice="HoneyComb120",
                           site=[2, 2], BatchSize=5)


def tv():
    return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.1, Θrelax=0.5, Θquench=0.0, Lattice="HoneyComb120",
                           site=[2, 2], BatchSize=5)


def tuv():
    return tUV_Hubbard_Para(Ht=1.0, Hu=4.0, Hv=0.5, Lattice="HoneyComb120", site=[2, 2], Δt=0.1, Θ=0.5, BatchSize=5)


MODELS = [("tU", tu, "OracleTU"), ("tV", tv, "OracleTV"), ("tUV", tuv, "OracleTUV")]


def philox_layout(O, m, seed, chain, sweep0, nsweeps):
    """The Philox draws the oracle itself would use, in the replay layout."""
    E, nt = elements_per_slice(m), m.Nt
    ch = np.zeros((nsweeps, 2, nt, E), dtype=np.uint8)
    u = np.zeros((nsweeps, 2, nt, E))
    for s in range(nsweeps):
        for d in (0, 1):
            for lt in range(1, nt + 1):
                for e in range(E):
                    ch[s, d, lt - 1, e], u[s, d, lt - 1, e] = O.draw(seed, chain, 0, sweep0 + s, lt | (d << 31), e)
    return ch, u


@pytest.mark.parametrize("name,mk,cls", MODELS)
def test_replaying_the_own_stream_reproduces_the_run(oracle_mod, name, mk, cls):
    m = mk()
    a = getattr(oracle_mod, cls)(m, seed=11)
    a.initial_s()
    b = getattr(oracle_mod, cls)(m, seed=99)          # another stream, replaying a's draws
    b.s = a.s.copy()
    b.set_replay(*philox_layout(oracle_mod, m, 11, 0, 0, 2))
    ra, rb = a.phy_update(2, True), b.phy_update(2, True)
    assert np.array_equal(a.s, b.s)
    assert np.abs(a.G - b.G).max() < 1e-14
    assert np.array_equal(ra, rb)


@pytest.mark.parametrize("name,mk,cls", MODELS)
def test_choice_selects_among_the_other_values_and_u_decides(oracle_mod, name, mk, cls):
    """u = 0 accepts every proposal with p > 0 (t-U: p = |det r|² γ'/γ > 0), so after the
    first forward slice every element holds allowed(s_old)[choice]; u = +inf rejects all."""
    m = mk()
    o = getattr(oracle_mod, cls)(m, seed=3)
    o.initial_s()
    s0 = o.s.copy()
    E, nt = elements_per_slice(m), m.Nt
    rng = np.random.default_rng(5)
    ch = rng.integers(0, 3, size=(1, 2, nt, E)).astype(np.uint8)
    u = np.full((1, 2, nt, E), np.inf)
    o.set_replay(ch, u)
    o.phy_update(1, False)
    assert np.array_equal(o.s, s0)                     # everything rejected
    if name != "tU":
        return                                         # t-V / t-U-V ratios may be ≤ 0 (det r real, not |.|²)
    o2 = getattr(oracle_mod, cls)(m, seed=3)
    o2.s = s0.copy()
    u[:] = 0.0
    o2.set_replay(ch, u)
    o2.init_chain()
    o2.update_layer(1, 0)
    flat0 = s0[:, 0]
    exp = np.array([[v for v in (1, 2, 3, 4) if v != x][c] for x, c in zip(flat0, ch[0, 0, 0])])
    assert np.array_equal(o2.s[:, 0], exp)


@pytest.mark.parametrize("name,mk,cls", MODELS)
def test_visiting_order_arrangement(oracle_mod, name, mk, cls):
    """Draws listed proposal by proposal as the reference visits them (forward lt = 1..Nt,
    backward lt = Nt..1; t-V families 3,2,1 forward and 1,2,3 backward) land where the replay
    layout expects them."""
    m = mk()
    E, nt = elements_per_slice(m), m.Nt
    ch, u = philox_layout(oracle_mod, m, 7, 0, 0, 1)
    seq_c, seq_u = [], []
    nb = m.Ns // 2
    for d, lts in ((0, range(1, nt + 1)), (1, range(nt, 0, -1))):
        for lt in lts:
            if name == "tV":
                fams = (3, 2, 1) if d == 0 else (1, 2, 3)
                elems = [i + (j - 1) * nb for j in fams for i in range(nb)]
            else:
                elems = range(E)
            for e in elems:
                seq_c.append(ch[0, d, lt - 1, e])
                seq_u.append(u[0, d, lt - 1, e])
    oc, ou = replay_from_visiting_order(m, seq_c, seq_u)
    assert np.array_equal(oc, ch) and np.array_equal(ou, u)
    assert field_shape(m)[-1] == nt


def test_sweeps_beyond_the_buffer_are_rejections(oracle_mod):
    m = tu()
    o = oracle_mod.OracleTU(m, seed=1)
    o.initial_s()
    ch, u = philox_layout(oracle_mod, m, 1, 0, 0, 1)
    u[:] = 0.0                                         # the replayed sweep accepts everything
    o.set_replay(ch, u)
    o.phy_update(1, False)
    s1 = o.s.copy()
    o.phy_update(1, False)                             # outside the buffer: all rejected
    assert np.array_equal(o.s, s1)
