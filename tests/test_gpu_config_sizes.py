"""The benchmark configurations at their full sizes through the GPU path (C ABI), with the
oracle where it finishes in seconds and size-independent properties elsewhere.

* Uneven batch (64 t-U chains, L = 6/9/12 by c mod 3, U = 2/4/6, seed c, 2Θ = 2L, all in one
  handle): one FP64 sweep, eight L = 6 chains and one L = 9 chain against the oracle (fields
  bit-exact, G 1e-10), and the mixed-precision batch's max |G_FP32 − G_FP64| per chain at the
  recomputes.
* Swap replicas at 9x9 (N = 162), Δτ = 0.05, 2Θ = 20 (Nt = 400), |A| = 81 (72 sites of the
  cells x = 1..4 and the 9 A sites at x = 5), B empty: the set-up (G(τ), G(Θ), G(τ,Θ), G(Θ,τ)
  of both replicas and det g) against the oracle for t-V and t-U-V; after a t-V sweep, G(τ)
  and G(Θ) are rank-(N − Np) projectors and the tracked det g equals the Grover determinant
  formed from scratch out of both replicas' G(Θ).
* Self-learning moves at 15x15 (N = 450), Nt = 240, 32 channels: network energy and loss
  gradient against the numpy restatement on the full lattice; log W(s') − log W(s) for one
  field change equals the log of the local-update ratio |1 + Δ(1 − G_ii)|² γ'/γ at that slice
  (the weight the sweep samples); a global move's decision replayed from the exact weights.
"""
import numpy as np
import pytest

from kapdqmc_amd.geometry import area_index
from kapdqmc_amd.models import tU_Hubbard_Para, tUV_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def uneven_groups():
    groups = {}
    for c in range(64):
        L, U = (6, 9, 12)[c % 3], (2.0, 4.0, 6.0)[(c // 3) % 3]
        m = tU_Hubbard_Para(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=float(L), Θquench=0.0, Lattice="HoneyComb120",
                            site=[L, L], BatchSize=10)
        groups.setdefault(L, []).append((c, m))
    return groups


def test_uneven_batch_fp64_against_oracle_and_mixed_error(oracle_mod):
    print("case: uneven 64-chain batch L=6/9/12, 2Theta=2L, one handle", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch, Initial_s
    groups = uneven_groups()
    ms = [None] * 64
    for g in groups.values():
        for c, m in g:
            ms[c] = m
    b = DQMCBatch(ms, list(range(64)), list(range(64)), delay=32)
    for c, m in enumerate(ms):
        b.set_fields(c, Initial_s(m, c, c))
    rb = b.phy_update(1, True)
    for c in [c for c, _ in groups[6][:8]] + [groups[9][0][0]]:   # eight L = 6 chains, one L = 9
        m = ms[c]
        o = oracle_mod.OracleTU(m, seed=c, chain=c)
        o.s = Initial_s(m, c, c)
        ro = o.phy_update(1, True)
        assert np.array_equal(b.get_fields(c), o.s), c
        assert rel(b.get_G(c), o.G) < 1e-10
        assert np.allclose(rb[c], ro, rtol=1e-9, atol=1e-11)
    b.close()
    mx = DQMCBatch(ms, list(range(64)), list(range(64)), delay=32, precision="mixed")
    for c, m in enumerate(ms):
        mx.set_fields(c, Initial_s(m, c, c))
    mx.phy_update(1, False)
    e_all = mx.precision_error()
    mx.close()
    for L, g in groups.items():
        e = e_all[[c for c, _ in g]]
        print(f"L={L}: max|G_FP32-G_FP64| per chain: median {np.median(e):.2e}, max {e.max():.2e}", flush=True)
        assert np.all(e > 0) and np.all(e < 1e-2)


def region81(m):
    iA = list(area_index(m.Lattice, m.site, ([1, 1], [4, 9]))) + [2 * (4 + 9 * y) + 1 for y in range(9)]
    iA = np.asarray(iA, dtype=np.int64)
    assert len(iA) == 81 and len(set(iA.tolist())) == 81
    return iA, np.zeros(0, dtype=np.int64)


def tv9():
    return tV_Hubbard_Para(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.05, Θrelax=10.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[9, 9], BatchSize=10)


def tuv9():
    return tUV_Hubbard_Para(Ht=1.0, Hu=4.0, Hv=0.5, Lattice="HoneyComb120", site=[9, 9], Δt=0.05, Θ=10.0,
                            BatchSize=10)


@pytest.mark.parametrize("mk", [tv9, tuv9], ids=["tV", "tUV"])
def test_scee_9x9_setup_against_oracle(oracle_mod, mk):
    print(f"case: SCEE set-up 9x9 Nt=400 |A|=81 {mk.__name__}", flush=True)
    from kapdqmc_amd.dqmc import Initial_s, SCEEBatch
    m = mk()
    assert m.Nt == 400 and m.Ns == 162
    iA, iB = region81(m)
    s1, s2 = Initial_s(m, 11), Initial_s(m, 12)
    o = oracle_mod.OracleSCEE2(m, iA.astype(np.int32), iB.astype(np.int32), 0.5, 2, s1, s2, seed=11)
    o.init()
    b = SCEEBatch([m], iA, iB, 0.5, 2, [11], [0])
    b.set_fields(0, 0, s1)
    b.set_fields(0, 1, s2)
    b.init()
    for r in (0, 1):
        for name, x, y in zip(("Gt", "G0", "Gt0", "G0t"), b.G4(0, r), o.G4(r)):
            assert rel(x, y) < 1e-9, (r, name, rel(x, y))
    d = b.detg()[0]
    assert abs(d[0] / o.detg(0) - 1) < 1e-8
    b.close()


def test_tv_scee_9x9_sweep_properties():
    print("case: t-V SCEE 9x9 Nt=400 |A|=81, one sweep", flush=True)
    from kapdqmc_amd.dqmc import Initial_s, SCEEBatch
    m = tv9()
    iA, iB = region81(m)
    b = SCEEBatch([m], iA, iB, 0.5, 2, [21], [0])
    b.set_fields(0, 0, Initial_s(m, 21))
    b.set_fields(0, 1, Initial_s(m, 22))
    O = b.sweeps(1)
    assert np.all(np.isfinite(O))
    A = iA - 1
    G0s = []
    for r in (0, 1):
        Gt, G0, Gt0, G0t = b.G4(0, r)
        for G in (Gt, G0):
            assert np.abs(G @ G - G).max() < 1e-8
            assert abs(np.trace(G).real - (m.Ns - m.Ns // 2)) < 1e-8
        G0s.append(G0.real)
    a, c = G0s[0][np.ix_(A, A)], G0s[1][np.ix_(A, A)]
    I = np.eye(len(A))
    fresh = np.linalg.det(a @ c + (I - a) @ (I - c))
    tracked = b.detg()[0, 0]
    print(f"det g tracked {tracked:.6e}, fresh {fresh:.6e}", flush=True)
    assert abs(tracked / fresh - 1) < 1e-7
    b.close()


def tu15():
    return tU_Hubbard_Para(Ht=1.0, Hu1=5.0, Hu2=5.0, Δt=0.1, Θrelax=12.0, Θquench=0.0, Lattice="HoneyComb120",
                           site=[15, 15], BatchSize=10)


def test_slmc_network_full_size_against_numpy():
    print("case: SLMC net 15x15 Nt=240 C=32 vs numpy", flush=True)
    from kapdqmc_amd import slmc
    from oracle import slmc_net as sn
    m = tu15()
    C = 32
    net = slmc.EffectiveModel.for_model(m, channels=C, seed=3, max_batch=2)
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.05, net.num_params).astype(np.float32)
    net.set_params(p)
    geo = sn.Geometry(15, 15, m.Nt)
    f = rng.integers(1, 5, size=(2, geo.n, m.Nt)).astype(np.uint8)
    fields = [np.asfortranarray(x) for x in f]
    pb = sn.bf16_weights(p, C)
    E = net.energy(fields)
    Eo = sn.energy(geo, pb, C, f)
    assert np.abs(E - Eo).max() < 1e-4 * np.abs(Eo).max()
    lw = rng.normal(0, 3, 2)
    loss, g = net.loss_grad(fields, lw)
    lo, go = sn.loss_grad(geo, pb, C, f, lw)
    assert abs(loss - lo) < 1e-3 * max(lo, 1.0)
    assert np.linalg.norm(g - go) < 5e-3 * np.linalg.norm(go)
    net.close()


def test_log_weight_difference_is_the_local_ratio_at_full_size():
    print("case: log W single flip vs local ratio, 15x15 Nt=240", flush=True)
    from kapdqmc_amd.dqmc import DQMCBatch
    m = tu15()
    b = DQMCBatch([m], [5], [0], delay=32)
    b.initial_s()
    b.phy_update(1, False)
    s = b.get_fields(0)
    lw0 = b.log_weight()[0]
    lt, i = 97, 123
    Gt = b.G4(0, lt, lt)[0]
    si = int(s[i, lt - 1])
    sx = 1 + (si % 4)
    d = np.exp(1j * m.α[lt - 1] * (m.η[sx - 1] - m.η[si - 1])) - 1.0
    ratio = abs(1 + d * (1 - Gt[i, i])) ** 2 * m.γ[sx - 1] / m.γ[si - 1]
    s2 = s.copy()
    s2[i, lt - 1] = sx
    b.set_fields(0, s2)
    lw1 = b.log_weight()[0]
    print(f"log W {lw0:.6f} -> {lw1:.6f}: diff {lw1 - lw0:.9f}, log ratio {np.log(ratio):.9f}", flush=True)
    assert abs((lw1 - lw0) - np.log(ratio)) < 1e-9 * abs(lw0) + 1e-8
    b.close()
