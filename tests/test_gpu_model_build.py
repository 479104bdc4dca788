"""Chains built from the library's C++ model constructor (kdq_model_build) sweep like chains
built by the Python host constructors: the same constants to rounding, a trial state equal
up to a rotation inside its span (the same G), so the same seeds give the same fields and G
to rounding. t-U (an uneven batch) and t-V."""
import numpy as np
import pytest

from kapdqmc_amd.models import native_model, tU_Hubbard_Para, tV_Hubbard_Para

pytestmark = pytest.mark.gpu


def run(models, sweeps=2, delay=16):
    from kapdqmc_amd.dqmc import DQMCBatch
    b = DQMCBatch(models, list(range(3, 3 + len(models))), list(range(len(models))), delay=delay)
    b.initial_s()
    rows = b.phy_update(sweeps, True)
    out = [b.get_fields(c) for c in range(b.n)], [b.get_G(c) for c in range(b.n)], rows
    b.close()
    return out


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_tu_native_models_sweep_like_host_models():
    print("case: t-U sweeps, native vs host model constants, L = 3, 4", flush=True)
    kws = [dict(Ht=1.0, Hu1=U, Hu2=U, Δt=0.1, Θrelax=1.0, Θquench=0.0, Lattice="HoneyComb120", site=[L, L],
                BatchSize=5) for L, U in ((3, 4.0), (4, 2.0))]
    f1, g1, r1 = run([native_model("tU", **kw) for kw in kws])
    f0, g0, r0 = run([tU_Hubbard_Para(**kw) for kw in kws])
    for c in range(len(kws)):
        assert np.array_equal(f1[c], f0[c]), c
        assert rel(g1[c], g0[c]) < 1e-9, (c, rel(g1[c], g0[c]))
    assert np.allclose(r1, r0, rtol=1e-8, atol=1e-10)


def test_tv_native_model_sweeps_like_host_model():
    print("case: t-V sweeps, native vs host model constants, L = 3", flush=True)
    kw = dict(Ht=1.0, Hv1=1.0, Hv2=1.0, Δt=0.05, Θrelax=0.5, Θquench=0.0, Lattice="HoneyComb120", site=[3, 3],
              BatchSize=5)
    f1, g1, r1 = run([native_model("tV", **kw)], delay=8)
    f0, g0, r0 = run([tV_Hubbard_Para(**kw)], delay=8)
    assert np.array_equal(f1[0], f0[0])
    assert rel(g1[0], g0[0]) < 1e-9
    assert np.allclose(r1, r0, rtol=1e-8, atol=1e-10)
