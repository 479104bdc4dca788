"""The autotuner picks a valid (delay, tile), leaves the chains' fields untouched, and the
tuned batch reproduces a batch created with that delay from the start."""
import json

import numpy as np
import pytest

from kapdqmc_amd.models import tU_Hubbard_Para
from kapdqmc_amd.rng import initial_fields

pytestmark = pytest.mark.gpu


def test_tune_restores_state_and_matches_fresh_batch(tmp_path):
    print("case: autotune L=6, 2 chains", flush=True)
    from kapdqmc_amd import autotune, dqmc
    ms = [tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                          Lattice="HoneyComb120", site=[6, 6], BatchSize=5)] * 2
    s0 = [initial_fields(3 + c, c, ms[0].Ns, ms[0].Nt) for c in range(2)]
    b = dqmc.DQMCBatch(ms, [3, 4], [0, 1], delay=16)
    for c in range(2):
        b.set_fields(c, s0[c])
    cache = str(tmp_path / "tune.json")
    best = autotune.tune(b, delays=(4, 16, 32), tiles=(0, 32), slices=2, cache=cache)
    assert best["delay"] in (4, 16, 32) and best["gemm_tile"] in (0, 32)
    assert json.load(open(cache))
    for c in range(2):
        assert np.array_equal(b.get_fields(c), s0[c])
    rows = b.phy_update(1, True)
    f = dqmc.DQMCBatch(ms, [3, 4], [0, 1], delay=best["delay"])
    f.set_option("gemm_tile", best["gemm_tile"])
    for c in range(2):
        f.set_fields(c, s0[c])
    assert np.array_equal(f.phy_update(1, True), rows)
    for c in range(2):
        assert np.array_equal(f.get_fields(c), b.get_fields(c))
    again = autotune.tune(b, cache=cache)   # served from the cache
    assert again["delay"] == best["delay"]
