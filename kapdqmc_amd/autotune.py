"""Autotuner for the decision block size k and the GEMM tile of a batch.

Times each candidate on the batch itself: `slices` slices of the forward sweep (wrap +
local update) with the device time of the wrap, decide and flush kernel families read from
HIP events (kdq_profile). The fields are saved before and restored after, and the chain is
re-initialised, so tuning leaves the Markov chains where they were. Results can be cached
in a JSON file keyed by the batch shape.
"""
from __future__ import annotations

import json
import os

import numpy as np

FAMILIES = ("wrap", "decide", "flush")


def _key(batch) -> str:
    ms = batch.models
    return json.dumps({"kind": batch.kind, "ns": [int(m.Ns) for m in ms], "nt": int(ms[0].Nt)})


def _time_slices(batch, slices: int) -> float:
    batch.init_chain()
    batch.profile(FAMILIES)
    nt = batch.models[0].Nt
    for lt in range(1, min(slices, nt) + 1):
        batch.wrap(lt, True)
        batch.update_layer(lt, 0, 0)
    prof = batch.profile_read()
    batch.profile([])
    return sum(prof[f][0] for f in FAMILIES)


def tune(batch, delays=(8, 16, 32, 64), tiles=(0, 32, 64), slices: int = 4, cache: str | None = None) -> dict:
    """Picks (delay, gemm_tile) for a t-U DQMCBatch, applies it and returns
    {"delay": k, "gemm_tile": t, "ms_per_slice": {...}}."""
    if batch.kind != "tU":
        raise ValueError("the block size is tunable for t-U batches only")
    key = _key(batch)
    if cache and os.path.exists(cache):
        with open(cache, encoding="utf-8") as fh:
            known = json.load(fh)
        if key in known:
            best = known[key]
            batch.set_option("delay", best["delay"])
            batch.set_option("gemm_tile", best["gemm_tile"])
            return best
    saved = [batch.get_fields(c).copy() for c in range(batch.n)]
    nmin = min(m.Ns for m in batch.models)
    timing = {}
    try:
        for k in delays:
            if k > 64 or k > nmin:
                continue
            batch.set_option("delay", k)
            batch.set_option("gemm_tile", 0)
            timing[f"delay={k}"] = _time_slices(batch, slices) / slices
            for c, s in enumerate(saved):
                batch.set_fields(c, s)
        k_best = min((k for k in delays if f"delay={k}" in timing), key=lambda k: timing[f"delay={k}"])
        batch.set_option("delay", k_best)
        for t in tiles:
            batch.set_option("gemm_tile", t)
            timing[f"tile={t}"] = _time_slices(batch, slices) / slices
            for c, s in enumerate(saved):
                batch.set_fields(c, s)
        t_best = min(tiles, key=lambda t: timing[f"tile={t}"])
        batch.set_option("gemm_tile", t_best)
    finally:
        for c, s in enumerate(saved):
            batch.set_fields(c, s)
        batch.init_chain()
    best = {"delay": int(k_best), "gemm_tile": int(t_best), "ms_per_slice": timing}
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache, encoding="utf-8") as fh:
                known = json.load(fh)
        known[key] = best
        with open(cache, "w", encoding="utf-8") as fh:
            json.dump(known, fh, indent=1)
    return best
