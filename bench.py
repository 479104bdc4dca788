#!/usr/bin/env python
"""Benchmark: projector-DQMC t-U sweep on one MI355X.

Workload (the plain FP64 GPU sweep of the drop-in path): half-filled t-U Hubbard model on
a 12 x 12 honeycomb lattice (HoneyComb120, N = 288, Np = 144), U/t = 4, Δτ = 0.1,
2Θ = 20 (Nt = 200 slices), stabilisation every 10 slices, one Markov chain, FP64
everywhere, local updates as delayed rank-k blocks (k = --delay), record = true (40
measurements per sweep as in the reference). A step is one sweep (forward + backward).

value = effective independent samples per second of the double occupancy D = Eu/N:
(sweeps/s) / (2 τ_int), τ_int by Sokal windowing over the timed sweeps.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def model_cfg2():
    from kapdqmc_amd.models import tU_Hubbard_Para
    return tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=10.0, Θquench=0.0,
                           Lattice="HoneyComb120", site=[12, 12], BatchSize=10)


def family_work(fam: str, N: int, Np: int):
    """Algorithmic work of one launch of a kernel family: (bound, amount, unit)."""
    if fam in ("wrap", "meas_wrap"):
        return "mfma", 4.0 * N ** 3, "flop"          # real N x N times complex N x N
    if fam == "chain":
        return "mfma", 4.0 * N * N * Np, "flop"
    if fam == "get_G":
        return "mfma", 8.0 * (2 * Np * Np * N + N * N * Np) / 3.0, "flop"  # mean of its 3 GEMMs
    if fam == "qr":
        return "mfma", 2 * 8.0 * (2 * N * Np * Np - 2 * Np ** 3 / 3.0), "flop"  # geqrf + orgqr
    if fam == "inv":
        return "mfma", 8.0 * Np ** 3, "flop"
    if fam == "flush":
        return "hbm", 32.0 * N * N, "byte"          # read + write one complex N x N
    if fam == "decide":
        return "hbm", 16.0 * 32 * 32, "byte"
    return "hbm", 16.0 * N * N, "byte"


def cpu_baseline(m, tau, nslices: int):
    """The oracle (CPU FP64 restatement of the reference algorithm, complex dense GEMMs and
    sequential Sherman-Morrison) on a bounded sample: chain initialisation + `nslices`
    forward slices, plus one measurement with 8 wraps (the mean over the 40 measurements of
    a sweep is 310/40 = 7.75 wraps). Extrapolated to a sweep and to ESS/s with the GPU
    chain's τ_int (same Markov chain)."""
    import numpy as np
    from oracle import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = oracle.OracleTU(m, seed=1)
    o.initial_s()
    t0 = time.perf_counter()
    o.init_chain()
    t_init = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.forward_slices(nslices)
    t_fw = time.perf_counter() - t0 - t_init
    t0 = time.perf_counter()
    o.measure(m.Nt // 2 + 8, o.G)
    t_meas = time.perf_counter() - t0
    t_sweep = t_init + (2 * m.Nt / nslices) * max(t_fw, 1e-9) + 40 * t_meas
    sweeps_s = 1.0 / t_sweep
    return {"value": sweeps_s / (2 * tau), "unit": "effective samples/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle (C FP64, OpenMP {threads} threads) on the same N=288 model: chain init "
                       f"{t_init:.2f}s + {nslices} forward slices {t_fw:.2f}s + 1 measurement "
                       f"(8 wraps) {t_meas:.2f}s, extrapolated to {t_sweep:.1f}s/sweep "
                       f"({sweeps_s:.4f} sweeps/s), ESS/s with the GPU chain's tau_int={tau:.2f}")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--delay", type=int, default=32)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-slices", type=int, default=20)
    ap.add_argument("--profile-all", action="store_true", help="time every kernel family (slower)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build measures one GPU only (independent chains shard by chain "
              "across GPUs with no exchange during sweeps); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        sys.exit(1)
    torch.cuda.set_device(0)
    from kapdqmc_amd import stats
    from kapdqmc_amd.dqmc import DQMCBatch
    from kapdqmc_amd.rng import initial_fields

    m = model_cfg2()
    N, Np = m.Ns, m.Ns // 2
    b = DQMCBatch([m], [args.seed], [0], delay=args.delay)
    b.set_fields(0, initial_fields(args.seed, 0, N, m.Nt))

    # warm-up with every family timed, to find the dominant kernel family
    b.profile(["wrap", "chain", "get_G", "decide", "flush", "qr", "inv", "meas_wrap", "measure"])
    b.phy_update(max(args.warmup, 1), True)
    prof_w = b.profile_read()
    dominant = max(prof_w, key=lambda f: prof_w[f][0])

    fams = list(prof_w) if args.profile_all else [dominant]
    b.profile(fams)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rows = b.phy_update(args.steps, True)[0]
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    prof = b.profile_read()
    dt = t1 - t0
    sweeps_s = args.steps / dt
    D = rows[:, 1] / N
    tau = stats.tau_int(D)
    ess_s = sweeps_s / (2 * tau)
    cnt = b.counters()[0]

    ms, nl = prof[dominant]
    bound, work, unit = family_work(dominant, N, Np)
    avg_s = ms / 1e3 / max(nl, 1)
    if bound == "mfma":
        achieved, peak, runit = work / avg_s / 1e12, PEAK_FP64_TFLOPS, "TFLOP/s"
    else:
        achieved, peak, runit = work / avg_s / 1e9, PEAK_HBM_GBS, "GB/s"
    roofline = {"bound": bound, "achieved": achieved, "peak": peak, "unit": runit,
                "frac": achieved / peak, "traffic": None, "kernel_family": dominant,
                "avg_launch_us": avg_s * 1e6, "launches": nl,
                "work_per_launch": f"{work:.4g} {unit}"}

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(m, tau, args.cpu_slices)
        except Exception as e:  # the baseline is reported, never required
            cpu = {"error": repr(e)}

    breakdown = {f: {"ms_per_sweep": prof_w[f][0] / max(args.warmup, 1),
                     "launches_per_sweep": prof_w[f][1] / max(args.warmup, 1)} for f in prof_w}
    out = {
        "metric": "effective samples/s at N=288 t-U: FP64 plain vs mixed+batched+SLMC, equal error",
        "value": ess_s,
        "unit": "effective samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox-seeded HS fields; model constants from the reference formulas)",
        "config": {"workload": "t-U honeycomb 12x12 N=288 Np=144 U=4 dtau=0.1 2Theta=20 Nt=200 BS=10, "
                               "plain FP64 sweep, 1 chain, record=true",
                   "delay_k": args.delay, "variant": "plain FP64 (mixed+batched+SLMC not built yet)"},
        "sweeps_per_s": sweeps_s,
        "tau_int_double_occupancy": tau,
        "double_occupancy_mean": float(D.mean()),
        "acceptance": float(cnt[1] / max(cnt[0], 1)),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernel_breakdown_warmup": breakdown,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
