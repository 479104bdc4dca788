"""Why the kinetic products (the wrap G ← eK G eK⁻¹, `WrapKV!` `src/tU/GreenMatrix.jl:56-86`)
stay in FP32 or FP64 and are not tried in BF16: the error of ONE wrap of a real chain's G
(t-U, 6x6 honeycomb, U = 4, G of the oracle after chain set-up) with the operands rounded to
each format and the products accumulated in FP32, as the CDNA4 matrix cores do for BF16 and
FP32 inputs, against the |ΔG| ≤ 1e-6 budget of the mixed-precision mode.

Measured: FP32 7.6e-7, BF16 1.3e-2, BF16x3 (each operand split into a BF16 high part and a
BF16 low part, three products) 1.6e-5. A wrap is applied Nt times per sweep between FP64
recomputations 10 slices apart, so only FP32 fits (and only just).
"""
import numpy as np
import torch

from kapdqmc_amd.models import tU_Hubbard_Para

BUDGET = 1e-6


def _round(x: np.ndarray, dt) -> np.ndarray:
    return torch.from_numpy(np.ascontiguousarray(x)).to(dt).double().numpy()


def _mm(A: np.ndarray, B: np.ndarray, dt) -> np.ndarray:
    """Real A times real B with operands rounded to dt and FP32 accumulation."""
    a = torch.from_numpy(_round(A, dt)).float()
    b = torch.from_numpy(_round(B, dt)).float()
    return (a @ b).double().numpy()


def _mm_split(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    Ah = _round(A, torch.bfloat16)
    Bh = _round(B, torch.bfloat16)
    bf = torch.bfloat16
    return _mm(Ah, Bh, bf) + _mm(Ah, B - Bh, bf) + _mm(A - Ah, Bh, bf)


def _wrap(eK, eKinv, G, mm):
    """eK G eKinv for complex G (planar: two real products per side, as on the GPU)."""
    T = mm(eK, G.real) + 1j * mm(eK, G.imag)
    return (mm(eKinv.T, T.real.T) + 1j * mm(eKinv.T, T.imag.T)).T


def test_kinetic_products_error_budget(oracle_mod):
    m = tU_Hubbard_Para(Ht=1.0, Hu1=4.0, Hu2=4.0, Δt=0.1, Θrelax=1.0, Θquench=0.0,
                        Lattice="HoneyComb120", site=[6, 6], BatchSize=10, Initial="H0")
    o = oracle_mod.OracleTU(m, seed=3)
    o.initial_s()
    o.init_chain()
    G = o.G
    ref = m.eK @ G @ m.eKinv
    err = {
        "fp32": np.abs(_wrap(m.eK, m.eKinv, G, lambda a, b: _mm(a, b, torch.float32)) - ref).max(),
        "bf16": np.abs(_wrap(m.eK, m.eKinv, G, lambda a, b: _mm(a, b, torch.bfloat16)) - ref).max(),
        "bf16x3": np.abs(_wrap(m.eK, m.eKinv, G, _mm_split) - ref).max(),
    }
    print({k: f"{v:.2e}" for k, v in err.items()})
    assert err["fp32"] < 2 * BUDGET
    assert err["bf16"] > 1e3 * BUDGET
    assert err["bf16x3"] > 5 * BUDGET
