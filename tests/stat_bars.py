"""Statistical bars for comparing sampled Monte Carlo averages with exact values (test helper).

Each chain of a batch is an independent Markov chain, so the per-chain means of several
observables are i.i.d. draws of a vector whose mean is the exact expectation. The joint test is
Hotelling's T² with the chain-to-chain covariance: (n − p) / (p (n − 1)) T² follows
F(p, n − p) under the hypothesis that the sampled distribution is the exact one. Requiring
its p-value to be ≥ 0.05 is the several-observable form of a 2σ bar (a 2σ bar applied to each
of p correlated observables alone would fail a correct sampler with probability up to
1 − 0.954^p); each observable is in addition kept within 3σ on its own.
"""
import numpy as np
from scipy import stats


def hotelling(per_chain: np.ndarray, exact) -> tuple:
    """(p-value, deviations in σ) of the per-chain means per_chain[n, p] against exact[p]."""
    X = np.asarray(per_chain, dtype=np.float64)
    n, p = X.shape
    d = X.mean(axis=0) - np.asarray(exact, dtype=np.float64)
    S = np.cov(X, rowvar=False).reshape(p, p)
    T2 = n * d @ np.linalg.solve(S, d)
    Fv = (n - p) / (p * (n - 1)) * T2
    sig = d / np.sqrt(np.diag(S) / n)
    return float(stats.f.sf(Fv, p, n - p)), sig


def assert_agrees(per_chain, exact, names=None):
    pval, sig = hotelling(per_chain, exact)
    names = names or [f"obs{k}" for k in range(len(sig))]
    text = ", ".join(f"{k} {s:+.2f}σ" for k, s in zip(names, sig))
    print(f"joint p = {pval:.3f}; {text}", flush=True)
    assert pval >= 0.05, f"sampled averages disagree with the exact values (p = {pval:.3g}): {text}"
    assert np.all(np.abs(sig) < 3.0), text
    return pval, sig
