"""Autocorrelation analysis of per-sweep observable series.

tau_int uses Sokal's automatic windowing (A. Sokal, "Monte Carlo methods in statistical
mechanics", Cargèse lectures 1996, §3): τ_int(M) = 1/2 + Σ_{t=1}^{M} ρ(t), with the
smallest window M ≥ c τ_int(M), c = 5. The effective number of independent samples of
a series of n sweeps is n / (2 τ_int).
"""
from __future__ import annotations

import numpy as np


def autocorr(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    y = x - x.mean()
    f = np.fft.rfft(y, 2 * n)
    ac = np.fft.irfft(f * np.conj(f))[:n]
    if ac[0] <= 0:
        return np.zeros(n)
    return ac / ac[0]


def tau_int(x, c: float = 5.0) -> float:
    """Integrated autocorrelation time (in sweeps); 0.5 for an uncorrelated series."""
    x = np.asarray(x, dtype=np.float64)
    if len(x) < 4:
        return 0.5
    rho = autocorr(x)
    tau = 0.5
    for M in range(1, len(x)):
        tau += rho[M]
        if M >= c * tau:
            break
    return max(tau, 0.5)


def tau_int_chains(X, c: float = 5.0) -> float:
    """τ_int of several independent chains of the same Markov process (rows of X, equal
    lengths): the autocovariances of the chains (each about its own mean) are averaged,
    normalised at lag 0, then windowed as in tau_int. Used where each chain alone is too
    short to resolve its own autocorrelation."""
    X = np.atleast_2d(np.asarray(X, dtype=np.float64))
    nch, n = X.shape
    if n < 4:
        return 0.5
    Y = X - X.mean(axis=1, keepdims=True)
    f = np.fft.rfft(Y, 2 * n, axis=1)
    acov = np.fft.irfft(f * np.conj(f), axis=1)[:, :n].mean(axis=0)
    if acov[0] <= 0:
        return 0.5
    rho = acov / acov[0]
    tau = 0.5
    for M in range(1, n):
        tau += rho[M]
        if M >= c * tau:
            break
    return max(tau, 0.5)


def binning_error(x, nbins: int = 16) -> float:
    """Standard error of the mean from nbins equal bins (cross-check of tau_int)."""
    x = np.asarray(x, dtype=np.float64)
    nb = min(nbins, len(x))
    if nb < 2:
        return float("nan")
    b = np.array([c.mean() for c in np.array_split(x, nb)])
    return b.std(ddof=1) / np.sqrt(nb)


def effective_samples(x, c: float = 5.0) -> float:
    return len(x) / (2.0 * tau_int(x, c))
