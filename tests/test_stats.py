"""tau_int / effective-sample estimator against the AR(1) closed form
τ_int = (1 + φ) / (2 (1 - φ)) and against binning."""
import numpy as np

from kapdqmc_amd import stats


def ar1(phi, n, seed=0):
    rs = np.random.default_rng(seed)
    e = rs.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return x


def test_uncorrelated():
    x = np.random.default_rng(1).standard_normal(50000)
    assert abs(stats.tau_int(x) - 0.5) < 0.05


def test_ar1_closed_form():
    for phi in (0.5, 0.8, 0.9):
        x = ar1(phi, 200000, seed=int(phi * 10))
        ref = (1 + phi) / (2 * (1 - phi))
        assert abs(stats.tau_int(x) / ref - 1) < 0.08, (phi, stats.tau_int(x), ref)


def test_binning_consistent_with_tau():
    x = ar1(0.8, 200000, seed=3)
    naive = x.std(ddof=1) / np.sqrt(len(x))
    err = stats.binning_error(x, 50)
    assert abs((err / naive) ** 2 / (2 * stats.tau_int(x)) - 1) < 0.3


def test_pooled_chains_match_closed_form():
    """64 short AR(1) chains (400 steps each): the pooled estimate recovers the closed form
    that a single 400-step chain resolves only loosely."""
    phi = 0.6
    X = np.stack([ar1(phi, 400, seed=100 + c) for c in range(64)])
    ref = (1 + phi) / (2 * (1 - phi))
    assert abs(stats.tau_int_chains(X) / ref - 1) < 0.12
    assert abs(stats.tau_int_chains(ar1(0.0, 400, seed=5)[None, :]) - 0.5) < 0.2
