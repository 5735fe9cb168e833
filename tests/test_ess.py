"""ESS estimator against the closed form of an AR(1) process."""
import numpy as np

from chimera_amd.ess import autocorr, ess


def _ar1(rho, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.standard_normal(n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - rho * rho)
    for t in range(1, n):
        x[t] = rho * x[t - 1] + e[t]
    return x


def test_ess_ar1_closed_form():
    n = 200_000
    for rho in (0.0, 0.5, 0.9, 0.98):
        want = n * (1 - rho) / (1 + rho)
        got = np.mean([ess(_ar1(rho, n, s)) for s in range(3)])
        assert abs(got / want - 1) < 0.12, (rho, got, want)


def test_autocorr_lag_values():
    x = _ar1(0.7, 100_000, 1)
    r = autocorr(x)
    assert abs(r[1] - 0.7) < 0.02 and abs(r[2] - 0.49) < 0.03


def test_ess_degenerate():
    assert ess(np.ones(10)) == 10.0
    assert ess([]) == 0.0
