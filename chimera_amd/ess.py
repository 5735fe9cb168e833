"""Effective sample size of a Markov-chain series (host analytics, no reference
counterpart: the reference reports no ESS).

`ess` uses Geyer's initial monotone sequence estimator (Geyer 1992, "Practical
Markov chain Monte Carlo", Statist. Sci. 7:473): the autocorrelations are summed in
consecutive pairs Gamma_m = rho_{2m} + rho_{2m+1} while they stay positive, made
non-increasing, and ESS = N / (1 + 2 sum_{t>=1} rho_t) = N / (-1 + 2 sum_m Gamma_m).
Autocorrelations come from an FFT of the zero-padded, centred series.
"""
from __future__ import annotations

import numpy as np


def autocorr(x):
    x = np.asarray(x, float)
    n = x.size
    x = x - x.mean()
    f = np.fft.rfft(x, n=2 * n)
    acov = np.fft.irfft(f * np.conj(f))[:n] / n
    if acov[0] <= 0:
        return np.zeros(n)
    return acov / acov[0]


def ess(x):
    """Effective sample size of a 1-D series (N for an uncorrelated one)."""
    x = np.asarray(x, float)
    n = x.size
    if n < 4 or np.ptp(x) == 0:
        return float(n) if n else 0.0
    rho = autocorr(x)
    m = n // 2
    G = rho[0:2 * m:2] + rho[1:2 * m:2]
    # initial positive sequence, then monotone
    pos = np.nonzero(G <= 0)[0]
    k = pos[0] if pos.size else G.size
    G = np.minimum.accumulate(G[:k]) if k else G[:0]
    tau = -1.0 + 2.0 * G.sum()
    tau = max(tau, 1.0 / n)
    return float(min(n / tau, n))


def ess_multi(series):
    """ESS of each column of (n, d) series; returns the array."""
    s = np.asarray(series, float)
    if s.ndim == 1:
        return np.array([ess(s)])
    return np.array([ess(s[:, j]) for j in range(s.shape[1])])
