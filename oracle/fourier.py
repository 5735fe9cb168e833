"""Windowed k -> R transform (oracle).

Follows exafs_gpu/fourier.py:21-80: R grid j*pi/(nfft dk) (:21-22), Athena Hanning
window (:25-33), window zeroed outside the model grid (:47-49), weight W(k) k^w,
linear resample from the model grid with index + clamped fraction (:55-62, 69-73),
numpy forward FFT (sign e^{-2ikR}) of length nfft, first nfft/2 points, times
dk/sqrt(pi) (:67, 75-80). chimera/multift.py:58-63 fans it out per channel
(edge-major) after adding chi_extra.

`dft` evaluates the same sum directly at arbitrary R (the workloads' 0.02 A grid).

Pinned by: np.fft (an FP64 library routine) equals the direct sum at the FFT grid
points; the Hanning closed form; a pure sin(2 k R0) input peaks at R0 (tests).
"""
from __future__ import annotations

import numpy as np

DK = 0.05
NFFT = 2048


def r_grid(nfft=NFFT, dk=DK):
    return np.arange(nfft // 2) * np.pi / (nfft * dk)


def hanning_window(k, kmin, kmax, dk_win=1.0):
    w = np.zeros_like(k)
    w[(k >= kmin + dk_win / 2) & (k <= kmax - dk_win / 2)] = 1.0
    lo = (k > kmin - dk_win / 2) & (k < kmin + dk_win / 2)
    w[lo] = 0.5 * (1 + np.sin(np.pi * (k[lo] - kmin) / dk_win))
    hi = (k > kmax - dk_win / 2) & (k < kmax + dk_win / 2)
    w[hi] = 0.5 * (1 - np.sin(np.pi * (k[hi] - kmax) / dk_win))
    return w


class Transform:
    """One channel: model grid -> uniform grid -> windowed FFT (fourier.py:36-80)."""

    def __init__(self, k_model, kmin=3.0, kmax=14.0, kweight=2, dk_win=1.0,
                 nfft=NFFT, dk=DK):
        self.nfft, self.dk, self.kweight = nfft, dk, kweight
        self.k_model = np.asarray(k_model, float)
        self.k_uniform = np.arange(nfft) * dk
        win = hanning_window(self.k_uniform, kmin, kmax, dk_win)
        win[self.k_uniform > self.k_model.max()] = 0.0
        win[self.k_uniform < self.k_model.min()] = 0.0
        self.weight = win * self.k_uniform ** kweight
        self.R = r_grid(nfft, dk)
        idx = np.searchsorted(self.k_model, self.k_uniform) - 1
        idx = np.clip(idx, 0, self.k_model.size - 2)
        span = np.maximum(self.k_model[idx + 1] - self.k_model[idx], 1e-12)
        frac = np.clip((self.k_uniform - self.k_model[idx]) / span, 0.0, 1.0)
        inside = (self.k_uniform >= self.k_model[0]) & (self.k_uniform <= self.k_model[-1])
        self.idx, self.frac, self.inside = idx, frac * inside, inside.astype(float)
        self.norm = dk / np.sqrt(np.pi)

    def resample(self, chi):
        a = chi[..., self.idx]
        b = chi[..., self.idx + 1]
        return (a + (b - a) * self.frac) * self.inside

    def to_R(self, chi):
        cu = self.resample(chi) * self.weight
        return np.fft.fft(cu, n=self.nfft, axis=-1)[..., : self.nfft // 2] * self.norm

    def dft(self, chi, R):
        """The same transform evaluated directly at the points R (any grid)."""
        cu = self.resample(chi) * self.weight
        E = np.exp(-2j * np.outer(self.k_uniform, np.asarray(R, float)))
        return (cu @ E) * self.norm


def multi_to_R(transforms, chan_edge, chi, chi_extra=None, R=None):
    """(..., nsp, nk) -> (..., nchan, nR), channel by channel (multift.py:58-63)."""
    if chi_extra is not None:
        chi = chi + chi_extra
    if R is None:
        out = [t.to_R(chi[..., e, :]) for t, e in zip(transforms, chan_edge)]
    else:
        out = [t.dft(chi[..., e, :], R) for t, e in zip(transforms, chan_edge)]
    return np.stack(out, axis=-2)
