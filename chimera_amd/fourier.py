"""Windowed k -> R transform on the device, as a DFT-matrix contraction.

Drop-in for exafs_gpu/fourier.py:21-129 (r_grid, hanning_window, FourierTransform).
The host part builds what the reference builds in its constructor (uniform grid,
Hanning window zeroed outside the model grid, W(k) k^w, resample index + clamped
fraction, norm dk/sqrt(pi)); the device folds resample, window and the e^{-2ikR}
kernel into one matrix per channel and `to_R` is a GEMM against it.  The result equals
numpy's forward FFT of the windowed, resampled chi at the FFT grid points
R_j = j pi/(nfft dk); `R` may also be any other grid (the sampler's 0.02 A window).

`resample` (the linear resample onto the uniform grid on its own) runs on the device
too.  Not provided: the Method-B adjoints `grad_resample` / `grad_magnitude`
(fourier.py:86-121; no caller on this path).
"""
from __future__ import annotations

import sys

import numpy as np

from . import _lib
from ._lib import check, f64, i32

DK = 0.05
NFFT = 2048


def r_grid(nfft=NFFT, dk=DK):
    return np.arange(nfft // 2) * np.pi / (nfft * dk)


def hanning_window(k, kmin, kmax, dk_win=1.0):
    """Athena-style Hanning window with cosine tapers of width dk_win (fourier.py:25-33)."""
    w = np.zeros_like(k)
    w[(k >= kmin + dk_win / 2) & (k <= kmax - dk_win / 2)] = 1.0
    lo = (k > kmin - dk_win / 2) & (k < kmin + dk_win / 2)
    w[lo] = 0.5 * (1 + np.sin(np.pi * (k[lo] - kmin) / dk_win))
    hi = (k > kmax - dk_win / 2) & (k < kmax + dk_win / 2)
    w[hi] = 0.5 * (1 - np.sin(np.pi * (k[hi] - kmax) / dk_win))
    return w


class FourierTransform:
    """One channel: chi on the model k grid -> complex chi(R)."""

    def __init__(self, k_model, kmin=3.0, kmax=14.0, kweight=2, dk_win=1.0, nfft=NFFT, dk=DK,
                 xp=None, R=None):
        self.xp = np
        self.nfft, self.dk, self.kweight = int(nfft), float(dk), int(kweight)
        self.k_model = np.asarray(k_model, float)
        self.k_uniform = np.arange(self.nfft) * self.dk
        self.kmin, self.kmax, self.dk_win = float(kmin), float(kmax), float(dk_win)
        win = hanning_window(self.k_uniform, kmin, kmax, dk_win)
        win[self.k_uniform > self.k_model.max()] = 0.0
        win[self.k_uniform < self.k_model.min()] = 0.0
        self.weight = win * self.k_uniform ** self.kweight
        self.R = r_grid(self.nfft, self.dk) if R is None else np.asarray(R, float)
        idx = np.searchsorted(self.k_model, self.k_uniform) - 1
        idx = np.clip(idx, 0, self.k_model.size - 2)
        span = np.maximum(self.k_model[idx + 1] - self.k_model[idx], 1e-12)
        frac = np.clip((self.k_uniform - self.k_model[idx]) / span, 0.0, 1.0)
        inside = (self.k_uniform >= self.k_model[0]) & (self.k_uniform <= self.k_model[-1])
        self._idx = i32(idx)
        self._frac = f64(frac * inside)
        self._inside = f64(inside.astype(np.float64))
        self._norm = self.dk / np.sqrt(np.pi)
        self._dev = None

    def _device(self):
        if self._dev is None:
            self._dev = _DeviceFT([self], np.zeros(1, np.int32))
        return self._dev

    def to_R(self, chi):
        """chi (..., nk) -> complex (..., nR)."""
        chi = np.asarray(chi, float)
        lead = chi.shape[:-1]
        out = self._device().apply(chi.reshape(-1, 1, chi.shape[-1]))
        return out.reshape(lead + (self.R.size,))

    def resample(self, chi):
        """chi (..., nk) on the model grid -> (..., nfft) on the uniform grid
        (fourier.py:69-73): linear interpolation with the clamped fraction, zero outside
        the model grid; on the device, NumPy's arithmetic bit for bit."""
        chi = f64(chi)
        lead = chi.shape[:-1]
        rows = chi.reshape(-1, chi.shape[-1])
        if rows.shape[1] != self.k_model.size:
            raise ValueError(f"chi has {rows.shape[1]} points, the model grid {self.k_model.size}")
        out = np.empty((rows.shape[0], self.nfft))
        check(_lib.load().chm_ft_resample(self._device().h, 0, rows.shape[0], f64(rows), out))
        return out.reshape(lead + (self.nfft,))

    def magnitude(self, chi):
        return np.abs(self.to_R(chi))

    def window_mask(self, rmin, rmax):
        return (self.R >= rmin) & (self.R <= rmax)

    def n_independent(self, rmin, rmax):
        """Nyquist/Stern estimate 2 dk dR / pi + 2 (fourier.py:127-129)."""
        return 2 * (self.kmax - self.kmin) * (rmax - rmin) / np.pi + 2


class _DeviceFT:
    """The device DFT matrices of a set of edge-major channels sharing one R grid."""

    def __init__(self, fts, chan_edge, ctx=None):
        self.ctx = ctx if ctx is not None else _lib.Context.default()
        fts = list(fts)
        R = fts[0].R
        self.nchan = len(fts)
        self.nk = fts[0].k_model.size
        nfft = fts[0].nfft
        for f in fts:
            assert f.k_model.size == self.nk and f.nfft == nfft, "channels must share grids"
        self.nR = R.size
        self.nsp = int(np.max(chan_edge)) + 1
        idx = i32(np.stack([f._idx for f in fts]))
        frac = f64(np.stack([f._frac for f in fts]))
        ins = f64(np.stack([f._inside for f in fts]))
        w = f64(np.stack([f.weight for f in fts]))
        h = _lib.vp()
        check(_lib.load().chm_ft_create(self.ctx.h, self.nchan, self.nk, nfft, self.nR,
                                        i32(chan_edge), idx, frac, ins, w,
                                        f64(fts[0].k_uniform), f64(R), float(fts[0]._norm),
                                        _lib.C.byref(h)))
        self.h = h

    def apply(self, chi, chi_extra=None):
        """chi (nrows, nsp, nk) -> complex (nrows, nchan, nR)."""
        chi = f64(chi)
        n = chi.shape[0]
        out = np.empty((n, self.nchan, self.nR, 2))
        ex = None if chi_extra is None else f64(chi_extra)
        check(_lib.load().chm_ft_apply(self.h, n, chi, None if ex is None else ex.ctypes.data, out))
        return out[..., 0] + 1j * out[..., 1]

    def apply_edge(self, edge, rows):
        """rows (nrows, nk) through edge's channels -> (nrows, nkw, nR, 2) real."""
        rows = f64(rows)
        n = rows.shape[0]
        nkw = self.nchan // self.nsp
        out = np.empty((n, nkw, self.nR, 2))
        check(_lib.load().chm_ft_apply_edge(self.h, int(edge), n, rows, out))
        return out

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at definition: module globals may already be gone at interpreter exit
        if _finalizing():
            return
        if getattr(self, "h", None) and _lib._lib is not None:
            _lib._lib.chm_ft_destroy(self.h)
            self.h = None
