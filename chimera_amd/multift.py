"""MultiEdgeFT: one transform per (edge, k weight) channel, edge-major.

Drop-in for chimera/multift.py:25-84.  All channels go through one device object: the
DFT matrices of every channel, applied to a batch of (nrows, nsp, nk) chi in one GEMM
per edge.  `chi_extra` (nsp, nk) is added before transforming (multift.py:58-61).
"""
from __future__ import annotations

import numpy as np

from .fourier import _DeviceFT


class MultiEdgeFT:
    def __init__(self, fts, chi_extra=None, chan_edge=None):
        self.fts = list(fts)
        self.chi_extra = None if chi_extra is None else np.ascontiguousarray(chi_extra, np.float64)
        R = self.fts[0].R
        for f in self.fts[1:]:
            assert np.allclose(f.R, R), "channels must share the R grid"
        self.R = R
        self.nchan = len(self.fts)
        self.chan_edge = (np.arange(self.nchan, dtype=np.int32) if chan_edge is None
                          else np.asarray(chan_edge, np.int32))
        assert self.chan_edge.shape == (self.nchan,)
        self.nsp = int(self.chan_edge.max()) + 1
        self.kmin = min(f.kmin for f in self.fts)
        self.kmax = max(f.kmax for f in self.fts)
        self.kweights = tuple(dict.fromkeys(f.kweight for f in self.fts))
        self.kweight = self.fts[0].kweight
        self.xp = np
        self._dev = None

    @classmethod
    def from_spectrum(cls, S):
        return cls(S.ft, chi_extra=getattr(S, "chi_ms", None),
                   chan_edge=getattr(S, "chan_edge", None))

    def device(self):
        if self._dev is None:
            self._dev = _DeviceFT(self.fts, self.chan_edge)
        return self._dev

    def to_R(self, chi):
        """chi (..., nsp, nk) -> complex (..., nchan, nR)."""
        chi = np.asarray(chi, float)
        lead = chi.shape[:-2]
        out = self.device().apply(chi.reshape((-1,) + chi.shape[-2:]), self.chi_extra)
        return out.reshape(lead + (self.nchan, self.R.size))

    def magnitude(self, chi):
        return np.abs(self.to_R(chi))

    def observe(self, chi, complex_out=False):
        return self.to_R(chi) if complex_out else self.magnitude(chi)

    def window_mask(self, rmin, rmax):
        return (self.R >= rmin) & (self.R <= rmax)

    def n_independent(self, rmin, rmax):
        return 2 * (self.kmax - self.kmin) * (rmax - rmin) / np.pi + 2
