"""Per-edge amplitude fit and R-factor of a model chi(R) against data (oracle).

Test infrastructure only (see the package header).  Follows chimera/model.py:497-512
(fit_scales: the closed-form least-squares amplitude per channel, or per edge over all
of its channels when the channel -> edge map is given) and chimera/model.py:535-540
(r_factor: sum |s m - d|^2 / sum |d|^2 over the fit window), for real or complex arrays
(nchan, nR).

Pinned (tests/test_oracle.py) by what the mathematics fixes, not by restating the
formula: the fitted amplitude minimises the residual (any perturbation raises it); a
model that matches the data up to an amplitude gives R = 0 and that amplitude; with
data = s0 m + n and n orthogonal to m on the window the fit returns s0 and
R = |n|^2 / (s0^2 |m|^2 + |n|^2) exactly; rescaling the data by c rescales the
amplitude by c and leaves R unchanged; per-edge fitting equals one fit over the edge's
channels laid end to end.
"""
from __future__ import annotations

import numpy as np


def fit_scales(model, data, mask, edge=None):
    """(nchan,) amplitudes: per channel s = Re<m, d> / <m, m> over the window, or one per
    edge from the summed inner products of its channels, repeated for each channel."""
    m = np.asarray(model)[:, mask]
    d = np.asarray(data)[:, mask]
    num = np.array([np.real(np.vdot(m[c], d[c])) for c in range(m.shape[0])])
    den = np.array([np.real(np.vdot(m[c], m[c])) for c in range(m.shape[0])])
    if edge is not None:
        edge = np.asarray(edge, int)
        num = np.array([num[edge == edge[c]].sum() for c in range(m.shape[0])])
        den = np.array([den[edge == edge[c]].sum() for c in range(m.shape[0])])
    return num / np.maximum(den, 1e-30)


def r_factor(model, data, mask, scales=None):
    """sum over the window of |s m - d|^2 divided by sum |d|^2."""
    m = np.asarray(model)
    if scales is not None:
        m = m * np.asarray(scales)[:, None]
    diff = (m - np.asarray(data))[:, mask]
    num = float(np.sum(diff.real ** 2 + diff.imag ** 2)) if np.iscomplexobj(diff) else float(np.sum(diff ** 2))
    dd = np.asarray(data)[:, mask]
    den = float(np.sum(dd.real ** 2 + dd.imag ** 2)) if np.iscomplexobj(dd) else float(np.sum(dd ** 2))
    return num / max(den, 1e-30)
