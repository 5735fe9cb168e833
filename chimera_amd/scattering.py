"""Scattering tables (host container) and the synthetic tables of the workloads.

ScatteringTables mirrors exafs_gpu/scattering.py:49-106 (f by scatterer, delta by
[absorber, scatterer], lambda, s02 on one k grid).  `synthetic` builds the seeded
synthetic form the workloads prescribe in place of FEFF output:

    f_b(k) = Z_b/26 exp(-(k-4)^2/18),  delta_ab(k) = -0.9 k + 1.2 + 0.05 Z_b,
    lambda(k) = 6 + 0.3 k,  S0^2 = 0.9

(the phase takes the scatterer's Z, mirroring the amplitude).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

Z_TABLE = {"Ti": 22, "V": 23, "Cr": 24, "Mn": 25, "Fe": 26, "Co": 27, "Ni": 28, "Cu": 29,
           "Zn": 30, "Al": 13, "Nb": 41, "Mo": 42, "Pd": 46, "Ta": 73, "W": 74, "Pt": 78,
           "Au": 79}


@dataclass
class ScatteringTables:
    k: np.ndarray
    f: np.ndarray
    delta: np.ndarray
    lam: np.ndarray
    s02: float = 0.85

    def __post_init__(self):
        n = self.f.shape[0]
        assert self.f.shape == (n, self.k.size), "f must be (nsp, nk)"
        assert self.delta.shape == (n, n, self.k.size), "delta must be (nsp, nsp, nk)"
        assert self.lam.shape == self.k.shape

    @property
    def nsp(self):
        return self.f.shape[0]

    def as_float64(self):
        return (np.ascontiguousarray(self.k, np.float64), np.ascontiguousarray(self.f, np.float64),
                np.ascontiguousarray(self.delta, np.float64),
                np.ascontiguousarray(self.lam, np.float64))


def k_grid(kmax=16.0, dk=0.05):
    return np.linspace(0.0, kmax, int(round(kmax / dk)) + 1)


def synthetic(elements, k=None, s02=0.9):
    """The workloads' seeded synthetic tables for `elements` (symbols or Z values)."""
    k = k_grid() if k is None else np.asarray(k, float)
    zs = np.array([Z_TABLE[e] if isinstance(e, str) else float(e) for e in elements], float)
    nsp = zs.size
    f = np.exp(-(k - 4.0) ** 2 / 18.0)[None, :] * (zs[:, None] / 26.0)
    delta = np.broadcast_to((-0.9 * k + 1.2)[None, None, :] + 0.05 * zs[None, :, None],
                            (nsp, nsp, k.size)).copy()
    lam = 6.0 + 0.3 * k
    return ScatteringTables(k=k, f=f, delta=delta, lam=lam, s02=float(s02))
