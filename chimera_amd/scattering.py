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


HARTREE_EV = 27.211386245988      # CODATA 2018
BOHR_A = 0.529177210903


def columns_to_record(cols, reff, deg=1.0, nleg=2, scatterer=None, header=""):
    """One path's feffNNNN.dat columns in this pipeline's convention
    (exafs_gpu/scattering.py:158-178).

    Columns: k, real[2*phc], mag[feff], phase[feff], red factor, lambda, real[p].  FEFF
    writes chi = amp/(k R^2) exp(-2R/lambda) sin(2kR + pha) with k measured from the
    Fermi level, amp = mag * red and pha = 2*phc + phase.  The path sum here runs on the
    internal momentum p, so the same term is rewritten on p: f(p) = amp p / k,
    delta(p) = unwrap(pha) - 2 (p - k) R_eff (the 2kR of FEFF is 2pR - 2(p-k)R), lambda
    unchanged.  Points with k <= 0.3 are dropped (the conversion divides by k).  The
    record also carries the energy of the p axis at k = 0 (the calculation's E0 offset)."""
    a = np.asarray(cols, np.float64)
    keep = a[:, 0] > 0.3
    k, phc, mag, pha, red, lam, p = (a[keep, c] for c in range(7))
    f = mag * red * p / k
    delta = np.unwrap(phc + pha) - 2.0 * (p - k) * float(reff)
    p0 = float(np.interp(0.0, k, p))
    e0 = p0 * p0 * HARTREE_EV * BOHR_A ** 2 / 2.0
    return dict(p=p, k=k, f=f, delta=delta, lam=lam, reff=float(reff), deg=float(deg),
                nleg=int(nleg), scatterer=scatterer, e0_offset_ev=e0, header=header)


def read_feffdat(fn):
    """Parse one feffNNNN.dat (exafs_gpu/scattering.py:181-190): the path line after the
    "nleg, deg, reff" label, the atom list, and the seven data columns under the
    "k  real[2*phc] ..." header."""
    with open(fn) as fh:
        lines = fh.read().split("\n")
    geo = next(i for i, ln in enumerate(lines) if "nleg, deg, reff" in ln)
    hdr = next(i for i, ln in enumerate(lines)
               if ln.strip().startswith("k ") and "real[2*phc]" in ln)
    nleg, deg, reff = (float(x) for x in lines[geo].split()[:3])
    atoms = [ln.split() for ln in lines[geo + 2:hdr]]
    scatterer = atoms[1][5] if len(atoms) > 1 else None
    rows = [[float(x) for x in ln.split()] for ln in lines[hdr + 1:] if ln.strip()]
    return columns_to_record(np.array(rows), reff, deg, int(nleg), scatterer,
                             header="\n".join(lines[:geo]))
