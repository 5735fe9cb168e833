"""Pair-count model Jacobian and Cramer-Rao bound in the linear-Gaussian case (oracle).

Follows chimera/shellmodel.py:450-593 for the part a closed form fixes: with the
per-edge amplitude fixed (profile_scale False) and no nuisance parameters, the
normalised residual is exactly linear in the pair counts N, so its Jacobian along a
short-range-order direction d is the transform of sum_{a,b,s} d[a,b,s] B[a,b,s] / N_a,
divided by sigma.  The posterior of the direction coefficients under the Gaussian prior
with precisions P is then N(., (J^T J + P)^-1), computed here by a QR factorisation of
the stacked system [J; sqrt(P)] (a different route from the normal equations).

Pinned by: the QR posterior equals the closed-form (J^T J + P)^-1 on random
well-conditioned systems, and the bound scales linearly with the noise when the prior
vanishes (tests).
"""
from __future__ import annotations

import numpy as np


def sro_jacobian(B, natoms_of, transforms, chan_edge, R, mask, sigma, shells, cplx=True):
    """(nres, ndir) Jacobian of (chi(R) - target)/sigma along the SRO directions
    (a < b) of `shells`; B (nsp, nsp, nshell, nk) the basis, natoms_of the absorber
    counts, transforms one per channel (oracle fourier.Transform), R the output grid."""
    nsp, _, nshell, nk = B.shape
    cols = []
    for s in shells:
        for a in range(nsp):
            for b in range(a + 1, nsp):
                dchi = np.zeros((nsp, nk))
                dchi[a] += (B[a, b, s] - B[a, a, s]) / natoms_of[a]
                dchi[b] += (B[b, a, s] - B[b, b, s]) / natoms_of[b]
                m = np.array([t.dft(dchi[e], R) for t, e in zip(transforms, chan_edge)])
                r = m[:, mask] / sigma[:, mask]
                cols.append(np.concatenate([r.real.ravel(), r.imag.ravel()]) if cplx else r.ravel())
    return np.array(cols).T


def posterior_cov(J, prec):
    """(J^T J + diag(prec))^-1 through QR of the stacked system [J; diag(sqrt(prec))]."""
    M = np.vstack([J, np.diag(np.sqrt(prec))])
    _, Rm = np.linalg.qr(M)
    Ri = np.linalg.solve(Rm, np.eye(Rm.shape[0]))
    return Ri @ Ri.T


def alpha_derivatives(natoms_of, z, nsp, s):
    """d alpha[a,b] / d coefficient of each SRO direction of shell s (alpha = 1 - N_ab /
    (N_a z_s c_b)): a direction (a, b) moves N_ab, N_ba up and N_aa, N_bb down by one."""
    c = np.asarray(natoms_of, float) / np.sum(natoms_of)
    out = []
    for a in range(nsp):
        for b in range(a + 1, nsp):
            dN = np.zeros((nsp, nsp))
            dN[a, b] += 1
            dN[b, a] += 1
            dN[a, a] -= 1
            dN[b, b] -= 1
            out.append(-dN / (np.asarray(natoms_of, float)[:, None] * z[s] * c[None, :]))
    return out
