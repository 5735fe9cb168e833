"""chi(k) in pair space and in count space (oracle).

chi_pairs follows ForwardModel.chi_reference (exafs_gpu/forward.py:265-293): for
every ordered pair the term of tables.pair_term, summed by absorber species; with
`normalise` divided by the number of ATOMS of that species (forward.py:183-187,
290-292). k = 0 contributes 0.

Count space follows chimera/shellmodel.py:290-336, 379-413: on the ideal lattice a
pair's term depends only on (absorber a, scatterer b, shell s), so
chi_a = sum_{b,s} N[a,b,s] B[a,b,s](k) (+ MS: sum_e C_e T_e(k), mslibrary.py:453-457).

Pinned by: the identity pair space == count space on the lattice (two different
summations of the same terms), the k = 0 convention, and the closed form of a single
isolated pair.
"""
from __future__ import annotations

import numpy as np


def chi_pairs(k, f, delta, lam, s02, pos, species, i_idx, j_idx, shift, sig2, shell=None,
              dr=None, c3=None, s2types=None, normalise=True):
    """(nsp, nk) chi of one configuration, pair by pair (forward.py:265-293)."""
    nsp = f.shape[0]
    species = np.asarray(species, np.int64)
    d = (pos[j_idx] + shift) - pos[i_idx]
    sa, sb = species[i_idx], species[j_idx]
    t = sa * nsp + sb
    npairs = i_idx.size
    shell = np.zeros(npairs, int) if shell is None else shell
    nshell = int(shell.max()) + 1
    dr = np.zeros((nshell, nsp, nsp)) if dr is None else dr
    c3 = np.zeros(nshell) if c3 is None else c3
    s2types = np.zeros((nshell, nsp, nsp)) if s2types is None else s2types
    r = np.linalg.norm(d, axis=1) + dr[shell, sa, sb]
    kk = np.asarray(k, float)
    nz = kk > 0
    kz = kk[nz]
    dflat = delta.reshape(nsp * nsp, -1)
    out = np.zeros((nsp, kk.size))
    s2p = np.broadcast_to(sig2, (npairs,)) + s2types[shell, sa, sb]
    ph3 = (4.0 / 3.0) * kz[None, :] ** 3 * c3[shell][:, None]
    # chunk over pairs to bound memory
    for p0 in range(0, npairs, 20000):
        sl = slice(p0, p0 + 20000)
        rr = r[sl][:, None]
        amp = (s02 * f[sb[sl]][:, nz] / (kz[None, :] * rr ** 2)
               * np.exp(-2.0 * rr / lam[nz][None, :] - 2.0 * kz[None, :] ** 2 * s2p[sl][:, None]))
        term = amp * np.sin(2.0 * kz[None, :] * rr + dflat[t[sl]][:, nz] - ph3[sl])
        for s in range(nsp):
            m = sa[sl] == s
            out[s, nz] += term[m].sum(0)
    if normalise:
        cnt = np.maximum(np.bincount(species, minlength=nsp), 1)
        out /= cnt[:, None]
    return out


def pair_counts(species, i_idx, j_idx, shell, nsp, nshell):
    """N[a, b, s] ordered pair counts (shellmodel.py:379-386)."""
    species = np.asarray(species, np.int64)
    N = np.zeros((nsp, nsp, nshell), np.int64)
    np.add.at(N, (species[i_idx], species[j_idx], shell), 1)
    return N


def ss_basis(k, f, delta, lam, s02, radii, sig2_shell, dr=None, c3=None, s2types=None):
    """B[a, b, s](k): the term of one (a, b, shell s) pair (shellmodel.py:290-336)."""
    from .tables import pair_term
    nsp = f.shape[0]
    nshell = len(radii)
    dr = np.zeros((nshell, nsp, nsp)) if dr is None else dr
    c3 = np.zeros(nshell) if c3 is None else c3
    s2types = np.zeros((nshell, nsp, nsp)) if s2types is None else s2types
    B = np.zeros((nsp, nsp, nshell, len(k)))
    for s in range(nshell):
        for a in range(nsp):
            for b in range(nsp):
                B[a, b, s] = pair_term(k, radii[s] + dr[s, a, b], f[b], delta[a, b], lam,
                                       sig2_shell[s] + s2types[s, a, b], s02, c3[s])
    return B


def chi_counts(N, B, C=None, T=None, nblk=None, natoms_of=None):
    """Unnormalised (or normalised by natoms_of) chi (nsp, nk) from counts."""
    nsp = N.shape[0]
    out = np.einsum("abs,absk->ak", N.astype(float), B)
    if C is not None:
        for a in range(nsp):
            out[a] += C[a * nblk:(a + 1) * nblk].astype(float) @ T[a * nblk:(a + 1) * nblk]
    if natoms_of is not None:
        out = out / np.maximum(natoms_of, 1)[:, None]
    return out
