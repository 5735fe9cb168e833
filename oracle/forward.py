"""chi(k) in pair space and in count space (oracle).

chi_pairs follows ForwardModel.chi_reference (exafs_gpu/forward.py:265-293): for
every ordered pair the term of tables.pair_term, summed by absorber species; with
`normalise` divided by the number of ATOMS of that species (forward.py:183-187,
290-292). k = 0 contributes 0.

Count space follows chimera/shellmodel.py:290-336, 379-413: on the ideal lattice a
pair's term depends only on (absorber a, scatterer b, shell s), so
chi_a = sum_{b,s} N[a,b,s] B[a,b,s](k) (+ MS: sum_e C_e T_e(k), mslibrary.py:453-457).

Adjoints (gradient-based refinement, exafs_gpu/forward.py:190-262 and kernels.py:
166-237, 514-660): grad_pos, chi_soft (relaxed occupancy, bilinear in the site
weights) and grad_soft.

Pinned by: the identity pair space == count space on the lattice (two different
summations of the same terms), the k = 0 convention, and the closed form of a single
isolated pair; the adjoints by central finite differences of chi_pairs / chi_soft and
by chi_soft == chi_pairs for one-hot weights.
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


def _pair_geometry(pos, i_idx, j_idx, shift):
    d = (pos[j_idx] + shift) - pos[i_idx]
    r = np.linalg.norm(d, axis=1)
    return r, d / r[:, None]


def grad_pos(k, f, delta, lam, s02, pos, species, i_idx, j_idx, shift, sig2, gchi, shell=None,
             dr=None, c3=None, s2types=None):
    """dL/dpos (natoms, 3) for L = sum_{a,k} gchi[a, k] chi_unnormalised[a, k]
    (ForwardModel.grad_pos, exafs_gpu/forward.py:209-234; kernels chi_adjoint_r and
    scatter_force, kernels.py:166-237): per pair dT/dr_eff summed over k with the
    absorber's gchi, then -g e to the absorber and +g e to the scatterer."""
    nsp = f.shape[0]
    species = np.asarray(species, np.int64)
    sa, sb = species[i_idx], species[j_idx]
    npairs = i_idx.size
    shell = np.zeros(npairs, int) if shell is None else shell
    nshell = int(shell.max()) + 1
    dr = np.zeros((nshell, nsp, nsp)) if dr is None else dr
    c3 = np.zeros(nshell) if c3 is None else c3
    s2types = np.zeros((nshell, nsp, nsp)) if s2types is None else s2types
    r0, e = _pair_geometry(pos, i_idx, j_idx, shift)
    r = r0 + dr[shell, sa, sb]
    kk = np.asarray(k, float)
    nz = kk > 0
    kz = kk[nz][None, :]
    dflat = delta.reshape(nsp * nsp, -1)
    s2p = np.broadcast_to(sig2, (npairs,)) + s2types[shell, sa, sb]
    gr = np.zeros(npairs)
    for p0 in range(0, npairs, 20000):
        sl = slice(p0, p0 + 20000)
        rr = r[sl][:, None]
        lz = lam[nz][None, :]
        amp = (s02 * f[sb[sl]][:, nz] / (kz * rr ** 2)
               * np.exp(-2.0 * rr / lz - 2.0 * kz ** 2 * s2p[sl][:, None]))
        th = 2.0 * kz * rr + dflat[sa[sl] * nsp + sb[sl]][:, nz] - (4.0 / 3.0) * kz ** 3 * c3[shell[sl]][:, None]
        dT = amp * ((-2.0 / rr - 2.0 / lz) * np.sin(th) + 2.0 * kz * np.cos(th))
        gr[sl] = (dT * gchi[sa[sl]][:, nz]).sum(1)
    gpos = np.zeros_like(pos, dtype=float)
    np.add.at(gpos, j_idx, gr[:, None] * e)
    np.add.at(gpos, i_idx, -gr[:, None] * e)
    return gpos


def chi_soft(k, f, delta, lam, s02, pos, w, i_idx, j_idx, shift, sig2, normalise=True):
    """Relaxed-occupancy chi (nsp, nk) (ForwardModel.chi_soft, forward.py:190-206;
    kernel chi_forward_soft, kernels.py:514-589): every pair contributes
    sum_{a,b} w_i[a] w_j[b] T_ab(r_ij, k), bilinear in the per-site weights w
    (natoms, nsp); normalised by max(sum_i w_i[a], 1).  No cumulant tables."""
    nsp = f.shape[0]
    r, _ = _pair_geometry(pos, i_idx, j_idx, shift)
    kk = np.asarray(k, float)
    nz = kk > 0
    kz = kk[nz][None, :]
    s2p = np.broadcast_to(sig2, (i_idx.size,))
    out = np.zeros((nsp, kk.size))
    for p0 in range(0, i_idx.size, 20000):
        sl = slice(p0, p0 + 20000)
        rr = r[sl][:, None]
        base = s02 / (kz * rr ** 2) * np.exp(-2.0 * rr / lam[nz][None, :] - 2.0 * kz ** 2 * s2p[sl][:, None])
        wi, wj = w[i_idx[sl]], w[j_idx[sl]]
        for a in range(nsp):
            for b in range(nsp):
                T = base * f[b][nz][None, :] * np.sin(2.0 * kz * rr + delta[a, b][nz][None, :])
                out[a, nz] += ((wi[:, a] * wj[:, b])[:, None] * T).sum(0)
    if normalise:
        out /= np.maximum(w.sum(0), 1.0)[:, None]
    return out


def grad_soft(k, f, delta, lam, s02, pos, w, i_idx, j_idx, shift, sig2, gchi, normalise=True):
    """(dL/dpos, dL/dw) of L = sum gchi * chi_soft (ForwardModel.grad_soft,
    forward.py:236-262; kernel chi_adjoint_soft, kernels.py:592-660), including the
    derivative of the 1/max(sum w, 1) normalisation as the reference applies it."""
    nsp = f.shape[0]
    r, e = _pair_geometry(pos, i_idx, j_idx, shift)
    kk = np.asarray(k, float)
    nz = kk > 0
    kz = kk[nz][None, :]
    lz = lam[nz][None, :]
    norm = np.maximum(w.sum(0), 1.0)
    g = gchi / norm[:, None] if normalise else gchi
    s2p = np.broadcast_to(sig2, (i_idx.size,))
    gr = np.zeros(i_idx.size)
    gw = np.zeros_like(w, dtype=float)
    for p0 in range(0, i_idx.size, 20000):
        sl = slice(p0, p0 + 20000)
        rr = r[sl][:, None]
        base = s02 / (kz * rr ** 2) * np.exp(-2.0 * rr / lz - 2.0 * kz ** 2 * s2p[sl][:, None])
        wi, wj = w[i_idx[sl]], w[j_idx[sl]]
        gwi = np.zeros((rr.shape[0], nsp))
        gwj = np.zeros((rr.shape[0], nsp))
        for a in range(nsp):
            ga = g[a][nz][None, :]
            for b in range(nsp):
                th = 2.0 * kz * rr + delta[a, b][nz][None, :]
                T = base * f[b][nz][None, :] * np.sin(th)
                dT = base * f[b][nz][None, :] * ((-2.0 / rr - 2.0 / lz) * np.sin(th) + 2.0 * kz * np.cos(th))
                gT = (ga * T).sum(1)
                gr[sl] += wi[:, a] * wj[:, b] * (ga * dT).sum(1)
                gwi[:, a] += wj[:, b] * gT
                gwj[:, b] += wi[:, a] * gT
        np.add.at(gw, i_idx[sl], gwi)
        np.add.at(gw, j_idx[sl], gwj)
    if normalise:
        chi_un = chi_soft(k, f, delta, lam, s02, pos, w, i_idx, j_idx, shift, sig2, normalise=False)
        gw -= ((gchi * chi_un).sum(1) / norm ** 2)[None, :]
    gpos = np.zeros_like(pos, dtype=float)
    np.add.at(gpos, j_idx, gr[:, None] * e)
    np.add.at(gpos, i_idx, -gr[:, None] * e)
    return gpos, gw
