"""Supercell, neighbour list, shells, reverse index and Warren-Cowley (oracle).

Follows exafs_gpu/lattice.py:205-295 (CrystalSupercell, 27-image minimum image),
chimera/model.py:199-203 (shell index = rank of the distance rounded to 1e-3 A),
methods/method_a_rmc.py:51-60 (reverse index), exafs_gpu/lattice.py:167-185
(Warren-Cowley), chimera/system.py:83-90 (integer_counts) and
chimera/model.py:451-459 (random_config).

Pinned by: fcc shell counts 12/6/24/12 and npairs = 54 N; rev an involution with
i[rev] = j; brute force on tiny cells; alpha = 0 in expectation for a random alloy
and the closed-form alpha of perfectly ordered L1_0 / B2 structures (tests).
"""
from __future__ import annotations

import numpy as np

LATTICES = {
    "sc": [[0, 0, 0]],
    "bcc": [[0, 0, 0], [.5, .5, .5]],
    "fcc": [[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]],
}


class Supercell:
    """Cubic-lattice periodic supercell (exafs_gpu/lattice.py:224-251)."""

    def __init__(self, lattice="fcc", ncell=(5, 5, 5), a0=3.56):
        basis = np.asarray(LATTICES[lattice], float)
        self.lattice = lattice
        self.ncell = tuple(int(n) for n in ncell)
        self.a0 = float(a0)
        A = np.eye(3) * self.a0
        self.H = np.diag(self.ncell).astype(float) @ A
        cells = np.array([(i, j, k) for i in range(self.ncell[0])
                          for j in range(self.ncell[1]) for k in range(self.ncell[2])], float)
        frac_conv = (cells[:, None, :] + basis[None, :, :]).reshape(-1, 3)
        self.ideal = np.ascontiguousarray(frac_conv @ A)
        self.natoms = self.ideal.shape[0]
        self.box = np.diag(self.H).copy()
        self._images = np.array([(i, j, k) for i in (-1, 0, 1)
                                 for j in (-1, 0, 1) for k in (-1, 0, 1)], float) @ self.H

    def neighbour_list(self, rmax):
        """Ordered pairs (i, j), i-major, j ascending; shift added to pos[j]
        (exafs_gpu/lattice.py:260-279)."""
        pos = self.ideal
        n = self.natoms
        I, J, SH = [], [], []
        for i in range(n):
            d0 = pos - pos[i]
            cand = d0[:, None, :] + self._images[None, :, :]
            n2 = (cand ** 2).sum(-1)
            best = n2.argmin(1)
            d = cand[np.arange(n), best]
            sh = self._images[best]
            r = np.linalg.norm(d, axis=1)
            sel = np.where((r <= rmax) & (r > 1e-6))[0]
            I.append(np.full(sel.size, i))
            J.append(sel)
            SH.append(sh[sel])
        i_idx = np.concatenate(I).astype(np.int32)
        j_idx = np.concatenate(J).astype(np.int32)
        shift = np.ascontiguousarray(np.concatenate(SH), np.float64)
        return i_idx, j_idx, shift


def pair_distances(pos, i_idx, j_idx, shift):
    return np.linalg.norm((pos[j_idx] + shift) - pos[i_idx], axis=1)


def shell_index(ideal, i_idx, j_idx, shift):
    """Shell of each pair: rank of its distance rounded to 1e-3 A, and the mean
    radius of each shell (chimera/model.py:199-206)."""
    d0 = pair_distances(ideal, i_idx, j_idx, shift)
    _, sh = np.unique(np.round(d0, 3), return_inverse=True)
    sh = sh.astype(np.int32)
    radii = np.array([d0[sh == s].mean() for s in range(int(sh.max()) + 1)])
    return sh, radii


def reverse_index(i_idx, j_idx, natoms):
    """rev[p] = index of (j, i) for p = (i, j) (methods/method_a_rmc.py:51-60)."""
    key = i_idx.astype(np.int64) * natoms + j_idx
    rkey = j_idx.astype(np.int64) * natoms + i_idx
    order = np.argsort(key)
    pos = np.searchsorted(key[order], rkey)
    rev = order[np.clip(pos, 0, key.size - 1)]
    assert np.array_equal(i_idx[rev], j_idx) and np.array_equal(j_idx[rev], i_idx)
    return rev.astype(np.int32)


def warren_cowley(species, i_idx, shell, j_idx, s, nsp):
    """alpha_AB = 1 - P(B|A)/c_B over the pairs of shell s
    (exafs_gpu/lattice.py:167-185, with the shell selected by index)."""
    m = shell == s
    a, b = species[i_idx[m]], species[j_idx[m]]
    c = np.bincount(species, minlength=nsp) / species.size
    alpha = np.zeros((nsp, nsp))
    for A in range(nsp):
        sel = a == A
        tot = int(sel.sum())
        if tot == 0:
            continue
        for B in range(nsp):
            alpha[A, B] = 1.0 - ((b[sel] == B).sum() / tot) / c[B]
    return alpha


def integer_counts(composition, natoms):
    """chimera/system.py:83-90."""
    comp = np.asarray(composition, float)
    comp = comp / comp.sum()
    cnt = np.round(comp * natoms).astype(int)
    cnt[-1] = natoms - cnt[:-1].sum()
    return cnt


def random_config(natoms, composition, seed):
    """chimera/model.py:451-459: repeat species by integer counts, PCG64 shuffle."""
    cnt = integer_counts(composition, natoms)
    sp = np.repeat(np.arange(len(cnt)), cnt).astype(np.int32)
    np.random.default_rng(seed).shuffle(sp)
    return sp
