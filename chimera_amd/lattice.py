"""Periodic cubic-lattice supercells and their fixed-topology neighbour lists (host setup).

Mirrors exafs_gpu/lattice.py:205-295 (CrystalSupercell, neighbour_list), shell indexing
of chimera/model.py:199-206 and Warren-Cowley of exafs_gpu/lattice.py:167-185.

Sites live on integer points in units of a0/2 (sc, bcc and fcc are all Bravais
lattices there), so the neighbours of every site are the same set of integer offsets:
the list is built from that stencil in O(N * ncoord) instead of the reference's
O(N^2) 27-image search.  The pair order is the reference's: i-major, j ascending,
with `shift` the periodic offset added to pos[j].
"""
from __future__ import annotations

import itertools

import numpy as np

BASIS = {
    "sc": [[0, 0, 0]],
    "bcc": [[0, 0, 0], [1, 1, 1]],
    "fcc": [[0, 0, 0], [0, 1, 1], [1, 0, 1], [1, 1, 0]],
}  # in units of a0/2


def lattice_offsets(lattice, radius_units):
    """Integer lattice vectors (units a0/2) with 0 < |v| <= radius_units."""
    R = int(np.floor(radius_units + 1e-9))
    g = np.array(list(itertools.product(range(-R, R + 1), repeat=3)), np.int64)
    if lattice == "fcc":
        keep = g.sum(1) % 2 == 0
    elif lattice == "bcc":
        keep = ((g % 2 == 0).all(1)) | ((g % 2 == 1).all(1))
    elif lattice == "sc":
        keep = (g % 2 == 0).all(1)
    else:
        raise ValueError(f"unknown lattice {lattice!r}; have sc, bcc, fcc")
    g = g[keep]
    n = np.sqrt((g ** 2).sum(1))
    return g[(n > 1e-9) & (n <= radius_units + 1e-9)]


class CrystalSupercell:
    """Supercell of `ncell` conventional cubic cells (lattice.py:224-251)."""

    def __init__(self, lattice="fcc", ncell=(5, 5, 5), a0=3.56):
        if lattice not in BASIS:
            raise ValueError(f"unknown lattice {lattice!r}; have {list(BASIS)}")
        self.lattice = lattice
        self.ncell = tuple(int(n) for n in ncell)
        self.a0 = float(a0)
        self.unit = self.a0 / 2.0
        basis = np.array(BASIS[lattice], np.int64)
        self.nbasis = len(basis)
        cells = np.array([(i, j, k) for i in range(self.ncell[0])
                          for j in range(self.ncell[1]) for k in range(self.ncell[2])], np.int64)
        self.xyz = (2 * cells[:, None, :] + basis[None]).reshape(-1, 3)   # a0/2 units
        self.natoms = self.xyz.shape[0]
        self.box_units = 2 * np.array(self.ncell, np.int64)
        self.H = np.diag(np.array(self.ncell, float) * self.a0)
        self.box = np.diag(self.H).copy()
        self.ideal = np.ascontiguousarray(self.xyz * self.unit)
        # basis index of a coordinate parity pattern (x&1 | y&1<<1 | z&1<<2)
        self.basis_of_parity = -np.ones(8, np.int32)
        for b, v in enumerate(basis):
            self.basis_of_parity[(v[0] & 1) | ((v[1] & 1) << 1) | ((v[2] & 1) << 2)] = b

    def site_of(self, xyz):
        """Site index of integer coordinates (wrapped into the box)."""
        x = np.mod(np.asarray(xyz, np.int64), self.box_units)
        par = (x[..., 0] & 1) | ((x[..., 1] & 1) << 1) | ((x[..., 2] & 1) << 2)
        b = self.basis_of_parity[par]
        cell = ((x[..., 0] >> 1) * self.ncell[1] + (x[..., 1] >> 1)) * self.ncell[2] + (x[..., 2] >> 1)
        return cell * self.nbasis + b

    def neighbour_list(self, rmax):
        """(i_idx, j_idx, shift) for all ordered pairs with |r_ij| <= rmax."""
        offs = lattice_offsets(self.lattice, rmax / self.unit)
        if 2 * rmax >= self.box.min() - 1e-9:
            raise ValueError("supercell too small: needs box > 2 rmax for a unique minimum image")
        n = self.natoms
        tgt = self.xyz[:, None, :] + offs[None, :, :]                # (n, nv, 3)
        j = self.site_of(tgt)
        wrapped = np.mod(tgt, self.box_units)
        shift_units = tgt - wrapped                                  # multiples of the box
        order = np.argsort(j, axis=1, kind="stable")
        j = np.take_along_axis(j, order, 1)
        shift_units = np.take_along_axis(shift_units, order[..., None], 1)
        i_idx = np.repeat(np.arange(n), offs.shape[0]).astype(np.int32)
        j_idx = j.reshape(-1).astype(np.int32)
        shift = np.ascontiguousarray(shift_units.reshape(-1, 3) * self.unit, np.float64)
        return i_idx, j_idx, shift

    def shells(self, i_idx, j_idx, shift):
        """(shell index per pair, mean radius per shell): distance rounded to 1e-3 A,
        ranked (chimera/model.py:199-206)."""
        d0 = np.linalg.norm((self.ideal[j_idx] + shift) - self.ideal[i_idx], axis=1)
        _, sh = np.unique(np.round(d0, 3), return_inverse=True)
        sh = sh.astype(np.int32)
        radii = np.array([d0[sh == s].mean() for s in range(int(sh.max()) + 1)])
        return sh, radii


def warren_cowley(species, i_idx, j_idx, shell, s, nsp):
    """alpha_AB = 1 - P(B|A)/c_B over the pairs of shell s (lattice.py:167-185)."""
    species = np.asarray(species)
    m = shell == s
    a, b = species[i_idx[m]], species[j_idx[m]]
    N = np.zeros((nsp, nsp))
    np.add.at(N, (a, b), 1.0)
    return alpha_from_counts(N, np.bincount(species, minlength=nsp))


def warren_cowley_radius(species, i_idx, j_idx, shift, cell, shell_r, tol=0.15, comp=None,
                         nsp=None):
    """alpha_AB = 1 - P(B|A)/c_B over the pairs whose ideal-lattice distance lies within
    `tol` of `shell_r` (exafs_gpu/lattice.py:167-185: the shell chosen by radius)."""
    species = np.asarray(species)
    n = int(nsp) if nsp is not None else int(species.max()) + 1
    i_idx, j_idx = np.asarray(i_idx), np.asarray(j_idx)
    d = np.linalg.norm(cell.ideal[j_idx] + np.asarray(shift, float) - cell.ideal[i_idx], axis=1)
    m = np.abs(d - float(shell_r)) < tol
    N = np.zeros((n, n))
    np.add.at(N, (species[i_idx[m]], species[j_idx[m]]), 1.0)
    if comp is None:
        return alpha_from_counts(N, np.bincount(species, minlength=n))
    c = np.asarray(comp, float)
    tot = N.sum(1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(tot > 0, 1.0 - (N / np.maximum(tot, 1)) / c[None, :], 0.0)


def alpha_from_counts(N, natoms_of):
    """Warren-Cowley matrix from ordered pair counts N[a, b] of one shell."""
    N = np.asarray(N, float)
    c = np.asarray(natoms_of, float) / np.sum(natoms_of)
    tot = N.sum(1, keepdims=True)
    with np.errstate(invalid="ignore", divide="ignore"):
        P = np.where(tot > 0, N / np.maximum(tot, 1), 0.0)
        out = np.where(tot > 0, 1.0 - P / c[None, :], 0.0)
    return out


def integer_counts(composition, natoms):
    """Species counts rounded with the last species taking the remainder
    (chimera/system.py:83-90)."""
    comp = np.asarray(composition, float)
    comp = comp / comp.sum()
    cnt = np.round(comp * natoms).astype(int)
    cnt[-1] = natoms - cnt[:-1].sum()
    return cnt


def random_config(natoms, composition, seed=0):
    """chimera/model.py:451-459."""
    cnt = integer_counts(composition, natoms)
    sp = np.repeat(np.arange(len(cnt)), cnt).astype(np.int32)
    np.random.default_rng(seed).shuffle(sp)
    return sp
