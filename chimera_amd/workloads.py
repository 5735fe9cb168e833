"""The benchmark workloads: supercell, tables, paths, synthetic target data, sampler.

  fcc_fe50ni50_500   fcc Fe50Ni50, 5x5x5 cells (500 atoms), a0 3.59 A, single
                     scattering to 5.2 A, 2 edges, k^1..3, 8 replicas
  fcc_crconi_4000    fcc CrCoNi, 10x10x10 cells (4000 atoms), a0 3.56 A, single
                     scattering to 5.5 A plus all 3-/4-leg paths to 5.5 A, 3 edges,
                     k^1..3, 32 replicas

Common: synthetic tables on k = 0..16 A^-1 step 0.05; S0^2 0.9; sigma^2 0.006 A^2 for
every pair; Hanning window 3-14 A^-1 (dk 1); complex chi(R) on R = 1.4..5.0 A step
0.02; target from a seeded reference configuration with first-shell alpha -0.08 plus
Gaussian noise of sd 0.002 k^-2; target / sigma built as chimera/fitrun.py:267-288
(the random configuration of seed 0 sets the per-edge scale and the model-error sigma).

Everything that evaluates the model runs on the device; the host holds the setup
(lattice, paths, the ordering walk that makes the reference configuration).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .forward import ForwardModel
from .fourier import FourierTransform
from .lattice import CrystalSupercell, integer_counts, random_config, warren_cowley
from .method_a_rmc import RMCSampler
from .mspaths import MSLibrary
from .multift import MultiEdgeFT
from .scattering import k_grid, synthetic


@dataclass
class Workload:
    name: str
    lattice: str
    ncell: tuple
    a0: float
    elements: tuple
    rmax: float
    ms: bool
    nrep: int
    kweights: tuple = (1, 2, 3)
    kmin: float = 3.0
    kmax: float = 14.0
    dk_win: float = 1.0
    rmin_fit: float = 1.4
    rmax_fit: float = 5.0
    dR: float = 0.02
    sigma2: float = 0.006
    alpha_target: float = -0.08
    noise: float = 0.002
    seed: int = 0
    edges: tuple = None      # measured edges (species indices); None = every species


WORKLOADS = {
    "fcc_fe50ni50_500": Workload("fcc_fe50ni50_500", "fcc", (5, 5, 5), 3.59, ("Fe", "Ni"),
                                 5.2, False, 8),
    "fcc_crconi_4000": Workload("fcc_crconi_4000", "fcc", (10, 10, 10), 3.56,
                                ("Cr", "Co", "Ni"), 5.5, True, 32),
}


def fit_R_grid(w: Workload):
    n = int(round((w.rmax_fit - w.rmin_fit) / w.dR)) + 1
    return w.rmin_fit + w.dR * np.arange(n)


def order_to_alpha(species, i_idx, j_idx, shell, nsp, target, seed, tol=0.002, max_iter=400000):
    """Swap unlike sites until the mean first-shell alpha of the unlike pairs reaches
    `target` (a seeded walk: a swap is kept only if it moves alpha toward the target;
    the walk stops within `tol` or at the first step across it)."""
    sp = np.array(species, np.int32)
    rng = np.random.default_rng(seed)
    natoms = sp.size
    m = shell == 0
    nb = j_idx[m].reshape(natoms, -1)            # first-shell neighbours, i-major

    def unlike_alpha(s):
        A = warren_cowley(s, i_idx, j_idx, shell, 0, nsp)
        return A[~np.eye(nsp, dtype=bool)].mean()

    def like(x, s_x, skip):
        return int(np.count_nonzero((sp[nb[x]] == s_x) & (nb[x] != skip)))

    cur = unlike_alpha(sp)
    want_fewer_like = cur > target
    for _ in range(max_iter):
        if abs(cur - target) < tol or (cur > target) != want_fewer_like:
            break
        a, b = (int(x) for x in rng.integers(0, natoms, 2))
        if sp[a] == sp[b]:
            continue
        before = like(a, sp[a], b) + like(b, sp[b], a)
        after = like(a, sp[b], b) + like(b, sp[a], a)
        if after != before and (after < before) == want_fewer_like:
            sp[a], sp[b] = sp[b], sp[a]
            cur = unlike_alpha(sp)
    return sp, cur


class Problem:
    """A built workload: model objects on the device, target and sigma, sampler."""

    def __init__(self, w: Workload, nrep=None, seed=None, device=0):
        self.w = w
        nrep = w.nrep if nrep is None else int(nrep)
        seed = w.seed if seed is None else int(seed)
        self.cell = CrystalSupercell(w.lattice, w.ncell, w.a0)
        self.i, self.j, self.shift = self.cell.neighbour_list(w.rmax)
        self.shell, self.radii = self.cell.shells(self.i, self.j, self.shift)
        self.tables = synthetic(w.elements, k_grid())
        self.k = self.tables.k
        self.nsp = len(w.elements)
        self.sig2 = np.full(self.i.size, w.sigma2)
        self.fm = ForwardModel(self.cell, self.tables, self.i, self.j, self.shift,
                               sigma2=self.sig2, shell=self.shell, device=device)
        self.ms = MSLibrary(self.cell, self.nsp, w.rmax, sigma2_nn=w.sigma2,
                            r_nn=float(self.radii[0])) if w.ms else None
        self.R = fit_R_grid(w)
        fts = [FourierTransform(self.k, w.kmin, w.kmax, kw, dk_win=w.dk_win, R=self.R)
               for e in range(self.nsp) for kw in w.kweights]
        self.chan_edge = np.repeat(np.arange(self.nsp), len(w.kweights)).astype(np.int32)
        self.ft = MultiEdgeFT(fts, chan_edge=self.chan_edge)
        self.mask = np.ones(self.R.size, bool)
        self.nchan = len(fts)
        comp = np.ones(self.nsp) / self.nsp
        self.cfg0 = random_config(self.cell.natoms, comp, 0)
        # placeholder target to get the device basis; replaced below
        zt = np.zeros((self.nchan, self.R.size), complex)
        self.sampler = RMCSampler(self.fm, self.ft, zt, self.mask, np.ones(zt.shape), nrep=nrep,
                                  seed=seed, swap_frac=1.0, ms_library=self.ms)
        self.sampler.init_state(self.cfg0, self.cell.ideal)
        # reference configuration and noisy data
        self.ref_species, self.ref_alpha = order_to_alpha(
            random_config(self.cell.natoms, comp, 12345), self.i, self.j, self.shell, self.nsp,
            w.alpha_target, seed=777)
        cnt = np.bincount(self.ref_species, minlength=self.nsp)
        chi_ref = self.sampler.chi_k_of(self.ref_species[None])[0] / np.maximum(cnt, 1)[:, None]
        rng = np.random.default_rng(2024)
        sd = np.zeros_like(self.k)
        sd[self.k > 0] = w.noise / self.k[self.k > 0] ** 2
        self.chi_data = chi_ref + rng.standard_normal(chi_ref.shape) * sd[None, :]
        self.target, self.sigma, self.scales0 = self._prepare()
        self.sampler.set_target(self.target, self.sigma)
        self.sampler.init_state(self.cfg0, self.cell.ideal)

    def _prepare(self):
        """chimera/fitrun.py:267-288: target = data / per-edge scale of the random
        configuration; sigma = that configuration's RMS residual / scale."""
        tgt = self.ft.to_R(self.chi_data[None])[0]                     # (nchan, nR)
        cnt = np.bincount(self.cfg0, minlength=self.nsp)
        chi0 = self.sampler.chi_k_of(self.cfg0[None])[0] / np.maximum(cnt, 1)[:, None]
        m0 = self.ft.to_R(chi0[None])[0]
        scales = fit_scales(m0, tgt, self.mask, self.chan_edge)
        resid0 = (m0 * scales[:, None] - tgt)[:, self.mask]
        sig_data = np.sqrt((np.abs(resid0) ** 2).mean(1) / 2.0)
        sig = sig_data / scales
        if self.w.edges is not None:
            # an element that is not a measured edge still scatters; its channels carry
            # no data: zero weight (the sampler's basis is per absorber species)
            measured = np.isin(self.chan_edge, np.asarray(self.w.edges))
            sig = np.where(measured, sig, 1e30)
        target = tgt / scales[:, None]
        sigR = np.repeat(sig[:, None], self.R.size, axis=1)
        return target, sigR, scales


def fit_scales(model, data, mask, edge=None):
    """chimera/model.py:497-512: least-squares amplitude per edge."""
    m, d = model[:, mask], data[:, mask]
    num = np.real((np.conj(m) * d).sum(axis=1))
    den = np.real((np.conj(m) * m).sum(axis=1))
    if edge is not None:
        edge = np.asarray(edge, int)
        ne = int(edge.max()) + 1
        num = np.bincount(edge, num, ne)[edge]
        den = np.bincount(edge, den, ne)[edge]
    return num / np.maximum(den, 1e-30)


def r_factor(model, data, mask, scales=None):
    """chimera/model.py:535-540."""
    m = model if scales is None else model * np.asarray(scales)[:, None]
    num = (np.abs(m[:, mask] - data[:, mask]) ** 2).sum()
    den = (np.abs(data[:, mask]) ** 2).sum()
    return float(num / max(den, 1e-30))


def build(name, **kw):
    return Problem(WORKLOADS[name], **kw)
