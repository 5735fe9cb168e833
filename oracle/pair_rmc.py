"""The reference sampler with every move it makes: swaps and single-atom displacements in
pair space, the Boltzmann energy prior, complex or |chi(R)| targets, any start (oracle).

THIS MODULE IS TEST INFRASTRUCTURE (the checker of the device pair-space sampler).

Follows methods/method_a_rmc.py and exafs_gpu/kernels.py:
  * proposal: swap flag u < swap_frac, site a, up to 8 partner candidates of another
    species; no partner (or no swap flag) -> displacement of a by a Gaussian step of
    disp_sigma (Box-Muller, two pairs of draws), the offset from the ideal site clamped
    to max_disp                                           kernels.py:362-423
  * affected directed pairs with dedup, their new distance and type
                                                          kernels.py:425-488
  * Delta chi: new term minus old term over the affected pairs, by absorber
                                                          kernels.py:243-325
  * MS delta of a swap through the per-site incidence   chimera/mslibrary.py:466-486
  * energy prior: V/2 per directed pair by (shell, type), weighted by 1/(kB T), not
    tempered                                             method_a_rmc.py:147-157, 310-326
  * chi^2 of a complex target (Re, Im separate residuals) or of |chi(R)|, profiled
    per-edge scale                                        method_a_rmc.py:180-227
  * Metropolis on x = beta dchi2 nfit + beta_E dE        method_a_rmc.py:278-285
  * commit (species of a and b; position of a for a displacement), best retention
    with positions                                        kernels.py:490-510,
                                                          method_a_rmc.py:287-307
  * replica exchange permuting whole states             method_a_rmc.py:332-373

Decisions shared with the device sampler (the reference's own cannot be replayed
without CuPy): the acceptance and exchange uniforms are the counter-based streams of
oracle.rmc, and a proposal's stream is keyed by the rung.

Parity: Delta chi is pinned by the full recompute (chi of the moved configuration minus
chi of the old one, summed pair by pair over all pairs: a dropped or doubled pair fails
it); the energy change by the energy difference of the two configurations; the
displacement's clamp by its closed form; the magnitude chi^2 by the minimum over the
scale (perturbation) and its value at the exact model (0).
"""
from __future__ import annotations

import math

import numpy as np

from .rmc import (GOLDEN, M64, accept_uniform, affected_pairs, exchange_uniforms, splitmix64,
                  u01)

KB_EV = 8.617333262e-5   # eV / K


def propose_move(species, pos, ideal, seed, stepc, c, swap_frac, disp_sigma, max_disp):
    """propose_select (kernels.py:362-423) for rung key c.

    Returns (a, b, new species of a, new species of b, is_swap, new position of a).  A
    displacement has b == a and unchanged species; a swap keeps a's position."""
    natoms = species.size
    h = splitmix64((seed ^ ((stepc * GOLDEN) & M64) ^ ((c << 32) & M64)) & M64)
    h = splitmix64(h)
    sw = u01(h) < swap_frac
    h = splitmix64(h)
    a = min(int(u01(h) * natoms), natoms - 1)
    sa = int(species[a])
    b, sb = a, sa
    if sw:
        for _ in range(8):
            h = splitmix64(h)
            cand = min(int(u01(h) * natoms), natoms - 1)
            if int(species[cand]) != sa:
                b, sb = cand, int(species[cand])
                break
    if sw and b != a:
        return a, b, sb, sa, 1, pos[a].copy()
    g = [0.0, 0.0, 0.0]
    for t in (0, 2):
        h = splitmix64(h)
        u1 = u01(h)
        h = splitmix64(h)
        u2 = u01(h)
        rad = math.sqrt(-2.0 * math.log(max(u1, 1e-300)))
        g[t] = rad * math.cos(6.283185307179586 * u2)
        if t + 1 < 3:
            g[t + 1] = rad * math.sin(6.283185307179586 * u2)
    o = [(float(pos[a, d]) + g[d] * disp_sigma) - float(ideal[a, d]) for d in range(3)]
    m = math.sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])
    sc = max_disp / m if m > max_disp else 1.0
    newp = np.array([float(ideal[a, d]) + o[d] * sc for d in range(3)])
    return a, a, sa, sa, 0, newp


def distances(pos, i_idx, j_idx, shift):
    """|pos[j] + shift - pos[i]| per pair, each operation rounded on its own."""
    d = (pos[j_idx] + shift) - pos[i_idx]
    return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])


def energy_table(V, T_ann, nshell, nsp):
    """(nshell * nsp * nsp,) V/2 per directed pair, and 1/(kB T) (method_a_rmc.py:147-157)."""
    V = np.asarray(V, float)
    tab = np.zeros((nshell, nsp * nsp))
    m = min(nshell, V.shape[0])
    tab[:m] = 0.5 * V[:m].reshape(m, nsp * nsp)
    return tab.reshape(-1), 1.0 / (KB_EV * float(T_ann))


class PairSampler:
    """methods/method_a_rmc.py:63-496 replica by replica on the CPU, pair space."""

    def __init__(self, prob, nrep=8, tmin=0.02, tmax=5.0, seed=0, swap_frac=0.75,
                 disp_sigma=0.015, max_disp=0.30, profile_scale=True, energy_prior=None,
                 magnitude=False, target_abs=None):
        """prob: oracle.rmc.Problem (tables, topology, transforms, window, MS).  With
        `magnitude` the target is |chi(R)| (`target_abs`, (nchan, nwin) real on the
        window; default |prob.tm|) and Re/Im are not separate residuals."""
        self.P = prob
        self.nrep = nrep
        self.beta = 1.0 / np.geomspace(tmin, tmax, nrep)
        self.seed = int(seed)
        self.swap_frac, self.disp_sigma, self.max_disp = float(swap_frac), float(disp_sigma), float(max_disp)
        self.profile = bool(profile_scale)
        self.magnitude = bool(magnitude)
        if self.magnitude:
            self.tm = np.abs(prob.tm) if target_abs is None else np.asarray(target_abs, float)
            self.nfit = int(prob.tm.shape[1]) * prob.nchan
        else:
            self.tm = prob.tm
            self.nfit = prob.nfit
        self.sm = prob.sm
        nsp = prob.nsp
        self.nshell = int(np.asarray(prob.shell).max()) + 1
        self.etab, self.ebeta = (None, 0.0)
        if energy_prior is not None:
            self.etab, self.ebeta = energy_table(energy_prior[0], energy_prior[1], self.nshell, nsp)
        self.stepc = 0
        self.iex = 0
        self.log = []

    # ------------------------------------------------------------------ state
    def chi_full(self, species, pos):
        """Unnormalised chi (nsp, nk) of one configuration from scratch (pair sum + MS)."""
        from .forward import chi_pairs
        P = self.P
        out = chi_pairs(P.k, P.f, P.delta, P.lam, P.s02, pos, species, P.i, P.j, P.shift, P.sig2,
                        shell=P.shell, dr=P.dr, c3=P.c3, s2types=P.s2types, normalise=False)
        if P.ms is not None:
            C = P.ms.counts(species)
            nb = P.ms.nblk
            for a in range(P.nsp):
                out[a] += C[a * nb:(a + 1) * nb].astype(float) @ P.T[a * nb:(a + 1) * nb]
        return out

    def init_state(self, species0, pos0):
        P = self.P
        n = self.nrep
        self.species = np.repeat(np.asarray(species0, np.int64)[None], n, 0)
        self.pos = np.repeat(np.asarray(pos0, float)[None], n, 0)
        r0 = distances(self.pos[0], P.i, P.j, P.shift)
        self.r = np.repeat(r0[None], n, 0)
        sp = self.species[0]
        self.ptype = np.repeat((sp[P.i] * P.nsp + sp[P.j])[None], n, 0)
        chi0 = self.chi_full(self.species[0], self.pos[0])
        self.chi = np.repeat(chi0[None], n, 0)
        self.cnt = np.bincount(self.species[0], minlength=P.nsp).astype(float)
        c2 = self.chi2_of_chi(chi0)
        self.chi2 = np.full(n, c2)
        self.best_chi2 = self.chi2.copy()
        self.best_sp = self.species.copy()
        self.best_pos = self.pos.copy()

    # ---------------------------------------------------------------- pieces
    def observe(self, chi_un):
        """The model on the target's footing over the window: complex or |.|."""
        m = self.P.to_R(chi_un / np.maximum(self.cnt, 1.0)[:, None])
        return np.abs(m) if self.magnitude else m

    def profiled_scale(self, m):
        """Per-edge least-squares amplitude, jointly over the edge's channels
        (method_a_rmc.py:190-209)."""
        P = self.P
        if self.magnitude:
            num = (m * self.tm).sum(-1)
            den = (m * m).sum(-1)
        else:
            num = (m.real * self.tm.real + m.imag * self.tm.imag).sum(-1)
            den = (m.real ** 2 + m.imag ** 2).sum(-1)
        num = num.reshape(P.nsp, P.nkw).sum(-1)
        den = den.reshape(P.nsp, P.nkw).sum(-1)
        return np.repeat(num / np.maximum(den, 1e-30), P.nkw)

    def chi2_of_chi(self, chi_un):
        """method_a_rmc.py:211-227."""
        m = self.observe(chi_un)
        if self.profile:
            m = m * self.profiled_scale(m)[:, None]
        d = (m - self.tm) / self.sm
        sq = (d.real ** 2 + d.imag ** 2) if not self.magnitude else d * d
        return float(sq.sum() / self.nfit)

    def terms(self, pairs, r, t):
        """Sum over `pairs` of the term at distance r and type t, by absorber (nsp, nk)
        (the per-pair expression of kernels.py:282-304)."""
        P = self.P
        nsp = P.nsp
        k = P.k
        nz = k > 0
        kz = k[nz][None, :]
        sa, sb = t // nsp, t % nsp
        sh = P.shell[pairs]
        re = (r + P.dr[sh, sa, sb])[:, None]
        s2 = (P.sig2[pairs] + P.s2types[sh, sa, sb])[:, None]
        amp = (P.s02 * P.f[sb][:, nz] / (kz * re * re)
               * np.exp(-2.0 * re / P.lam[nz][None, :] - 2.0 * kz * kz * s2))
        term = amp * np.sin(2.0 * kz * re + P.delta[sa, sb][:, nz]
                            - (4.0 / 3.0) * kz ** 3 * P.c3[sh][:, None])
        out = np.zeros((nsp, k.size))
        for s in range(nsp):
            out[s, nz] = term[sa == s].sum(0)
        return out

    def move_pairs(self, c, a, b, na, nb, newp):
        """Affected pairs of replica c's move, their new distance and type
        (propose_build, kernels.py:425-488)."""
        P = self.P
        sp = self.species[c]
        aff = affected_pairs(a, b, P.nbr, P.rev, P.i, P.j)
        ii, jj = P.i[aff], P.j[aff]
        if a == b:
            pos = self.pos[c].copy()
            pos[a] = newp
            d = (pos[jj] + P.shift[aff]) - pos[ii]
            r_new = np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])
        else:
            r_new = self.r[c, aff].copy()
        spn = sp.copy()
        spn[a], spn[b] = na, nb
        t_new = spn[ii] * P.nsp + spn[jj]
        return aff, r_new, t_new

    def delta_chi(self, c, a, b, na, nb, newp):
        """Unnormalised Delta chi (nsp, nk) of replica c's move, its affected pairs and
        the MS entry-count change."""
        P = self.P
        aff, r_new, t_new = self.move_pairs(c, a, b, na, nb, newp)
        out = self.terms(aff, r_new, t_new) - self.terms(aff, self.r[c, aff], self.ptype[c, aff])
        dC = None
        if P.ms is not None and a != b:
            dC = P.ms.delta_counts(self.species[c], a, b, na, nb)
            nbk = P.ms.nblk
            for s in range(P.nsp):
                out[s] += dC[s * nbk:(s + 1) * nbk].astype(float) @ P.T[s * nbk:(s + 1) * nbk]
        return out, aff, r_new, t_new

    def energy_delta(self, c, aff, t_new):
        """method_a_rmc.py:310-319: sum over the affected pairs of new minus old V/2."""
        nn = self.P.nsp ** 2
        sh = self.P.shell[aff]
        return float((self.etab[sh * nn + t_new] - self.etab[sh * nn + self.ptype[c, aff]]).sum())

    def energy(self, species):
        """method_a_rmc.py:321-326: pair-interaction energy (eV) of a configuration."""
        P = self.P
        sp = np.asarray(species, np.int64)
        return float(self.etab[P.shell * P.nsp ** 2 + sp[P.i] * P.nsp + sp[P.j]].sum())

    # ------------------------------------------------------------------- step
    def step(self):
        self.stepc += 1
        rec = []
        acc_all = np.zeros(self.nrep, bool)
        for c in range(self.nrep):
            a, b, na, nb, sw, newp = propose_move(self.species[c], self.pos[c], self.P.ideal,
                                                  self.seed, self.stepc, c, self.swap_frac,
                                                  self.disp_sigma, self.max_disp)
            dchi, aff, r_new, t_new = self.delta_chi(c, a, b, na, nb, newp)
            trial = self.chi[c] + dchi
            c2 = self.chi2_of_chi(trial)
            x = self.beta[c] * (c2 - self.chi2[c]) * self.nfit
            if self.etab is not None:
                x = x + self.ebeta * self.energy_delta(c, aff, t_new)
            de = self.energy_delta(c, aff, t_new) if self.etab is not None else 0.0
            u = accept_uniform(self.seed, self.stepc, c)
            acc = bool((x <= 0) or (u < math.exp(-min(x, 60.0))))
            if acc:
                self.r[c, aff] = r_new
                self.ptype[c, aff] = t_new
                self.species[c, a], self.species[c, b] = na, nb
                if not sw:
                    self.pos[c, a] = newp
                self.chi[c] = trial
                self.chi2[c] = c2
            acc_all[c] = acc
            rec.append((a, b, sw, int(acc), c2, x, u, newp, de))
        imp = self.chi2 < self.best_chi2
        self.best_chi2 = np.where(imp, self.chi2, self.best_chi2)
        self.best_sp[imp] = self.species[imp]
        self.best_pos[imp] = self.pos[imp]
        self.log.append(rec)
        return acc_all

    def replica_exchange(self):
        """method_a_rmc.py:332-373 with the counter-based uniforms; whole states move."""
        self.iex += 1
        n = self.nrep
        start, us = exchange_uniforms(self.seed, self.iex, n - 1)
        order = np.arange(n)
        nsw = 0
        for i in range(start, n - 1, 2):
            d = (self.beta[i] - self.beta[i + 1]) * (self.chi2[order[i]] - self.chi2[order[i + 1]]) * self.nfit
            if d >= 0 or us[i] < math.exp(max(d, -60.0)):
                order[i], order[i + 1] = order[i + 1], order[i]
                nsw += 1
        if nsw:
            for name in ("species", "pos", "r", "ptype", "chi", "chi2", "best_chi2", "best_sp",
                         "best_pos"):
                setattr(self, name, getattr(self, name)[order])
        return order

    def run(self, nsteps, exchange_every=25):
        for it in range(nsteps):
            self.step()
            if exchange_every and (it + 1) % exchange_every == 0:
                self.replica_exchange()
