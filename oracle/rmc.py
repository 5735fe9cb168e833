"""The plain parallel-tempering RMC swap sampler, step for step (oracle, CPU baseline).

Follows methods/method_a_rmc.py and exafs_gpu/kernels.py:
  * splitmix64 / u01                                   kernels.py:352-360
  * propose_select (swap: up to 8 uniform candidates)  kernels.py:362-423
  * affected directed pairs with dedup (propose_build) kernels.py:425-488
  * Delta chi over the affected pairs, old and new term kernels.py:243-325
  * MS delta through the per-site incidence            chimera/mslibrary.py:466-486
  * full windowed transform of the trial chi           chimera/multift.py:58-63
  * per-edge profiled scale and reduced chi^2          method_a_rmc.py:190-227
  * Metropolis x = beta dchi2 nfit                      method_a_rmc.py:278-285
  * best retention                                      method_a_rmc.py:301-307
  * replica exchange, alternating neighbour pairs       method_a_rmc.py:332-373
  * ladder beta = 1/geomspace(tmin, tmax, nrep)         method_a_rmc.py:118
  * ladder retuning from window exchange rates         method_a_rmc.py:376-425
  * collection rungs nearest collect_beta               method_a_rmc.py:457-476

Decisions that differ from the reference (the reference's own choice cannot be
replayed here; both are recorded):
  * swap-only: when all 8 candidates repeat a's species the reference makes a
    displacement move (kernels.py:403-421); here that proposal is a NULL move
    (state unchanged, counted as rejected).
  * acceptance and exchange uniforms come from counter-based splitmix64 streams
    (`accept_uniform`, `exchange_uniforms`), not CuPy XORWOW / NumPy PCG64.
  * the RNG of a proposal is keyed by the RUNG index (the reference keys it by the
    state index, which equals the rung index because it permutes states).

Parity: splitmix64 is pinned by the published reference outputs of Vigna's
generator (tests); the proposal stream by brute-force restatement; the chi^2 by
the closed form and the profiled-scale minimum (perturbation test); exchange by
detailed balance on a two-rung toy.
"""
from __future__ import annotations

import time

import numpy as np

from .fourier import Transform
from .tables import pair_term

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
ACC_SALT = 0xD1B54A32D192ED03
EXC_SALT = 0x94D049BB133111EB


def splitmix64(x):
    """kernels.py:352-357 (Vigna's splitmix64 finaliser on x + golden)."""
    x = (x + GOLDEN) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


def u01(h):
    """kernels.py:358-360."""
    return float(h >> 11) * (1.0 / 9007199254740992.0)


def propose(species, seed, stepc, c, swap_frac=1.0):
    """propose_select (kernels.py:362-398) for replica key c.

    Returns (a, b, new_a, new_b, is_swap); a null move has b == a, is_swap 0."""
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
    real = sw and b != a
    if not real:
        return a, a, sa, sa, 0
    return a, b, sb, sa, 1


def accept_uniform(seed, stepc, c):
    h = splitmix64((seed ^ ((stepc * GOLDEN) & M64) ^ ((c << 32) & M64) ^ ACC_SALT) & M64)
    return u01(splitmix64(h))


def exchange_uniforms(seed, iex, npairs):
    """(start parity, uniforms per adjacent pair) of exchange number iex."""
    hx = splitmix64((seed ^ ((iex * GOLDEN) & M64) ^ EXC_SALT) & M64)
    start = int(hx >> 63)
    us = [u01(splitmix64(hx ^ (((i + 1) << 32) & M64))) for i in range(npairs)]
    return start, us


def adapt_ladder(beta, window_exch):
    """Temperature ladder retuned from the exchange rates of a window
    (method_a_rmc.py:376-398, ladder part; the displacement-step part has no swap-only
    effect).  Working in log beta with both endpoints fixed: a neighbour gap whose
    exchange rate r (clipped to [0.02, 0.98]) is above the mean rate r_bar widens, one
    below it narrows, each by sqrt(r / r_bar); then every gap is rescaled by the one
    factor that restores the original span."""
    lb = np.log(np.asarray(beta, np.float64))
    r = np.clip(np.asarray(window_exch, np.float64), 0.02, 0.98)
    gaps = np.diff(lb) * (r / max(r.mean(), 1e-6)) ** 0.5
    gaps = gaps * ((lb[-1] - lb[0]) / gaps.sum())
    return np.exp(np.concatenate([[lb[0]], lb[0] + np.cumsum(gaps)]))


def affected_pairs(a, b, nbr, rev, i_idx, j_idx):
    """propose_build's affected directed pairs with dedup (kernels.py:448-468)."""
    out = []
    for blk, site in ((0, a), (1, a), (2, b), (3, b)):
        for p in nbr[site]:
            q = p if blk in (0, 2) else rev[p]
            if blk >= 2:
                if b == a:
                    continue
                other = j_idx[q] if blk == 2 else i_idx[q]
                if other == a:
                    continue
            out.append(int(q))
    return np.array(out, np.int64)


class Problem:
    """Everything the sampler needs about one fit (topology, tables, target)."""

    def __init__(self, k, f, delta, lam, s02, ideal, i_idx, j_idx, shift, shell, sig2,
                 transforms, chan_edge, target, sigma, R=None, mask=None,
                 ms=None, T=None, dr=None, c3=None, s2types=None, chi_extra=None):
        """dr / c3 / s2types: the cumulant tables of exafs_gpu/forward.py:104-145
        (r_eff = r + dr[shell, a, b], sigma^2 + s2types[shell, a, b], phase
        - 4/3 k^3 c3[shell]); chi_extra (nsp, nk): the fixed term added to the
        normalised chi before transforming (chimera/multift.py:58-61)."""
        self.k, self.f, self.delta, self.lam, self.s02 = k, f, delta, lam, s02
        self.nsp = f.shape[0]
        nshell = int(np.asarray(shell).max()) + 1
        self.dr = np.zeros((nshell, self.nsp, self.nsp)) if dr is None else np.asarray(dr, float)
        self.c3 = np.zeros(nshell) if c3 is None else np.asarray(c3, float)
        self.s2types = (np.zeros((nshell, self.nsp, self.nsp)) if s2types is None
                        else np.asarray(s2types, float))
        self.chi_extra = None if chi_extra is None else np.asarray(chi_extra, float)
        self.ideal = ideal
        self.i, self.j, self.shift, self.shell = i_idx, j_idx, shift, shell
        self.sig2 = np.broadcast_to(np.asarray(sig2, float), i_idx.shape).copy()
        self.natoms = ideal.shape[0]
        counts = np.bincount(i_idx, minlength=self.natoms)
        assert counts.min() == counts.max(), "expected uniform coordination"
        self.ncoord = int(counts[0])
        self.nbr = np.arange(i_idx.size).reshape(self.natoms, self.ncoord)
        from .lattice import reverse_index
        self.rev = reverse_index(i_idx, j_idx, self.natoms)
        self.r = np.linalg.norm((ideal[j_idx] + shift) - ideal[i_idx], axis=1)
        self.transforms = transforms
        self.chan_edge = np.asarray(chan_edge)
        self.nchan = len(transforms)
        self.nkw = self.nchan // self.nsp
        self.R = R
        if R is None:
            self.mask = np.asarray(mask, bool)
            self.tm = target[:, self.mask]
            self.sm = sigma[:, self.mask]
        else:
            self.mask = None
            self.tm = target
            self.sm = sigma
        self.nfit = int(self.tm.shape[1]) * self.nchan * 2
        self.ms, self.T = ms, T
        # for the CPU baseline: the DFT restricted to the window's nonzero k points
        if R is not None:
            self._E = []
            for t in transforms:
                nzk = np.nonzero(t.weight)[0]
                self._E.append((nzk, t.weight[nzk][:, None]
                                * np.exp(-2j * np.outer(t.k_uniform[nzk], R)) * t.norm))

    # ---------------------------------------------------------------- pieces
    def term(self, kk, r, t, s2):
        nsp = self.nsp
        sa, sb = t // nsp, t % nsp
        return pair_term(kk, r, self.f[sb], self.delta[sa, sb], self.lam, s2, self.s02)

    def _terms(self, pairs, t):
        """Sum of the pair terms of `pairs` with types `t`, by absorber (nsp, nk)."""
        nsp = self.nsp
        k = self.k
        nz = k > 0
        kz = k[nz][None, :]
        sa, sb = t // nsp, t % nsp
        sh = self.shell[pairs]
        r = (self.r[pairs] + self.dr[sh, sa, sb])[:, None]
        s2 = (self.sig2[pairs] + self.s2types[sh, sa, sb])[:, None]
        amp = (self.s02 * self.f[sb][:, nz] / (kz * r * r)
               * np.exp(-2.0 * r / self.lam[nz][None, :] - 2.0 * kz * kz * s2))
        term = amp * np.sin(2.0 * kz * r + self.delta[sa, sb][:, nz]
                            - (4.0 / 3.0) * kz ** 3 * self.c3[sh][:, None])
        out = np.zeros((nsp, k.size))
        for s in range(nsp):
            out[s, nz] = term[sa == s].sum(0)
        return out

    def chi_full(self, species):
        from .forward import chi_pairs
        out = chi_pairs(self.k, self.f, self.delta, self.lam, self.s02, self.ideal, species,
                        self.i, self.j, self.shift, self.sig2, shell=self.shell, dr=self.dr,
                        c3=self.c3, s2types=self.s2types, normalise=False)
        if self.ms is not None:
            C = self.ms.counts(species)
            nb = self.ms.nblk
            for a in range(self.nsp):
                out[a] += C[a * nb:(a + 1) * nb].astype(float) @ self.T[a * nb:(a + 1) * nb]
        return out

    def to_R(self, chi):
        """(nsp, nk) normalised chi -> (nchan, nwin) complex on the fit window."""
        if self.chi_extra is not None:
            chi = chi + self.chi_extra
        if self.R is None:
            out = [t.to_R(chi[e])[self.mask] for t, e in zip(self.transforms, self.chan_edge)]
        else:
            out = []
            for (nzk, E), t, e in zip(self._E, self.transforms, self.chan_edge):
                cu = t.resample(chi[e])
                out.append(cu[nzk] @ E)
        return np.array(out)

    def profiled_scale(self, m):
        """method_a_rmc.py:190-209 (complex target)."""
        num = (m.real * self.tm.real + m.imag * self.tm.imag).sum(-1)
        den = (m.real ** 2 + m.imag ** 2).sum(-1)
        num = num.reshape(self.nsp, self.nkw).sum(-1)
        den = den.reshape(self.nsp, self.nkw).sum(-1)
        return np.repeat(num / np.maximum(den, 1e-30), self.nkw)

    def chi2(self, chi_un, cnt, profile=True):
        """method_a_rmc.py:211-227; also returns the window transform."""
        chi = chi_un / np.maximum(cnt, 1.0)[:, None]
        m = self.to_R(chi)
        if profile:
            m = m * self.profiled_scale(m)[:, None]
        d = (m - self.tm) / self.sm
        return float((d.real ** 2 + d.imag ** 2).sum() / self.nfit)

    def delta_chi(self, species, a, b, na, nb):
        """Unnormalised Delta chi (nsp, nk) of the swap (kernels.py:243-325 plus
        mslibrary.py:466-486)."""
        out = np.zeros((self.nsp, self.k.size))
        if a == b:
            return out, None
        aff = affected_pairs(a, b, self.nbr, self.rev, self.i, self.j)
        nsp = self.nsp
        sp_new = species.copy()
        sp_new[a], sp_new[b] = na, nb
        ii, jj = self.i[aff], self.j[aff]
        t_old = species[ii].astype(np.int64) * nsp + species[jj]
        t_new = sp_new[ii].astype(np.int64) * nsp + sp_new[jj]
        out += self._terms(aff, t_new)
        out -= self._terms(aff, t_old)
        dC = None
        if self.ms is not None:
            dC = self.ms.delta_counts(species, a, b, na, nb)
            nbk = self.ms.nblk
            for s in range(nsp):
                out[s] += dC[s * nbk:(s + 1) * nbk].astype(float) @ self.T[s * nbk:(s + 1) * nbk]
        return out, dC


class PlainSampler:
    """The reference sampler loop on the CPU, replica by replica (the CPU baseline
    and the checker of the device sampler)."""

    def __init__(self, prob: Problem, nrep=8, tmin=0.02, tmax=5.0, seed=0,
                 collect_beta=0.5, n_collect=4, profile_scale=True):
        self.P = prob
        self.nrep = nrep
        self.beta = 1.0 / np.geomspace(tmin, tmax, nrep)
        self.seed = int(seed)
        self.profile = profile_scale
        self.collect_beta, self.n_collect = collect_beta, n_collect
        self.stepc = 0
        self.iex = 0
        self.log = []

    def init_state(self, species0):
        P = self.P
        self.species = np.repeat(np.asarray(species0, np.int64)[None], self.nrep, 0)
        chi0 = P.chi_full(self.species[0])
        self.chi = np.repeat(chi0[None], self.nrep, 0)
        self.cnt = np.bincount(self.species[0], minlength=P.nsp).astype(float)
        c2 = P.chi2(chi0, self.cnt, self.profile)
        self.chi2 = np.full(self.nrep, c2)
        self.best_chi2 = self.chi2.copy()
        self.best_sp = self.species.copy()

    def step(self):
        P = self.P
        self.stepc += 1
        acc_all = np.zeros(self.nrep, bool)
        rec = []
        for c in range(self.nrep):
            sp = self.species[c]
            a, b, na, nb, sw = propose(sp, self.seed, self.stepc, c)
            u = accept_uniform(self.seed, self.stepc, c)
            if not sw:
                rec.append((a, b, 0, self.chi2[c], 0.0, u))
                continue
            dchi, _ = P.delta_chi(sp, a, b, na, nb)
            trial = self.chi[c] + dchi
            c2 = P.chi2(trial, self.cnt, self.profile)
            x = self.beta[c] * (c2 - self.chi2[c]) * P.nfit
            acc = (x <= 0) or (u < np.exp(-min(x, 60.0)))
            if acc:
                sp[a], sp[b] = na, nb
                self.chi[c] = trial
                self.chi2[c] = c2
            acc_all[c] = acc
            rec.append((a, b, int(acc), c2, x, u))
        imp = self.chi2 < self.best_chi2
        self.best_chi2 = np.where(imp, self.chi2, self.best_chi2)
        self.best_sp[imp] = self.species[imp]
        self.log.append(rec)
        return acc_all

    def replica_exchange(self, return_rates=False):
        """method_a_rmc.py:332-373 with the counter-based uniforms.  With return_rates,
        also this sweep's 0/1 outcome per neighbour pair, the pairs not tried holding
        the mean over the tried ones (method_a_rmc.py:366-372)."""
        self.iex += 1
        n = self.nrep
        start, us = exchange_uniforms(self.seed, self.iex, n - 1)
        order = np.arange(n)
        nsw = 0
        outcome = np.full(n - 1, np.nan)
        for i in range(start, n - 1, 2):
            d = (self.beta[i] - self.beta[i + 1]) * (self.chi2[order[i]] - self.chi2[order[i + 1]]) * self.P.nfit
            outcome[i] = 0.0
            if d >= 0 or us[i] < np.exp(max(d, -60.0)):
                order[i], order[i + 1] = order[i + 1], order[i]
                nsw += 1
                outcome[i] = 1.0
        if nsw:
            self.species = self.species[order]
            self.chi = self.chi[order]
            self.chi2 = self.chi2[order]
            self.best_chi2 = self.best_chi2[order]
            self.best_sp = self.best_sp[order]
        if return_rates:
            tried = ~np.isnan(outcome)
            fill = outcome[tried].mean() if tried.any() else 0.3
            return order, np.where(tried, outcome, fill)
        return order

    def run(self, nsteps, exchange_every=25, adapt_every=0):
        """method_a_rmc.py:401-425: steps, exchange every `exchange_every`, and with
        `adapt_every` the ladder retuned from each window's exchange rates."""
        t0 = time.perf_counter()
        window = []
        for it in range(nsteps):
            self.step()
            if exchange_every and (it + 1) % exchange_every == 0:
                _, rates = self.replica_exchange(return_rates=True)
                if adapt_every:
                    window.append(rates)
            if adapt_every and (it + 1) % adapt_every == 0:
                if window and self.nrep >= 4:
                    self.beta = adapt_ladder(self.beta, np.mean(window, axis=0))
                window = []
        return time.perf_counter() - t0

    def collect_rungs(self):
        k = min(self.n_collect, self.nrep)
        return np.argsort(np.abs(np.log(self.beta) - np.log(self.collect_beta)))[:k]
