"""RMCSampler: parallel-tempering reverse Monte Carlo swap sampler, device-resident.

Drop-in for methods/method_a_rmc.py:63-496 for the reference's swap-only use
(chimera/fitrun.py:307-315 passes swap_frac=1.0): same constructor, init_state / step /
run / replica_exchange / best / scales / collect_rungs / posterior_wc, and the result
attributes samples, history, walltime, exchange_rate, best_verified, chi2, species,
beta, chi.  Arrays come back as NumPy in rung order, as the reference holds them.

What runs where: every step (proposal, gather of the pairs and paths through the two
swapped sites, trial chi(R), profiled chi^2, Metropolis, commit) and every replica
exchange run on the GPU (libchimera_hip: K1-K4, replayed from a captured hipGraph);
the host only launches and reads back at log / collection points.

Count space: on the ideal lattice every pair term is B[a,b,shell](k) and every
multiple-scattering path term is T_entry(k), so a configuration's chi(R) is its
integer counts times the transformed basis.  The basis is built on the device (term
rows, then the DFT-matrix transform of each edge) when init_state fixes the
composition (the per-edge normalisation by absorber count).

Decisions (recorded, the reference's own choice cannot be replayed without CuPy):
  * when all 8 candidates repeat a's species the proposal is a null move (the
    reference falls back to a displacement, kernels.py:403-421);
  * acceptance and exchange uniforms are counter-based splitmix64 streams keyed by
    (seed, step, rung) and (seed, exchange number);
  * exchange swaps temperatures between slots, not configurations.
"""
from __future__ import annotations

import inspect
import sys
import time

import numpy as np

from . import _lib
from ._lib import check, f64, i32
from .lattice import alpha_from_counts


def state_sections(blob):
    """{name: (byte offset, byte count)} of the sections of a save_state blob (the
    header written by chm_sampler_export: magic, version, eight int32 sizes, step count,
    seed, then 17 offsets and 17 byte counts)."""
    b = np.frombuffer(bytes(blob) if not isinstance(blob, np.ndarray) else blob.tobytes(), np.uint8)
    magic, version = b[:8].view(np.uint32)
    if magic != 0x53534D43 or version != 1:
        raise ValueError("not a sampler state blob")
    nsect = int(b[8:40].view(np.int32)[7])
    tab = b[56:56 + 16 * nsect].view(np.int64)
    return {name: (int(tab[q]), int(tab[nsect + q]))
            for q, name in enumerate(RMCSampler._STATE_SECTIONS[:nsect])}


def _beta_ladder(tmin, tmax, nrep):
    return 1.0 / np.geomspace(tmin, tmax, nrep)


def _needs_pair_space(target, swap_frac, energy_prior):
    """Whether a sampler with these arguments needs the pair-space engine: displacement
    moves, the energy prior or an |chi(R)| (real) target."""
    return swap_frac != 1.0 or energy_prior is not None or not np.iscomplexobj(target)


class RMCSampler:
    engine = "count"

    def __new__(cls, *args, **kwargs):
        # The reference's full move set (displacements, energy prior, |chi(R)| targets) runs
        # on the pair-space engine; the swap-only chain on a complex target on count space.
        if cls is RMCSampler:
            b = _SIG.bind(None, *args, **kwargs)
            b.apply_defaults()
            a = b.arguments
            if _needs_pair_space(a["target"], a["swap_frac"], a["energy_prior"]):
                from .pair_rmc import PairRMCSampler
                cls = PairRMCSampler
        return object.__new__(cls)

    def __init__(self, fm, ft, target, mask, sigma_R, nrep=48, tmin=0.02, tmax=5.0, seed=0,
                 swap_frac=0.75, disp_sigma=0.015, max_disp=0.30, nsp=None,
                 collect_after=None, collect_every=25, adapt_every=0, target_accept=0.35,
                 profile_scale=True, collect_beta=0.5, n_collect=4, energy_prior=None,
                 ms_library=None, precision="f64", reanchor_every=512, accumulator="f64",
                 paired=True):
        self._ctor = {k: v for k, v in locals().items() if k not in ("self", "__class__")}
        """Arguments as methods/method_a_rmc.py:64-69, plus `precision`: "f64" (the trial
        chi(R) recomputed from the integer counts in FP64 every step), "bf16x3" (the
        change of chi(R) in split-BF16 MFMA with FP32 accumulation, added to a running
        chi(R) re-anchored in FP64 every `reanchor_every` accepted moves of a replica, 0 =
        never) or "f32" (the change in FP32 FMA, same running chi(R)); `accumulator`:
        "f64" (default) or "f32" (the running chi(R) rounded to FP32 after every update;
        for precision studies).  chi^2 and the acceptance test are FP64 in every mode.
        `paired` (bf16x3 with the uniform proposal): two steps per transform and
        per-replica launch, by pre-fetching the trial of the second step for both outcomes
        of the first; the chain is the same move for move."""
        if precision not in ("f64", "bf16x3", "f32"):
            raise ValueError("precision must be 'f64', 'bf16x3' or 'f32'")
        if accumulator not in ("f64", "f32") or (precision == "f64" and accumulator == "f32"):
            raise ValueError("accumulator must be 'f64' or 'f32' (an incremental precision only)")
        self.precision = precision
        self.accumulator = accumulator
        self.reanchor_every = int(reanchor_every)
        self.paired = bool(paired)
        if swap_frac != 1.0:
            raise ValueError("the device sampler makes swap moves only: pass swap_frac=1.0 "
                             "(the reference's fit driver does, chimera/fitrun.py:218)")
        if energy_prior is not None:
            raise ValueError("energy_prior is not supported by the device sampler")
        self.adapt_every = int(adapt_every)
        self.target_accept = float(target_accept)
        self.disp_sigma = float(disp_sigma)
        self.fm, self.ft = fm, ft
        self.nsp = int(nsp) if nsp is not None else int(fm.nsp)
        self.nrep = int(nrep)
        self.collect_after = collect_after
        self.collect_every = int(collect_every)
        self.collect_beta = None if collect_beta is None else float(collect_beta)
        self.n_collect = int(n_collect)
        self.profile_scale = bool(profile_scale)
        self.samples = []
        self.seed = int(seed)
        self.swap_frac = swap_frac
        self.natoms = fm.natoms
        self.npairs = fm.npairs
        self.nk = fm.nk
        if ms_library is not None and not hasattr(ms_library, "stencil"):
            from .mspaths import MSLibrary
            ms_library = MSLibrary.from_protocol(ms_library, fm.cell)
        self.ms = ms_library

        i_np, j_np = np.asarray(fm.i_idx), np.asarray(fm.j_idx)
        counts = np.bincount(i_np, minlength=self.natoms)
        assert counts.min() == counts.max(), "expected uniform coordination"
        assert np.all(np.diff(i_np) >= 0), "pairs must be i-sorted"
        self.ncoord = int(counts[0])
        key = i_np.astype(np.int64) * self.natoms + j_np
        rkey = j_np.astype(np.int64) * self.natoms + i_np
        assert np.array_equal(np.sort(key), np.sort(rkey)), "neighbour list is not symmetric"
        self._beta_rung = _beta_ladder(tmin, tmax, self.nrep)

        self.complex = bool(np.iscomplexobj(target))
        if not self.complex:
            raise ValueError("the device sampler needs a complex chi(R) target (chi(k) data)")
        target = np.asarray(target)
        self.mask = np.asarray(mask, bool)
        self.nchan = int(target.shape[0])
        self.nkw = self.nchan // self.nsp
        assert self.nkw * self.nsp == self.nchan, \
            f"{self.nchan} channels do not divide into {self.nsp} edges"
        chan_edge = getattr(ft, "chan_edge", None)
        if chan_edge is not None:
            assert np.array_equal(np.asarray(chan_edge), np.repeat(np.arange(self.nsp), self.nkw)), \
                "channels must be ordered edge-major"
        self.nwin = int(self.mask.sum())
        self.nfit = self.nwin * self.nchan * 2
        self._tm = target[:, self.mask]
        self._sm = np.asarray(sigma_R)[:, self.mask]
        self._exchange_rate = None
        self._rate_pending = False
        self._S = None
        self._comp = None
        self._log_cap = 0
        self._build_hooks = []   # weak methods re-applied to every new device sampler
        self._best_verified = None
        self._verify_pending = False

    # ------------------------------------------------------------------ basis
    def _lattice_geometry(self):
        cell = self.fm.cell
        a0 = float(cell.a0)
        unit = a0 / 2.0
        X = np.asarray(cell.ideal, float) / unit
        Xi = np.rint(X).astype(np.int64)
        assert np.allclose(X, Xi, atol=1e-6), "sites are not lattice points"
        ncell = np.array(cell.ncell, np.int64)
        box = 2 * ncell
        Xi %= box
        par = (Xi[:, 0] & 1) | ((Xi[:, 1] & 1) << 1) | ((Xi[:, 2] & 1) << 2)
        nbasis = self.natoms // int(np.prod(ncell))
        bpar = -np.ones(8, np.int32)
        cellidx = ((Xi[:, 0] >> 1) * ncell[1] + (Xi[:, 1] >> 1)) * ncell[2] + (Xi[:, 2] >> 1)
        b = np.arange(self.natoms) - cellidx * nbasis
        for p_, b_ in zip(par, b):
            if bpar[p_] < 0:
                bpar[p_] = b_
            assert bpar[p_] == b_, "site order is not cell-major with a fixed basis order"
        xyz = (Xi[:, 0] | (Xi[:, 1] << 10) | (Xi[:, 2] << 20)).astype(np.int32)
        return xyz, box, ncell, nbasis, bpar

    def _shells(self):
        fm = self.fm
        d = np.linalg.norm((fm.cell.ideal[fm.j_idx] + fm.d_shift) - fm.cell.ideal[fm.i_idx], axis=1)
        sh = np.asarray(fm.shell)
        radii, s2 = [], []
        for s in range(fm.nshell):
            m = sh == s
            r_s = d[m]
            if r_s.size == 0:
                radii.append(1.0)
                s2.append(0.0)
                continue
            if np.ptp(r_s) > 1e-9 * r_s.mean():
                raise ValueError("pairs of one shell differ in length: the count-space sampler "
                                 "needs the ideal lattice with one radius per shell")
            sg = np.asarray(fm.d_sig2)[m]
            if np.ptp(sg) > 1e-12 * max(abs(sg.mean()), 1e-30):
                raise ValueError("sigma^2 differs within a shell")
            radii.append(float(r_s.mean()))
            s2.append(float(sg.mean()))
        return np.array(radii), np.array(s2)

    def _build(self, species0):
        fm = self.fm
        nsp, nk = self.nsp, self.nk
        natoms_of = np.bincount(species0, minlength=nsp)
        self._comp = natoms_of
        radii, s2 = self._shells()
        B = fm.ss_basis(radii, s2)                         # (nsp, nsp, nshell, nk)
        nshell = fm.nshell
        if self.ms is not None:
            T = self.ms.table(fm)
            nblk = self.ms.nblk
        else:
            T, nblk = None, 0
        extra = getattr(self.ft, "chi_extra", None)
        has_const = extra is not None
        nrow_e = nsp * nshell + nblk + (1 if has_const else 0)
        dev = self.ft.device() if hasattr(self.ft, "device") else None
        if dev is None:
            raise ValueError("ft must be a chimera_amd MultiEdgeFT")
        ncol_e = self.nkw * self.nwin * 2
        basis_R = np.zeros((nsp, nrow_e, ncol_e))
        basis_k = np.zeros((nsp, nrow_e, nk))
        for e in range(nsp):
            rows = [B[e].reshape(nsp * nshell, nk)]
            if T is not None:
                rows.append(T[e * nblk:(e + 1) * nblk])
            rk = np.concatenate(rows, 0)
            basis_k[e, :rk.shape[0]] = rk
            scaled = rk / max(natoms_of[e], 1)
            if has_const:
                scaled = np.concatenate([scaled, np.asarray(extra)[e][None]], 0)
            outR = dev.apply_edge(e, scaled)                # (nrow_e, nkw, nR, 2)
            basis_R[e] = outR[:, :, self.mask, :].reshape(nrow_e, ncol_e)
        self._basis_k = basis_k
        self._basis_R = basis_R
        # geometry
        xyz, box, ncell, nbasis, bpar = self._lattice_geometry()
        g = _lib.Geom()
        g.natoms, g.nsp, g.ncoord, g.nshell = self.natoms, nsp, self.ncoord, nshell
        g.nblk = nblk
        if self.ms is not None:
            stencil = np.ascontiguousarray(self.ms.stencil, np.int8)
            g.nstencil, g.nabsrec = self.ms.nstencil, self.ms.nabsrec
        else:
            stencil = None
            g.nstencil = g.nabsrec = 0
        g.nkw, g.nwin = self.nkw, self.nwin
        for d in range(3):
            g.box[d] = int(box[d])
            g.ncell[d] = int(ncell[d])
        g.nbasis = int(nbasis)
        for q in range(8):
            g.basis_of_parity[q] = int(bpar[q])
        g.has_const = 1 if has_const else 0
        tgt = np.stack([self._tm.real, self._tm.imag], -1).reshape(-1)
        nbr_j = i32(np.asarray(fm.j_idx).reshape(self.natoms, self.ncoord))
        nbr_sh = i32(np.asarray(fm.shell).reshape(self.natoms, self.ncoord))
        h = _lib.vp()
        check(_lib.load().chm_sampler_create(
            fm.ctx.h, _lib.C.byref(g), self.nrep, nk, xyz, nbr_j, nbr_sh,
            None if stencil is None else stencil.ctypes.data, f64(basis_R), f64(basis_k),
            f64(tgt), f64(self._sm.reshape(-1)), int(self.profile_scale), _lib.C.byref(h)))
        self._S = h
        self._stencil_keep = stencil
        self.nrow_e = nrow_e
        self.nshell = nshell

    # ------------------------------------------------------------------ state
    def init_state(self, species0, pos0):
        self._settle()
        species0 = i32(species0)
        pos0 = np.asarray(pos0, float)
        if not np.allclose(pos0, self.fm.cell.ideal, atol=1e-9):
            # an off-lattice start: the chain continues on the pair-space engine
            self._to_pair_space()
            return self.init_state(species0, pos0)
        comp = np.bincount(species0, minlength=self.nsp)
        if self._S is None or not np.array_equal(comp, self._comp):
            self._destroy()
            self._build(species0)
            if self._log_cap:
                self.set_recording(self._log_cap, self._series_cap)
            for ref in list(self._build_hooks):
                hook = ref()
                if hook is None:
                    self._build_hooks.remove(ref)
                else:
                    hook(self)
        mode = {"f64": 0, "bf16x3": 1, "f32": 2}[self.precision]
        if getattr(self, "accumulator", "f64") == "f32":
            mode |= 16
        check(_lib.load().chm_sampler_set_precision(self._S, mode, self.reanchor_every))
        check(_lib.load().chm_sampler_set_pairs(self._S, int(getattr(self, "paired", True))))
        check(_lib.load().chm_sampler_init(self._S, species0, f64(self._beta_rung),
                                           int(self.seed) & ((1 << 64) - 1)))
        self.samples = []
        self._t_run = 0.0

    def _to_pair_space(self):
        """Become a PairRMCSampler with the same constructor arguments (collected samples
        are kept)."""
        from .pair_rmc import PairRMCSampler
        samples = self.samples
        self._destroy()
        self.__class__ = PairRMCSampler
        PairRMCSampler.__init__(self, **self._ctor)
        self.samples = samples

    def set_recording(self, log_steps=0, series_steps=0, series_rung=None):
        """Keep the move log (a, b, accepted, rung per step and replica) and the
        pair-count series of one rung (default: nearest collect_beta) on the device."""
        self._log_cap, self._series_cap = int(log_steps), int(series_steps)
        if series_rung is None:
            series_rung = int(self.collect_rungs()[0]) if series_steps else -1
        self._series_rung = series_rung
        if self._S is not None:
            check(_lib.load().chm_sampler_set_recording(self._S, self._log_cap, self._series_cap,
                                                        int(series_rung)))

    # ------------------------------------------------------------- checkpoint
    _STATE_SECTIONS = ("species", "counts", "chi2_rows", "running_chiR", "best_chi2",
                       "best_species", "rung_of", "slot_of", "beta", "beta_slot", "step_counter",
                       "exchange_counter", "accepted_per_rung", "accepted_since_anchor",
                       "reanchors", "exchange_attempts", "exchange_accepts")

    def save_state(self):
        """The whole chain state between runs as bytes (species, counts, chi^2, running
        chi(R), best state, rung/slot maps, ladder, counters, step count, seed).
        load_state on a sampler of the same problem, precision and composition (made
        and initialised the same way) continues the chain bit for bit."""
        self._settle()
        lib = _lib.load()
        n = _lib.C.c_int64()
        check(lib.chm_sampler_state_size(self._S, _lib.C.byref(n)))
        buf = np.zeros(n.value, np.uint8)
        check(lib.chm_sampler_export(self._S, buf.ctypes.data, buf.nbytes))
        return buf.tobytes()

    def load_state(self, blob):
        """Continue from a blob written by save_state (see there).  Collected samples
        are kept; the ladder is taken from the blob."""
        self._settle()
        if self._S is None:
            raise RuntimeError("init_state first (the blob replaces its state)")
        buf = np.frombuffer(bytes(blob), np.uint8).copy()
        check(_lib.load().chm_sampler_import(self._S, buf.ctypes.data, buf.nbytes))
        sec = state_sections(buf)
        off, nb = sec["beta"]
        self._beta_rung = buf[off:off + nb].view(np.float64).copy()

    def _read(self, what, shape, dtype):
        out = np.empty(shape, dtype)
        check(_lib.load().chm_sampler_read(self._S, what, out.ctypes.data, out.nbytes))
        return out

    # Device calls of the run loop, exchange bookkeeping and history marks (the pair-space
    # engine, chimera_amd.pair_rmc, supplies its own).
    def _lib_run(self, n, E):
        check(_lib.load().chm_sampler_run(self._S, int(n), int(E)))

    def _lib_set_beta(self, ladder):
        check(_lib.load().chm_sampler_set_beta(self._S, f64(ladder)))

    def _lib_exchange(self):
        check(_lib.load().chm_sampler_exchange(self._S))

    def _lib_reserve_marks(self, n):
        check(_lib.load().chm_sampler_reserve_marks(self._S, int(n)))

    def _lib_mark(self, i):
        check(_lib.load().chm_sampler_mark(self._S, int(i)))

    def _lib_read_marks(self, n):
        mk = np.empty((n, 3))
        check(_lib.load().chm_sampler_read_marks(self._S, int(n), mk))
        return mk

    def _steps_done(self):
        return int(self._read(10, (1,), np.int64)[0])

    def _exch_raw(self):
        """(2, nrep - 1) int64 exchange attempts and accepts per neighbour pair."""
        return self._read(8, (2, max(self.nrep - 1, 1)), np.int64)

    # ------------------------------------------------------------- properties
    @property
    def species(self):
        return self._read(0, (self.nrep, self.natoms), np.int32)

    @property
    def chi2(self):
        return self._read(1, (self.nrep,), np.float64)

    @property
    def best_chi2(self):
        return self._read(3, (self.nrep,), np.float64)

    @property
    def counts(self):
        return self._read(2, (self.nrep, self.nsp * self.nrow_e), np.int32)

    @property
    def chi(self):
        """Unnormalised chi(k) (nrep, nsp, nk), rung order."""
        return self._read(11, (self.nrep, self.nsp, self.nk), np.float64)

    @property
    def beta(self):
        return self._beta_rung.copy()

    @property
    def count(self):
        return np.repeat(self._comp[None].astype(float), self.nrep, 0)

    def move_log(self, nsteps):
        return self._read(6, (self._log_cap, self.nrep, 4), np.int32)[:nsteps]

    def pair_series(self, nsteps):
        nss = self.nsp * self.nsp * self.nshell
        return self._read(7, (self._series_cap, nss), np.int32)[:nsteps].reshape(
            -1, self.nsp, self.nsp, self.nshell)

    def accepted_per_rung(self):
        return self._read(9, (self.nrep,), np.int64)

    def reanchors(self):
        """FP64 re-anchors of the running chi(R) per slot (incremental precision)."""
        return self._read(14, (self.nrep,), np.int64)

    def running_chiR(self):
        """(nrep, nsp*nkw*nwin*2) running chi(R) accumulator, slot order."""
        return self._read(13, (self.nrep, self.nsp * self.nkw * self.nwin * 2), np.float64)

    def exact_chiR(self):
        """(nrep, nsp*nkw*nwin*2) chi(R) of every slot recomputed on the host in FP64 from
        its integer counts (the reference value of the running accumulator)."""
        C = self.counts_slot_order().astype(np.float64).reshape(self.nrep, self.nsp, self.nrow_e)
        return np.concatenate([C[:, e] @ self._basis_R[e] for e in range(self.nsp)], axis=1)

    def counts_slot_order(self):
        rung_of = self._read(5, (self.nrep,), np.int32)
        return self.counts[rung_of]

    # ------------------------------------------------------------------ moves
    def step(self):
        """One swap step for every replica; returns the accepted flags (rung order)."""
        self._settle()
        before = self.accepted_per_rung()
        self._lib_run(1, 0)
        return (self.accepted_per_rung() - before) > 0

    def replica_exchange(self, return_rates=False):
        """One exchange sweep on the device (method_a_rmc.py:332-373).  Returns the
        number of swaps (and per-pair 0/1 rates of this sweep, NaN where not tried)."""
        self._settle()
        ea0 = self._exch_raw()
        self._lib_exchange()
        ea1 = self._exch_raw()
        tried = (ea1[0] - ea0[0]) > 0
        ok = (ea1[1] - ea0[1]) > 0
        nsw = int(ok.sum())
        if return_rates:
            rates = np.where(tried, ok.astype(float), np.nan)
            filled = np.where(np.isnan(rates), np.nanmean(rates) if tried.any() else 0.3, rates)
            return nsw, filled
        return nsw

    def adapt(self, window_acc, window_exch):
        """Retune the temperature ladder (method_a_rmc.py:376-398).  In log beta, with the
        hottest and coldest rungs fixed, each neighbour gap is scaled by the square root
        of its exchange rate over the window (clipped to [0.02, 0.98]) relative to the
        mean rate, then all gaps by the one factor that restores the span: pairs that
        rarely exchange move closer, pairs that exchange freely move apart.  The new
        ladder goes to the device; each slot takes the beta of the rung it holds.
        The displacement step is retuned as the reference does, though swap-only chains
        make no displacement moves."""
        if window_acc is not None and window_acc > 0:
            f = min(max((window_acc / self.target_accept) ** 0.35, 0.7), 1.4)
            self.disp_sigma = min(max(self.disp_sigma * f, 1e-4), 0.15)
        if window_exch is None or self.nrep < 4:
            return
        logb = np.log(self._beta_rung)
        rate = np.clip(np.asarray(window_exch, np.float64), 0.02, 0.98)
        step = np.diff(logb) * (rate / max(rate.mean(), 1e-6)) ** 0.5
        step = step * ((logb[-1] - logb[0]) / step.sum())
        ladder = np.exp(np.concatenate([[logb[0]], logb[0] + np.cumsum(step)]))
        self._lib_set_beta(ladder)
        self._beta_rung = ladder

    def _exchange_outcomes(self, since):
        """0/1 outcome per neighbour pair of the exchange sweep made since the attempt /
        accept counters read `since`, untried pairs filled with the mean of the tried
        ones (method_a_rmc.py:366-372); returns (rates, counters now)."""
        ea = self._exch_raw()
        tried = (ea[0] - since[0]) > 0
        ok = (ea[1] - since[1]) > 0
        fill = ok[tried].mean() if tried.any() else 0.3
        return np.where(tried, ok.astype(np.float64), fill), ea

    def run(self, nsteps, exchange_every=25, log_every=500, callback=None):
        """Run the chain on the device (method_a_rmc.py:401-455).  With adapt_every the
        run stops at every exchange to read that sweep's per-pair outcomes and retunes
        the ladder every adapt_every steps, as the reference's loop does.

        The history tuples (step, best chi^2, acceptance, seconds) are those of the
        reference's log points (step 1 and every log_every steps).  Without a callback
        they are snapshots taken on the device stream and read once at the end (the
        seconds from the device clock), so logging costs no host round trip; with a
        callback the run stops at every log point to call it.  best_verified (the
        retained best's chi^2 recomputed from scratch) is evaluated when first read."""
        hist = []
        t0 = time.time()
        self._best_verified, self._verify_pending = None, False
        self._rate_pending = False    # recomputed from the cumulative counters at the end
        it = 0
        E = int(exchange_every)
        A = self.adapt_every
        marks = set()
        if self.collect_after is not None:
            ce = self.collect_every
            first = int(self.collect_after)
            first += (-(first + 1)) % ce                  # first s >= collect_after, (s+1) % ce == 0
            marks.update(range(first, nsteps, ce))
        logs = set(range(log_every - 1, nsteps, log_every)) | ({0} if nsteps > 0 else set())
        marks.update(logs)
        if A:
            h0 = self._steps_done()
            if E:
                marks.update(range((-(h0 + 1)) % E, nsteps, E))
            marks.update(range(A - 1, nsteps, A))
            win_exch = []
            ea_last = self._exch_raw()
            acc_win = self.accepted_per_rung().sum()
        sync_logs = callback is not None
        log_steps = []
        if logs:
            if sync_logs:
                acc0 = self.accepted_per_rung().sum()
            else:
                self._lib_reserve_marks(len(logs) + 1)
                self._lib_mark(0)
        for m in sorted(marks):
            n = m + 1 - it
            if n > 0:
                self._lib_run(n, E)
                it = m + 1
            if A and E and (h0 + m + 1) % E == 0:
                rates, ea_last = self._exchange_outcomes(ea_last)
                win_exch.append(rates)
            if A and (m + 1) % A == 0:
                acc_now = self.accepted_per_rung().sum()
                wa = float(acc_now - acc_win) / (A * self.nrep)
                self.adapt(wa, np.mean(win_exch, axis=0) if win_exch else None)
                win_exch, acc_win = [], acc_now
            if (self.collect_after is not None and m >= self.collect_after
                    and (m + 1) % self.collect_every == 0):
                self._collect()
            if m in logs:
                if sync_logs:
                    best = float(self.best_chi2.min())
                    acc = (self.accepted_per_rung().sum() - acc0) / ((m + 1) * self.nrep)
                    hist.append((m + 1, best, float(acc), time.time() - t0))
                    callback(m + 1, best, hist[-1][2])
                else:
                    log_steps.append(m + 1)
                    self._lib_mark(len(log_steps))
        if it < nsteps:
            self._lib_run(nsteps - it, int(exchange_every))
        if log_steps:
            mk = self._lib_read_marks(len(log_steps) + 1)
            for q, st in enumerate(log_steps, 1):
                hist.append((st, float(mk[q, 0]), float((mk[q, 1] - mk[0, 1]) / (st * self.nrep)),
                             float(mk[q, 2])))
        self.walltime = time.time() - t0
        self.history = hist
        self._verify_pending = self._rate_pending = True
        return hist

    # Replica-exchange bookkeeping (method_a_rmc.py:332-373, 451-454): the attempt and
    # accept counters are live; exchange_rate is the value at the end of the last run, read
    # from the device when first used (a run needs no host round trip to report it).
    def _exchange_counts(self):
        m = max(self.nrep - 1, 1)
        if self._S is None:
            return np.zeros((2, m))
        return self._exch_raw().astype(float)

    @property
    def exch_attempts(self):
        return self._exchange_counts()[0]

    @property
    def exch_accepted(self):
        return self._exchange_counts()[1]

    @property
    def exchange_rate(self):
        if self._rate_pending:
            att, acc = self._exchange_counts()
            with np.errstate(invalid="ignore", divide="ignore"):
                self._exchange_rate = np.where(att > 0, acc / np.maximum(att, 1), np.nan)
            self._rate_pending = False
        return self._exchange_rate

    @property
    def best_verified(self):
        """chi^2 of the best configuration retained by the last run, recomputed from
        scratch (method_a_rmc.py:439-448).  Evaluated on first read (a pair-space path
        sum, the multiple-scattering counts and the transform), so a run that nobody
        verifies costs nothing; anything that changes the chain first settles it."""
        if self._verify_pending:
            self._best_verified = self.verify_best()
            self._verify_pending = False
        return self._best_verified

    @best_verified.setter
    def best_verified(self, value):
        self._best_verified, self._verify_pending = value, False

    def _settle(self):
        # the values the last run would have reported, before the chain can change them
        if self._S is not None:
            if self._verify_pending:
                self.best_verified
            if self._rate_pending:
                self.exchange_rate

    def verify_best(self):
        """chi^2 of the retained best, recomputed from scratch: pair-space FP64 path sum
        + multiple-scattering counts + transform (method_a_rmc.py:439-448)."""
        bsp, _, _ = self.best()
        return float(self.chi2_of(bsp[None])[0])

    def chi_k_of(self, species):
        """Unnormalised chi(k) (nconf, nsp, nk) of configurations from scratch on the
        device: pair-space path sum plus multiple scattering from the entry counts."""
        species = i32(species).reshape(-1, self.natoms)
        n = species.shape[0]
        pos = np.repeat(self.fm.cell.ideal[None], n, 0)
        chi = self.fm.chi(pos, species, normalise=False)
        if self.ms is not None:
            C = np.empty((n, self.nsp * self.nrow_e), np.int32)
            check(_lib.load().chm_sampler_counts(self._S, n, species, C))
            C.reshape(n, self.nsp, self.nrow_e)[:, :, :self.nsp * self.nshell] = 0
            chi_ms = np.empty((n, self.nsp, self.nk))
            check(_lib.load().chm_sampler_chik(self._S, n, C, chi_ms))
            chi += chi_ms
        return chi

    def counts_of(self, species):
        """Basis-row counts (nconf, nsp*nrow_e) of configurations, on the device."""
        species = i32(species).reshape(-1, self.natoms)
        C = np.empty((species.shape[0], self.nsp * self.nrow_e), np.int32)
        check(_lib.load().chm_sampler_counts(self._S, species.shape[0], species, C))
        return C

    def set_target(self, target, sigma_R):
        """Replace the target and sigma (same channels and window); the state is reset
        by the next init_state."""
        target = np.asarray(target)
        assert target.shape[0] == self.nchan
        self._tm = target[:, self.mask]
        self._sm = np.asarray(sigma_R)[:, self.mask]
        self._destroy()
        self._comp = None

    def chi2_of(self, species):
        """Reduced chi^2 of configurations (nconf, natoms), from scratch on the device."""
        species = i32(species).reshape(-1, self.natoms)
        n = species.shape[0]
        chi = self.chi_k_of(species)
        cnt = np.stack([(species == s).sum(1) for s in range(self.nsp)], 1)
        chi = chi / np.maximum(cnt, 1)[:, :, None]
        z = self.ft.to_R(chi)[:, :, self.mask]               # (n, nchan, nwin)
        zz = np.ascontiguousarray(np.stack([z.real, z.imag], -1).reshape(n, -1))
        out = np.empty(n)
        check(_lib.load().chm_sampler_chi2(self._S, n, zz, out))
        return out

    def scales(self, chi_un=None):
        """Profiled per-edge amplitudes (nrep, nchan) of the current states."""
        chi = self.chi / np.maximum(self.count, 1.0)[:, :, None]
        m = self.ft.to_R(chi)[:, :, self.mask]
        num = (m.real * self._tm.real[None] + m.imag * self._tm.imag[None]).sum(-1)
        den = (m.real ** 2 + m.imag ** 2).sum(-1)
        num = np.repeat(num.reshape(-1, self.nsp, self.nkw).sum(-1), self.nkw, 1)
        den = np.repeat(den.reshape(-1, self.nsp, self.nkw).sum(-1), self.nkw, 1)
        return num / np.maximum(den, 1e-30)

    # ------------------------------------------------------------- posterior
    def collect_rungs(self):
        beta = self._beta_rung
        k = min(self.n_collect, self.nrep)
        if self.collect_beta is None:
            return np.argsort(beta)[::-1][:k]
        return np.argsort(np.abs(np.log(beta) - np.log(self.collect_beta)))[:k]

    def _collect(self):
        idx = self.collect_rungs()
        sp = self.species[idx]
        c2 = self.chi2[idx]
        for a, b in zip(sp, c2):
            self.samples.append((a.astype(np.int32), float(b)))

    def posterior_wc(self, shell_r, i_idx=None, j_idx=None, shift=None, cell=None, nsp=None,
                     tol=0.15):
        """Warren-Cowley alpha (upper triangle) over the collected samples, for the pairs
        within `tol` of the radius `shell_r`: (mean, std, n) (method_a_rmc.py:477-491).
        The pair lists and cell default to the forward model's."""
        from .lattice import warren_cowley_radius
        if not self.samples:
            raise RuntimeError("no samples collected; set collect_after in the constructor")
        fm = self.fm
        i_idx = fm.i_idx if i_idx is None else i_idx
        j_idx = fm.j_idx if j_idx is None else j_idx
        shift = fm.d_shift if shift is None else shift
        cell = fm.cell if cell is None else cell
        n = self.nsp if nsp is None else int(nsp)
        iu = np.triu_indices(n)
        arr = np.array([warren_cowley_radius(sp, i_idx, j_idx, shift, cell, shell_r, tol=tol,
                                             nsp=n)[iu] for sp, _ in self.samples])
        return arr.mean(0), arr.std(0), len(arr)

    def alpha_series(self, nsteps, shell=0):
        """(nsteps, nsp, nsp) Warren-Cowley alpha of the recorded rung, from its pair
        counts after every step."""
        N = self.pair_series(nsteps)[..., shell].astype(float)
        return np.array([alpha_from_counts(Nt, self._comp) for Nt in N])

    def best(self):
        bc = self.best_chi2
        b = int(bc.argmin())
        sp = self._read(4, (self.nrep, self.natoms), np.int32)[b]
        return sp, np.asarray(self.fm.cell.ideal).copy(), float(bc[b])

    # ------------------------------------------------------------- teardown
    def _destroy(self):
        # (a constructor that raised leaves no handle)
        if getattr(self, "_S", None) is not None and _lib._lib is not None:
            _lib._lib.chm_sampler_destroy(self._S)
        self._S = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # bound at definition: module globals may already be gone at interpreter exit
        if _finalizing():
            return
        self._destroy()


_SIG = inspect.signature(RMCSampler.__init__)
