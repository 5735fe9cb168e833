"""PairRMCSampler: the reference sampler with every move it makes, on the pair-space
device engine (libchimera_hip chm_psampler_*).

Drop-in for methods/method_a_rmc.py:63-496 where the count-space engine does not apply:
displacement moves (swap_frac < 1, the default 0.75, and the displacement a swap falls
back to when 8 draws find no partner, kernels.py:403-421), the Boltzmann energy prior
(method_a_rmc.py:147-157, 310-326), an |chi(R)| target (a real target,
method_a_rmc.py:186-188) and starts off the ideal lattice.  RMCSampler(...) returns one
of these by itself for such arguments, and a count-space sampler turns into one when it
is started off the lattice.

Per replica the device keeps what the reference keeps: species, positions, the distance
and type of every pair, the unnormalised chi(k), chi^2, the best state with positions.
A step evaluates only the terms of the pairs the move touches (new minus old, as
kernels.py:243-325), transforms the trial chi on the window (one FP64 GEMM per edge
against the DFT matrix of resample + window + FFT), profiles the per-edge scale and
takes the Metropolis decision, all on the device; exchange periods replay a captured
hipGraph.  Arrays come back as NumPy in rung order.

Decisions shared with the count-space engine (the reference's own cannot be replayed
without CuPy): the acceptance and exchange uniforms are counter-based splitmix64 streams
keyed by (seed, step, rung) and (seed, exchange number); exchange swaps temperatures
between slots, not configurations.
"""
from __future__ import annotations

import sys

import numpy as np

from . import _lib
from ._lib import check, f64, i32
from .method_a_rmc import RMCSampler, _beta_ladder

KB_EV = 8.617333262e-5   # eV / K


def ms_entry_counts(ms, geom, species):
    """(nsp * nblk,) multiple-scattering entry counts of one configuration from the
    absorber-role stencil records (SupercellMS.counts, chimera/mslibrary.py:436-451):
    for every site as absorber and every record, the entry s_abs * nblk + cls_off +
    sum_q s_q nsp^q of the slot species, with the record's weight."""
    xyz, box, ncell, nbasis, bpar = geom
    sp = np.asarray(species, np.int64)
    n = sp.size
    x0, y0, z0 = xyz & 1023, (xyz >> 10) & 1023, (xyz >> 20) & 1023
    C = np.zeros(ms.nsp * ms.nblk, np.int64)
    for rec in np.asarray(ms.stencil[:ms.nabsrec], np.int64):
        nslot, w = int(rec[12]), int(rec[13])
        cls_off = int(rec[14] & 0xFF) | (int(rec[15]) << 8)
        e = sp * ms.nblk + cls_off
        pw = 1
        for q in range(1, nslot + 1):
            X = (x0 + rec[3 * q]) % box[0]
            Y = (y0 + rec[3 * q + 1]) % box[1]
            Z = (z0 + rec[3 * q + 2]) % box[2]
            par = (X & 1) | ((Y & 1) << 1) | ((Z & 1) << 2)
            site = (((X >> 1) * ncell[1] + (Y >> 1)) * ncell[2] + (Z >> 1)) * nbasis + bpar[par]
            e = e + sp[site] * pw
            pw *= ms.nsp
        C += np.bincount(e, minlength=C.size) * w
    assert n == xyz.size
    return C


class PairRMCSampler(RMCSampler):
    engine = "pair"

    def __init__(self, fm, ft, target, mask, sigma_R, nrep=48, tmin=0.02, tmax=5.0, seed=0,
                 swap_frac=0.75, disp_sigma=0.015, max_disp=0.30, nsp=None,
                 collect_after=None, collect_every=25, adapt_every=0, target_accept=0.35,
                 profile_scale=True, collect_beta=0.5, n_collect=4, energy_prior=None,
                 ms_library=None, precision="f64", reanchor_every=512, accumulator="f64",
                 paired=True):
        """Arguments as methods/method_a_rmc.py:64-69.  The pair-space engine computes in
        FP64 only: `precision` must be "f64" (`reanchor_every`, `accumulator` and
        `paired` belong to the count-space engine and are ignored)."""
        self._ctor = {k: v for k, v in locals().items() if k not in ("self", "__class__")}
        if precision != "f64":
            raise ValueError("the pair-space sampler computes in FP64 (precision='f64')")
        self.precision = "f64"
        self.fm, self.ft = fm, ft
        self.nsp = int(nsp) if nsp is not None else int(fm.nsp)
        self.nrep = int(nrep)
        self.seed = int(seed)
        self.swap_frac = float(swap_frac)
        self.disp_sigma = float(disp_sigma)
        self.max_disp = float(max_disp)
        self.collect_after = collect_after
        self.collect_every = int(collect_every)
        self.adapt_every = int(adapt_every)
        self.target_accept = float(target_accept)
        self.profile_scale = bool(profile_scale)
        self.collect_beta = None if collect_beta is None else float(collect_beta)
        self.n_collect = int(n_collect)
        self.samples = []
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
        order = np.argsort(key)
        rkey = j_np.astype(np.int64) * self.natoms + i_np
        pos_ = np.searchsorted(key[order], rkey)
        self.rev = i32(order[np.clip(pos_, 0, key.size - 1)])
        assert np.array_equal(i_np[self.rev], j_np) and np.array_equal(j_np[self.rev], i_np), \
            "neighbour list is not symmetric"
        self._beta_rung = _beta_ladder(tmin, tmax, self.nrep)
        self.complex = bool(np.iscomplexobj(target))
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
        self.nfit = self.nwin * self.nchan * (2 if self.complex else 1)
        self._tm = target[:, self.mask]
        self._sm = np.asarray(sigma_R)[:, self.mask]
        nshell = int(np.asarray(fm.shell).max()) + 1
        self.energy_beta = 0.0
        self._etab = None
        if energy_prior is not None:
            V, T_ann = energy_prior
            V = np.asarray(V, float)
            tab = np.zeros((nshell, self.nsp * self.nsp))
            m = min(nshell, V.shape[0])
            tab[:m] = 0.5 * V[:m].reshape(m, self.nsp * self.nsp)
            self._etab = tab.reshape(-1)
            self.energy_beta = 1.0 / (KB_EV * float(T_ann))
        self._S = None
        self._comp = None
        self._log_cap = 0
        self._exchange_rate = None
        self._rate_pending = False
        self._best_verified = None
        self._verify_pending = False
        self._build_hooks = []

    # ------------------------------------------------------------------ setup
    def _transform_rows(self):
        """Te (nsp, nk, nkw*nwin*2): the window transform of a unit chi at every k; Rx:
        that of chi_extra."""
        dev = self.ft.device() if hasattr(self.ft, "device") else None
        if dev is None:
            raise ValueError("ft must be a chimera_amd MultiEdgeFT")
        nk, ncol_e = self.nk, self.nkw * self.nwin * 2
        eye = np.eye(nk)
        Te = np.empty((self.nsp, nk, ncol_e))
        for e in range(self.nsp):
            Te[e] = dev.apply_edge(e, eye)[:, :, self.mask, :].reshape(nk, ncol_e)
        extra = getattr(self.ft, "chi_extra", None)
        Rx = None
        if extra is not None:
            Rx = np.stack([dev.apply_edge(e, np.asarray(extra)[e][None])[0][:, self.mask, :].reshape(-1)
                           for e in range(self.nsp)])
        return Te, Rx

    def _create(self):
        fm = self.fm
        Te, Rx = self._transform_rows()
        self._Rx_keep = None if Rx is None else f64(Rx)
        if self.complex:
            tgt = np.stack([self._tm.real, self._tm.imag], -1).reshape(-1)
        else:
            tgt = np.asarray(self._tm, float).reshape(-1)
        h = _lib.vp()
        lib = _lib.load()
        check(lib.chm_psampler_create(
            fm.ctx.h, self.nrep, f64(fm.cell.ideal), self.rev, self.nkw, self.nwin, f64(Te),
            None if Rx is None else self._Rx_keep.ctypes.data, f64(tgt),
            f64(self._sm.reshape(-1)), 0 if self.complex else 1, int(self.profile_scale),
            _lib.C.byref(h)))
        self._S = h
        if self._etab is not None:
            self._etab_keep = f64(self._etab)
            check(lib.chm_psampler_set_prior(h, self._etab_keep.ctypes.data, self.energy_beta))
        if self.ms is not None:
            self._setup_ms()

    def _setup_ms(self):
        xyz, box, ncell, nbasis, bpar = self._lattice_geometry()
        self._geom = (np.asarray(xyz, np.int64), box, ncell, nbasis, bpar)
        g = _lib.Geom()
        g.natoms, g.nsp, g.ncoord = self.natoms, self.nsp, self.ncoord
        g.nshell = int(self.fm.nshell)
        g.nblk = self.ms.nblk
        g.nstencil, g.nabsrec = self.ms.nstencil, self.ms.nabsrec
        g.nkw, g.nwin = self.nkw, self.nwin
        for d in range(3):
            g.box[d] = int(box[d])
            g.ncell[d] = int(ncell[d])
        g.nbasis = int(nbasis)
        for q in range(8):
            g.basis_of_parity[q] = int(bpar[q])
        g.has_const = 0
        self._ms_T = f64(self.ms.table(self.fm))
        self._ms_sten = np.ascontiguousarray(self.ms.stencil, np.int8)
        check(_lib.load().chm_psampler_set_ms(self._S, _lib.C.byref(g), i32(xyz), self._ms_sten,
                                              self._ms_T))

    def _chi_ms(self, species):
        """Unnormalised MS chi (nsp, nk) of one configuration from its entry counts."""
        C = ms_entry_counts(self.ms, self._geom, species).astype(float)
        nb = self.ms.nblk
        return np.stack([C[a * nb:(a + 1) * nb] @ self._ms_T[a * nb:(a + 1) * nb]
                         for a in range(self.nsp)])

    # ------------------------------------------------------------------ state
    def init_state(self, species0, pos0):
        self._settle()
        species0 = i32(species0)
        pos0 = f64(pos0).reshape(self.natoms, 3)
        if self._S is None:
            self._create()
            if self._log_cap:
                self.set_recording(self._log_cap)
        self._push_moves()
        self._comp = np.bincount(species0, minlength=self.nsp)
        chi_ms = None if self.ms is None else f64(self._chi_ms(species0))
        check(_lib.load().chm_psampler_init(self._S, species0, pos0,
                                            None if chi_ms is None else chi_ms.ctypes.data,
                                            f64(self._beta_rung), int(self.seed) & ((1 << 64) - 1)))
        self._chi_ms_keep = chi_ms
        self.samples = []

    def _push_moves(self):
        check(_lib.load().chm_psampler_set_moves(self._S, self.swap_frac, self.disp_sigma, self.max_disp))

    def set_recording(self, log_steps=0, series_steps=0, series_rung=None):
        """Keep the move log (a, b, is-swap | accepted << 1, slot per rung and step) and the
        trial chi^2 of the last log_steps steps on the device."""
        if series_steps:
            raise ValueError("the pair-space sampler records the move log only")
        self._log_cap = int(log_steps)
        if self._S is not None:
            check(_lib.load().chm_psampler_set_recording(self._S, self._log_cap))

    def move_log(self, nsteps):
        return self._read(12, (self._log_cap, self.nrep, 4), np.int32)[:nsteps]

    def trial_chi2_log(self, nsteps):
        return self._read(13, (self._log_cap, self.nrep), np.float64)[:nsteps]

    def save_state(self):
        raise NotImplementedError("checkpoints are kept by the count-space sampler only")

    def load_state(self, blob):
        raise NotImplementedError("checkpoints are kept by the count-space sampler only")

    def _read(self, what, shape, dtype):
        out = np.empty(shape, dtype)
        check(_lib.load().chm_psampler_read(self._S, what, out.ctypes.data, out.nbytes))
        return out

    # ------------------------------------------------------------------ hooks
    def _lib_run(self, n, E):
        self._push_moves()
        check(_lib.load().chm_psampler_run(self._S, int(n), int(E)))

    def _lib_set_beta(self, ladder):
        check(_lib.load().chm_psampler_set_beta(self._S, f64(ladder)))

    def _lib_exchange(self):
        check(_lib.load().chm_psampler_exchange(self._S))

    def _lib_reserve_marks(self, n):
        check(_lib.load().chm_psampler_reserve_marks(self._S, int(n)))

    def _lib_mark(self, i):
        check(_lib.load().chm_psampler_mark(self._S, int(i)))

    def _lib_read_marks(self, n):
        mk = np.empty((n, 3))
        check(_lib.load().chm_psampler_read_marks(self._S, int(n), mk))
        return mk

    def _steps_done(self):
        return int(self._read(11, (1,), np.int64)[0])

    def _exch_raw(self):
        return self._read(10, (2, max(self.nrep - 1, 1)), np.int64)

    # ------------------------------------------------------------- properties
    @property
    def species(self):
        return self._read(0, (self.nrep, self.natoms), np.int32)

    @property
    def chi2(self):
        return self._read(1, (self.nrep,), np.float64)

    @property
    def best_chi2(self):
        return self._read(2, (self.nrep,), np.float64)

    @property
    def pos(self):
        return self._read(4, (self.nrep, self.natoms, 3), np.float64)

    @property
    def chi(self):
        """Unnormalised chi(k) (nrep, nsp, nk), rung order."""
        return self._read(6, (self.nrep, self.nsp, self.nk), np.float64)

    @property
    def r(self):
        return self._read(7, (self.nrep, self.npairs), np.float64)

    @property
    def ptype(self):
        return self._read(8, (self.nrep, self.npairs), np.int32)

    @property
    def count(self):
        return np.repeat(self._comp[None].astype(float), self.nrep, 0)

    def accepted_per_rung(self):
        return self._read(9, (self.nrep,), np.int64)

    # ------------------------------------------------------------------ moves
    def adapt(self, window_acc, window_exch):
        """method_a_rmc.py:376-398: the displacement step moves toward target_accept by
        (acc / target)^0.35 clipped to [0.7, 1.4] (and the step to [1e-4, 0.15]); the
        ladder as the count-space sampler retunes it."""
        super().adapt(window_acc, window_exch)
        self._push_moves()

    def energy_delta(self):
        """Energy change (eV) of the move each configuration last proposed, under the
        pair interactions (method_a_rmc.py:310-319), in the current rung order (the value
        travels with its configuration through exchanges)."""
        if self._etab is None:
            raise RuntimeError("no energy_prior was given")
        return self._read(15, (self.nrep,), np.float64)

    def energy(self, species=None):
        """Pair-interaction energy (eV) of every replica's configuration
        (method_a_rmc.py:321-326), or of the given configurations."""
        if self._etab is None:
            raise RuntimeError("no energy_prior was given")
        sp = self.species if species is None else i32(species).reshape(-1, self.natoms)
        nn = self.nsp * self.nsp
        sh = np.asarray(self.fm.shell, np.int64)
        t = sp[:, self.fm.i_idx].astype(np.int64) * self.nsp + sp[:, self.fm.j_idx]
        return self._etab[sh[None, :] * nn + t].sum(axis=1)

    # ------------------------------------------------------------- results
    def chi_k_of(self, species, pos=None):
        """Unnormalised chi(k) (nconf, nsp, nk) of configurations from scratch on the
        device (pair path sum at the given positions, default ideal; plus MS)."""
        species = i32(species).reshape(-1, self.natoms)
        n = species.shape[0]
        pos = (np.repeat(self.fm.cell.ideal[None], n, 0) if pos is None
               else f64(pos).reshape(n, self.natoms, 3))
        chi = self.fm.chi(pos, species, normalise=False)
        if self.ms is not None:
            chi = chi + np.stack([self._chi_ms(sp) for sp in species])
        return chi

    def chi2_of(self, species, pos=None):
        """Reduced chi^2 of configurations (with positions, default ideal), from scratch."""
        chi = f64(self.chi_k_of(species, pos))
        out = np.empty(chi.shape[0])
        check(_lib.load().chm_psampler_chi2(self._S, chi.shape[0], chi, out))
        return out

    def best(self):
        bc = self.best_chi2
        b = int(bc.argmin())
        sp = self._read(3, (self.nrep, self.natoms), np.int32)[b]
        pos = self._read(5, (self.nrep, self.natoms, 3), np.float64)[b]
        return sp, pos, float(bc[b])

    def verify_best(self):
        """chi^2 of the retained best (species and positions), recomputed from scratch:
        pair path sum + MS counts + transform (method_a_rmc.py:439-448)."""
        bsp, bpos, _ = self.best()
        return float(self.chi2_of(bsp[None], bpos[None])[0])

    def scales(self, chi_un=None):
        """Profiled per-edge amplitudes (nrep, nchan) of the current states."""
        chi = (self.chi if chi_un is None else np.asarray(chi_un)) / np.maximum(self.count, 1.0)[:, :, None]
        m = self.ft.to_R(chi)[:, :, self.mask]
        if self.complex:
            num = (m.real * self._tm.real[None] + m.imag * self._tm.imag[None]).sum(-1)
            den = (m.real ** 2 + m.imag ** 2).sum(-1)
        else:
            m = np.abs(m)
            num = (m * self._tm[None]).sum(-1)
            den = (m * m).sum(-1)
        num = np.repeat(num.reshape(-1, self.nsp, self.nkw).sum(-1), self.nkw, 1)
        den = np.repeat(den.reshape(-1, self.nsp, self.nkw).sum(-1), self.nkw, 1)
        return num / np.maximum(den, 1e-30)

    def set_target(self, target, sigma_R):
        """Replace the target and sigma (same channels and window); the state is reset by
        the next init_state."""
        target = np.asarray(target)
        assert target.shape[0] == self.nchan
        if bool(np.iscomplexobj(target)) != self.complex:
            raise ValueError("the new target must be complex (or real) as the old one")
        self._tm = target[:, self.mask]
        self._sm = np.asarray(sigma_R)[:, self.mask]
        self._destroy()

    # ------------------------------------------------------------- teardown
    def _destroy(self):
        if getattr(self, "_S", None) is not None and _lib._lib is not None:
            _lib._lib.chm_psampler_destroy(self._S)
        self._S = None

    def __del__(self, _finalizing=sys.is_finalizing):
        if _finalizing():
            return
        self._destroy()
