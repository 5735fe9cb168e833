"""Spectrum: configuration -> chi(k) / chi(R), with cheap calibration updates.

Drop-in for the calibration-parameter path of chimera/model.py:150-449 (Spectrum), the
caller that builds the sampler's inputs in chimera/fitrun.py:250-333:

  * set_a0       lattice constant: ideal sites, cell and every pair shift rescaled
                 (model.py:255-268), the path library rescaled with it;
  * set_sigma_scale  sigma^2 = reference sigma^2 x global scale x per-shell factor
                 (model.py:270-282);
  * set_dr / set_c3  bond-length offsets by pair type and third cumulants per shell
                 (model.py:284-297), forwarded to the ForwardModel's device tables;
  * set_ms       the composition-averaged multiple-scattering chi(k) (model.py:299-324,
                 chimera/mslibrary.py:304-333): entry table rows at the current lattice and
                 sigma^2, weighted by the random-alloy entry occurrences, added to every
                 chi (the sampler takes it as a constant basis row); or, with an exact path
                 library attached (model.py:326-343), the library's own rows;
  * set_window   window, k weights and per-edge E0 (model.py:352-372): edge e's channels
                 transform chi against k_exp(e) = sqrt(k^2 - ETOK dE0_e) (model.py:120-130),
                 a non-uniform axis, resampled by the device transform's interpolation.

Every evaluation runs on the device (path sum, table rows, transforms); the host holds
the setup.  Tables are the seeded synthetic form (FEFF / native table generation is not
part of this path); sigma^2 is either one value for every pair or the reference's
harmonic model (nearest-neighbour bond-stretch + transverse springs, thermal average).
"""
from __future__ import annotations

import numpy as np

from .forward import ForwardModel, thermal_sigma2
from .fourier import FourierTransform
from .lattice import CrystalSupercell, integer_counts
from .mspaths import MSLibrary
from .multift import MultiEdgeFT
from .scattering import synthetic

ETOK = 0.2624682843        # k^2 [A^-2] per eV: 2 m_e / hbar^2
NKMODEL = 320
K_LO, K_HI = 0.5, 18.0


def k_exp_axis(k_phys, dE0, k_table=None):
    """The experimenter's k axis given their E0 error dE0 (eV) (model.py:120-130):
    k_exp^2 = k^2 - ETOK dE0, floored at 1e-4 A^-2.  `k_table` (a per-edge k(p) from
    the tables) replaces k_phys when the system maps k through the tables."""
    k0 = np.asarray(k_phys if k_table is None else k_table, np.float64)
    return np.sqrt(np.maximum(k0 * k0 - ETOK * float(dE0), 1e-4))


def as_kweights(kweight):
    """One k weight or several, as a tuple of ints (model.py's as_kweights)."""
    if np.isscalar(kweight):
        return (int(kweight),)
    out = tuple(int(w) for w in kweight)
    if not out:
        raise ValueError("need at least one k weight")
    return out


def force_constant_matrix(cell, i_idx, j_idx, shift, k_bond=2.6, k_angle=0.35, nn_cut=None):
    """Harmonic force constants (3N, 3N), eV/A^2 (exafs_gpu/lattice.py:97-111): every
    nearest-neighbour bond carries a stretch spring k_bond along the bond and a
    transverse spring k_angle across it."""
    i_idx, j_idx = np.asarray(i_idx), np.asarray(j_idx)
    d = cell.ideal[j_idx] + np.asarray(shift, float) - cell.ideal[i_idx]
    r = np.linalg.norm(d, axis=1)
    cut = (r.min() * 1.05) if nn_cut is None else float(nn_cut)
    m = r <= cut
    e = d[m] / r[m][:, None]
    ee = e[:, :, None] * e[:, None, :]
    blk = k_bond * ee + k_angle * (np.eye(3)[None] - ee)           # (nb, 3, 3)
    n = cell.natoms
    Phi = np.zeros((n, 3, n, 3))
    ii, jj = i_idx[m], j_idx[m]
    np.add.at(Phi, (ii, slice(None), jj, slice(None)), -blk)
    np.add.at(Phi, (ii, slice(None), ii, slice(None)), blk)
    return Phi.reshape(3 * n, 3 * n)


class Spectrum:
    """Configuration -> chi(R), with the calibration parameters of chimera/model.py:150-449.

    System fields are passed directly (`elements`, `composition`, `lattice`, `a0_ref`);
    `sigma2` None uses the reference's thermal sigma^2 (`temperature`, `k_bond`,
    `k_angle`), a number sets it for every pair.  `k_phys` defaults to the reference's
    model grid (0.5-18 A^-1, 320 points); `R` is the output R grid (None: the FFT grid
    of the transform, as the reference's)."""

    def __init__(self, a0=None, dE0=0.0, kmin=3.0, kmax=14.0, kweight=2, sigma_scale=1.0,
                 k_bond=2.6, k_angle=0.35, temperature=300.0, ncell=(5, 5, 5), rmax=6.2,
                 dk_win=1.0, s02=1.0, ms=True, ms_amp=1.0, ms_c=1.3, sigma_shell_scale=None,
                 dr=None, c3=None, elements=("Fe", "Ni"), composition=None, lattice="fcc",
                 a0_ref=3.59, sigma2=None, k_phys=None, R=None, ms_rmax=None, device=0):
        self.elements = tuple(elements)
        self.nsp = len(self.elements)
        if self.nsp < 2:
            raise ValueError("need at least two species")
        comp = np.ones(self.nsp) if composition is None else np.asarray(composition, float)
        self.composition = comp / comp.sum()
        self.a0_ref = float(a0_ref)
        self.cell = CrystalSupercell(lattice, ncell, a0=self.a0_ref)
        self.i, self.j, self.shift0 = self.cell.neighbour_list(rmax)
        self.ideal0 = self.cell.ideal.copy()
        self.H0 = self.cell.H.copy()
        self.s02, self.dk_win, self.rmax = float(s02), float(dk_win), float(rmax)
        d0 = np.linalg.norm((self.ideal0[self.j] + self.shift0) - self.ideal0[self.i], axis=1)
        _, sh = np.unique(np.round(d0, 3), return_inverse=True)
        self.shell_idx = sh.astype(np.int32)
        self.nshell = int(self.shell_idx.max()) + 1
        self.shells_ref = np.array([d0[self.shell_idx == s].mean() for s in range(self.nshell)])
        self.shell_mult = np.array([(self.shell_idx == s).sum() / self.cell.natoms
                                    for s in range(self.nshell)])
        if sigma2 is None:
            self.Phi = force_constant_matrix(self.cell, self.i, self.j, self.shift0,
                                             k_bond=k_bond, k_angle=k_angle)
            self.sigma2_ref = thermal_sigma2(self.Phi, self.cell, self.i, self.j, self.shift0,
                                             temperature)
        else:
            self.Phi = None
            self.sigma2_ref = np.full(self.i.size, float(sigma2))
        self.sigma2_shell_ref = np.array([self.sigma2_ref[self.shell_idx == s].mean()
                                          for s in range(self.nshell)])
        self.k_phys = (np.linspace(K_LO, K_HI, NKMODEL) if k_phys is None
                       else np.asarray(k_phys, np.float64))
        self.tables = synthetic(self.elements, self.k_phys, s02=self.s02)
        self.fm = ForwardModel(self.cell, self.tables, self.i, self.j, self.shift0,
                               sigma2=self.sigma2_ref, shell=self.shell_idx, device=device)
        self._R_out = None if R is None else np.asarray(R, np.float64)
        self.R = None
        # multiple scattering: the library's classes found once at the reference lattice
        self.ms_paths = (MSLibrary(self.cell, self.nsp, self.rmax if ms_rmax is None else ms_rmax,
                                   sigma2_nn=float(self.sigma2_shell_ref[0]),
                                   r_nn=float(self.shells_ref[0]), c_ms=ms_c, amp=ms_amp)
                         if ms else None)
        self.ms_library = None
        self.chi_ms = None
        self.ms_amp, self.ms_c = float(ms_amp), float(ms_c)
        self.sigma_shell_scale = np.ones(self.nshell)
        self.kmin, self.kmax = kmin, kmax
        self.kweights = as_kweights(kweight)
        self.dE0 = np.atleast_1d(np.asarray(dE0, float)) * np.ones(self.nsp)
        self.set_a0(self.a0_ref if a0 is None else a0)
        self.set_sigma_scale(sigma_scale, sigma_shell_scale)
        self.set_dr(dr)
        self.set_c3(c3)
        self.set_ms(ms_amp, ms_c)
        self.set_window(kmin, kmax, kweight, dE0)

    # ------------------------------------------------------------- parameters
    def set_a0(self, a0):
        """Rescale the lattice to a0 (model.py:255-268)."""
        self.a0 = float(a0)
        s = self.a0 / self.a0_ref
        self.cell.ideal = np.ascontiguousarray(self.ideal0 * s)
        self.cell.H = self.H0 * s
        self.cell.box = np.diag(self.cell.H).copy()
        self.cell.a0 = self.a0
        self.cell.unit = self.a0 / 2.0
        self.shift = self.shift0 * s
        self.shells = self.shells_ref * s
        self.fm.d_shift = self.shift
        if self._ms_on():
            self.set_ms()
        return self

    def set_sigma_scale(self, s, per_shell=None):
        """Global sigma^2 multiplier, optionally times a per-shell factor (model.py:270-282)."""
        self.sigma_scale = float(s)
        if per_shell is not None:
            ps = np.asarray(per_shell, float).ravel()
            if ps.size != self.nshell:
                raise ValueError(f"need {self.nshell} per-shell factors")
            self.sigma_shell_scale = ps.copy()
        self.sigma2 = self.sigma2_ref * self.sigma_scale * self.sigma_shell_scale[self.shell_idx]
        self.fm.d_sig2 = self.sigma2
        if self._ms_on():
            self.set_ms(self.ms_amp, self.ms_c)
        return self

    def set_dr(self, dr=None):
        """Nearest-neighbour bond-length offsets by pair type, (nsp, nsp) A (model.py:284-289)."""
        self.dr = (np.zeros((self.nsp, self.nsp)) if dr is None
                   else np.asarray(dr, float).reshape(self.nsp, self.nsp))
        self.fm.set_dr(self.dr)
        return self

    def set_c3(self, c3=None):
        """Third cumulant per shell (A^3); a scalar means the first shell (model.py:291-297)."""
        c3 = np.zeros(self.nshell) if c3 is None else np.atleast_1d(np.asarray(c3, float))
        full = np.zeros(self.nshell)
        full[:c3.size] = c3[:self.nshell]
        self.c3 = full
        self.fm.set_c3(full)
        return self

    def sigma2_nn(self):
        """Mean sigma^2 of the nearest-neighbour pairs (model.py:345-348)."""
        d = np.linalg.norm((self.cell.ideal[self.j] + self.shift) - self.cell.ideal[self.i], axis=1)
        return float(self.sigma2[np.abs(d - self.shells[0]) < 0.05].mean())

    def _ms_params(self):
        return dict(sigma2_nn=self.sigma2_nn(), r_nn=float(self.shells[0]), c_ms=self.ms_c,
                    amp=self.ms_amp, a_scale=float(self.shells[0] / self.shells_ref[0]))

    def set_ms(self, amp=None, c_ms=None):
        """Rebuild the composition-averaged multiple-scattering chi(k), or, with a path
        library attached, the library's rows at the current lattice and sigma^2
        (model.py:299-324)."""
        if amp is not None:
            self.ms_amp = float(amp)
        if c_ms is not None:
            self.ms_c = float(c_ms)
        if self.ms_paths is None:
            self.chi_ms = None
            return self
        self.ms_paths.set_params(**self._ms_params())
        if self.ms_library is not None:
            self.chi_ms = None
            return self
        self.chi_ms = self.ms_paths.chi_fast(self.fm, self.composition)
        return self

    def attach_ms_library(self, library=True):
        """Exact multiple scattering (model.py:326-343): the configuration's own path
        entries replace the composition-averaged term; the sampler takes the library
        through `ms_library`.  None / False detaches."""
        if not library:
            self.ms_library = None
            return self.set_ms()
        if self.ms_paths is None:
            raise ValueError("this Spectrum was built with ms=False")
        self.ms_library = self.ms_paths
        return self.set_ms()

    def _ms_on(self):
        return self.ms_paths is not None and (self.chi_ms is not None or self.ms_library is not None)

    def set_window(self, kmin=None, kmax=None, kweight=None, dE0=None):
        """Window, k weight(s) and per-edge E0 (model.py:352-372): one transform per
        (edge, k weight), edge-major; edge e transforms against k_exp(e)."""
        if kmin is not None:
            self.kmin = float(kmin) if np.isscalar(kmin) else np.asarray(kmin, float)
        if kmax is not None:
            self.kmax = float(kmax) if np.isscalar(kmax) else np.asarray(kmax, float)
        if kweight is not None:
            self.kweights = as_kweights(kweight)
        if dE0 is not None:
            self.dE0 = np.atleast_1d(np.asarray(dE0, float)) * np.ones(self.nsp)
        lo, hi = np.atleast_1d(self.kmin), np.atleast_1d(self.kmax)
        self.ft = []
        for e in range(self.nsp):
            ke = self.k_exp(e)
            for w in self.kweights:
                self.ft.append(FourierTransform(ke, float(lo[e % lo.size]), float(hi[e % hi.size]),
                                                w, dk_win=self.dk_win, R=self._R_out))
        self.chan_edge = np.repeat(np.arange(self.nsp), len(self.kweights)).astype(np.int32)
        self.nchan = len(self.ft)
        self.R = self.ft[0].R
        self._mft = None
        return self

    def k_exp(self, e):
        """The experimental k axis of edge e (its E0 error applied)."""
        return k_exp_axis(self.k_phys, self.dE0[e])

    @property
    def kweight(self):
        return self.kweights[0] if len(self.kweights) == 1 else list(self.kweights)

    def params(self):
        return dict(elements=list(self.elements), a0=self.a0, dE0=[float(x) for x in self.dE0],
                    kmin=self.kmin, kmax=self.kmax, kweight=self.kweight,
                    sigma_scale=self.sigma_scale, sigma_shell_scale=self.sigma_shell_scale.tolist(),
                    dr=self.dr.tolist(), c3=self.c3.tolist(),
                    multiple_scattering=self.ms_paths is not None, ms_amp=self.ms_amp,
                    ms_c=self.ms_c, nn_distance=float(self.shells[0]))

    # ------------------------------------------------------------- evaluation
    def multi_ft(self):
        """MultiEdgeFT over this Spectrum's channels with chi_ms as the fixed term
        (chimera/multift.py:53-56)."""
        if self._mft is None:
            self._mft = MultiEdgeFT.from_spectrum(self)
        return self._mft

    def chi_k(self, species, pos=None):
        """chi(k) (nsp, nk) on k_phys, per absorbing atom, multiple scattering included
        (model.py:405-432)."""
        sp = np.atleast_2d(np.asarray(species, np.int32))
        p = self.cell.ideal[None] if pos is None else np.asarray(pos, float).reshape(1, -1, 3)
        chi = self.fm.chi(p, sp)[0]
        if self.chi_ms is not None:
            chi = chi + self.chi_ms
        if self.ms_library is not None:
            raise ValueError("with an exact path library the configuration's own paths are "
                             "counted by the sampler: use RMCSampler.chi_k_of")
        return chi

    def chi_R(self, species, pos=None, complex_out=False):
        """chi(R) (nchan, nR) for one configuration, |chi(R)| by default (model.py:434-445).
        chi_ms is already inside chi_k, so the channels transform it once."""
        chi = self.chi_k(species, pos)
        out = MultiEdgeFT(self.ft, chan_edge=self.chan_edge).to_R(chi[None])[0]
        return out if complex_out else np.abs(out)

    def random_config(self, seed=0, composition=None):
        """A random arrangement at the composition (model.py:451-459)."""
        n = self.cell.natoms
        comp = self.composition if composition is None else composition
        sp = np.repeat(np.arange(self.nsp), integer_counts(comp, n)).astype(np.int32)
        np.random.default_rng(seed).shuffle(sp)
        return sp
