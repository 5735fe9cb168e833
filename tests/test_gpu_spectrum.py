"""The calibration-parameter path (Spectrum: a0, sigma^2 by shell, dr, C3, per-edge E0,
composition-averaged or exact multiple scattering) on the device, against the oracle
built independently at the calibrated geometry."""
import sys

import numpy as np
import pytest

from oracle import forward as OF
from oracle import fourier as OT
from oracle import lattice as OL
from oracle import mspaths as OM
from oracle import rmc as OR
from oracle import tables as OTab

pytestmark = pytest.mark.gpu

A0_REF, A0 = 3.59, 3.63
NCELL = (3, 3, 3)
DE0 = [1.5, -2.0]
SHELL_SCALE = [1.0, 1.1, 0.9, 1.05]
DR = [[0.01, -0.02], [-0.02, 0.015]]
C3 = 2e-4
MS_AMP, MS_C = 0.8, 1.4
R_FIT = np.arange(1.4, 5.0001, 0.02)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _spectrum(exact_ms=False):
    from chimera_amd.scattering import k_grid
    from chimera_amd.spectrum import Spectrum
    _say(f"[build] Spectrum fcc {NCELL} exact_ms={exact_ms}")
    S = Spectrum(elements=("Fe", "Ni"), ncell=NCELL, rmax=5.2, ms_rmax=5.2, a0_ref=A0_REF,
                 sigma2=0.006, s02=0.9, kweight=(1, 2, 3), k_phys=k_grid(), R=R_FIT)
    if exact_ms:
        S.attach_ms_library()
    S.set_a0(A0)
    S.set_sigma_scale(1.2, per_shell=SHELL_SCALE)
    S.set_dr(DR)
    S.set_c3(C3)
    S.set_ms(amp=MS_AMP, c_ms=MS_C)
    S.set_window(dE0=DE0)
    return S


class _Oracle:
    """The same calibrated model from the oracle's own lattice, tables and paths."""

    def __init__(self, exact_ms=False):
        self.k, self.f, self.d, self.lam, self.s02 = OTab.synthetic_ss(OTab.k_grid(), [26, 28], s02=0.9)
        cell = OL.Supercell("fcc", NCELL, A0)
        self.cell = cell
        self.i, self.j, self.sh = cell.neighbour_list(5.2)
        self.shell, self.radii = OL.shell_index(cell.ideal, self.i, self.j, self.sh)
        self.sig2 = 0.006 * 1.2 * np.asarray(SHELL_SCALE)[self.shell]
        ns = self.radii.size
        self.dr = np.zeros((ns, 2, 2))
        self.dr[0] = DR
        self.c3 = np.zeros(ns)
        self.c3[0] = C3
        self.trs = [OT.Transform(OTab.k_exp_axis(self.k, DE0[e]), 3.0, 14.0, w, 1.0)
                    for e in range(2) for w in (1, 2, 3)]
        self.ce = np.repeat(np.arange(2), 3)
        # paths: the library's classes at the reference lattice, lengths scaled to A0
        cl = OM.enumerate_classes("fcc", A0_REF, 5.2)
        unit = A0 / 2.0
        s2nn = 0.006 * 1.2 * SHELL_SCALE[0]
        self.T = OTab.ms_entry_table(self.k, [26, 28], self.f, self.lam, OM.scattering_sequences(cl, 2),
                                     [c.reff_units() * unit for c in cl], s02=0.9, sigma2_nn=s2nn,
                                     r_nn=float(self.radii[0]), c_ms=MS_C, amp=MS_AMP)
        self.cl = cl
        self.nblk = int(sum(2 ** c.nslot for c in cl))
        E = OM.expected_counts(cl, 2, [0.5, 0.5])
        nb = self.nblk
        self.chi_ms = None if exact_ms else np.stack(
            [E[a * nb:(a + 1) * nb] @ self.T[a * nb:(a + 1) * nb] for a in range(2)])
        self.ms = OM.SupercellPaths(cl, 2, cell.ideal, A0, NCELL) if exact_ms else None

    def chi(self, sp):
        """Normalised chi(k): pair space + (averaged MS, or the configuration's own paths)."""
        out = OF.chi_pairs(self.k, self.f, self.d, self.lam, self.s02, self.cell.ideal, sp,
                           self.i, self.j, self.sh, self.sig2, shell=self.shell, dr=self.dr,
                           c3=self.c3, normalise=False)
        if self.ms is not None:
            C = self.ms.counts(sp).astype(float)
            nb = self.nblk
            for a in range(2):
                out[a] += C[a * nb:(a + 1) * nb] @ self.T[a * nb:(a + 1) * nb]
        out /= np.maximum(np.bincount(np.asarray(sp), minlength=2), 1)[:, None]
        return out if self.chi_ms is None else out + self.chi_ms

    def problem(self, target, sigma):
        return OR.Problem(self.k, self.f, self.d, self.lam, self.s02, self.cell.ideal, self.i,
                          self.j, self.sh, self.shell, self.sig2, self.trs, self.ce, target, sigma,
                          R=R_FIT, ms=self.ms, T=None if self.ms is None else self.T, dr=self.dr,
                          c3=self.c3, chi_extra=self.chi_ms)


@pytest.fixture(scope="module")
def averaged():
    return _spectrum(False), _Oracle(False)


def test_geometry_and_sigma2_follow_calibration(averaged):
    S, O = averaged
    _say("[test] calibrated geometry")
    np.testing.assert_allclose(S.cell.ideal, O.cell.ideal, rtol=0, atol=1e-12)
    assert np.array_equal(S.i, O.i) and np.array_equal(S.j, O.j)
    np.testing.assert_allclose(S.shift, O.sh, rtol=0, atol=1e-12)
    np.testing.assert_allclose(S.shells, O.radii, rtol=1e-14)
    np.testing.assert_allclose(S.fm.d_sig2, O.sig2, rtol=1e-15)
    np.testing.assert_allclose(S.sigma2_nn(), 0.006 * 1.2, rtol=1e-15)


def test_chi_k_with_averaged_ms(averaged):
    S, O = averaged
    _say("[test] chi(k) calibrated + averaged MS")
    sp = S.random_config(3)
    got = S.chi_k(sp)
    ref = O.chi(sp)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    np.testing.assert_allclose(S.chi_ms, O.chi_ms, rtol=0, atol=1e-12 * np.abs(O.chi_ms).max())


def test_chi_R_per_edge_E0(averaged):
    S, O = averaged
    _say("[test] chi(R) with per-edge E0 axes")
    sp = S.random_config(4)
    got = S.chi_R(sp, complex_out=True)
    ref = OT.multi_to_R(O.trs, O.ce, O.chi(sp), R=R_FIT)
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()
    # the E0 shift moves the transform: the zero-shift axis gives a different chi(R)
    t0 = OT.Transform(O.k, 3.0, 14.0, 2, 1.0)
    assert np.abs(t0.dft(O.chi(sp)[0], R_FIT) - ref[1]).max() > 1e-3 * np.abs(ref[1]).max()


def _target(S, O, seed):
    rng = np.random.default_rng(seed)
    chi_t = O.chi(S.random_config(seed))
    tgt = OT.multi_to_R(O.trs, O.ce, chi_t, R=R_FIT)
    tgt = tgt + 0.01 * np.abs(tgt).max() * (rng.standard_normal(tgt.shape)
                                            + 1j * rng.standard_normal(tgt.shape))
    return tgt, np.full(tgt.shape, 0.02 * np.abs(tgt).max())


def _trajectory(S, O, nsteps, exch, seed, ms_library=None):
    from chimera_amd.method_a_rmc import RMCSampler
    tgt, sig = _target(S, O, 77)
    mask = np.ones(R_FIT.size, bool)
    smp = RMCSampler(S.fm, S.multi_ft(), tgt, mask, sig, nrep=6, seed=seed, swap_frac=1.0,
                     ms_library=ms_library)
    smp.set_recording(log_steps=nsteps)
    sp0 = S.random_config(5)
    smp.init_state(sp0, S.cell.ideal)
    _say(f"[run] device sampler {nsteps} steps")
    smp.run(nsteps, exchange_every=exch, log_every=10 ** 9)
    log = smp.move_log(nsteps)
    R = OR.PlainSampler(O.problem(tgt, sig), nrep=6, seed=seed)
    R.init_state(sp0)
    _say("[run] oracle sampler")
    R.run(nsteps, exchange_every=exch)
    for t in range(nsteps):
        by_rung = {int(r[3]): r for r in log[t]}
        for rung in range(6):
            a, b, acc, c2, x, u = R.log[t][rung]
            g = by_rung[rung]
            assert (int(g[0]), int(g[1]), int(g[2])) == (a, b, acc), f"step {t} rung {rung}"
    assert np.array_equal(smp.species, R.species)
    np.testing.assert_allclose(smp.chi2, R.chi2, rtol=1e-9)
    assert smp.accepted_per_rung().sum() > 0
    np.testing.assert_allclose(smp.best_verified, smp.best_chi2.min(), rtol=1e-9)
    return smp


def test_sampler_with_calibrated_spectrum(averaged):
    """The device sampler on a calibrated Spectrum (a0, sigma^2 by shell, dr, C3,
    per-edge E0, averaged MS as chi_extra) follows the oracle move by move."""
    S, O = averaged
    _trajectory(S, O, 40, 10, seed=21)


def test_sampler_with_exact_library_at_scaled_lattice():
    """Exact multiple scattering attached, lattice rescaled: the library's rows follow
    the lattice and sigma^2; the chain follows the oracle move by move."""
    S = _spectrum(True)
    O = _Oracle(True)
    assert S.chi_ms is None and S.ms_library is not None
    _trajectory(S, O, 30, 10, seed=8, ms_library=S.ms_library)


def test_thermal_sigma2_spectrum_runs():
    """The reference's default sigma^2 (harmonic springs at 300 K): one value per shell
    by symmetry, positive, growing with distance; the sampler runs on it."""
    from chimera_amd.scattering import k_grid
    from chimera_amd.method_a_rmc import RMCSampler
    from chimera_amd.spectrum import Spectrum
    _say("[build] thermal Spectrum")
    S = Spectrum(elements=("Fe", "Ni"), ncell=NCELL, rmax=5.2, ms=False, a0_ref=A0_REF,
                 k_phys=k_grid(), R=R_FIT, kweight=2)
    s2 = S.sigma2_shell_ref
    for s in range(S.nshell):
        v = S.sigma2_ref[S.shell_idx == s]
        assert np.ptp(v) <= 1e-12 * v.mean()
    assert np.all(s2 > 0) and s2[0] < s2[-1]
    sp = S.random_config(1)
    tgt = S.chi_R(S.random_config(2), complex_out=True)
    smp = RMCSampler(S.fm, S.multi_ft(), tgt, np.ones(R_FIT.size, bool),
                     np.full(tgt.shape, 0.05 * np.abs(tgt).max()), nrep=4, seed=1, swap_frac=1.0)
    smp.init_state(sp, S.cell.ideal)
    smp.run(20, exchange_every=5, log_every=10 ** 9)
    np.testing.assert_allclose(smp.chi2, smp.chi2_of(smp.species), rtol=1e-9)
