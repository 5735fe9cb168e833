"""The drop-in boundary against the reference's interface, through the C ABI: the
resample of FourierTransform, a reference-style ms_library object, the exact keyword set
of the reference's fit driver and the in-place device-attribute updates its callers make."""
import sys

import numpy as np
import pytest

from oracle import forward as OF
from oracle import fourier as OT

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tern():
    from chimera_amd import workloads as W
    _say("[build] boundary t108")
    return W.Problem(W.Workload("bd108", "fcc", (3, 3, 3), 3.59, ("Cr", "Co", "Ni"), 5.2, True, 6))


@pytest.mark.parametrize("grid", ["uniform", "shifted", "coarse"])
def test_resample_equals_oracle_bit_for_bit(grid):
    """FourierTransform.resample (exafs_gpu/fourier.py:69-73) on the device equals the
    oracle's NumPy resample exactly, on the model grid, a shifted grid (k_min > 0, so
    the clamp and the outside-the-grid zeros are exercised) and a coarse non-uniform one."""
    from chimera_amd.fourier import FourierTransform
    k = {"uniform": np.arange(321) * 0.05,
         "shifted": 0.7 + np.arange(250) * 0.05,
         "coarse": np.sort(np.r_[0.0, np.random.default_rng(1).uniform(0.1, 17.0, 150)])}[grid]
    rng = np.random.default_rng(2)
    chi = rng.standard_normal((4, 2, k.size))
    ft = FourierTransform(k, 3.0, 14.0, 2)
    ref = OT.Transform(k, 3.0, 14.0, 2)
    got = ft.resample(chi)
    want = ref.resample(chi)
    assert got.shape == (4, 2, 2048)
    assert np.array_equal(got, want)


def test_reference_style_ms_library_object(tern):
    """An object of the reference's ms_library protocol (chimera/mslibrary.py:341-492:
    lib.classes with representatives, nsp, nblk and the entry table T left by
    set_params) passed as ms_library runs the same chain as the library it stands for."""
    P = tern
    T = P.ms.table(P.fm)

    class _Lib:
        classes = [type("Cls", (), {"rep": c.rep.copy()})() for c in P.ms.classes]
        nblk = P.ms.nblk

    class _SupercellMS:          # the attributes the reference's SupercellMS carries
        lib = _Lib()
        nsp = P.nsp
        nblk = P.ms.nblk

        def set_params(self, *a, **kw):
            self.T = T
            return self

    from chimera_amd.method_a_rmc import RMCSampler
    obj = _SupercellMS().set_params(0.006, float(P.radii[0]))
    runs = []
    for lib in (P.ms, obj):
        S = RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma, nrep=6, seed=4, swap_frac=1.0,
                       ms_library=lib)
        S.init_state(P.cfg0, P.cell.ideal)
        S.set_recording(log_steps=40)
        S.init_state(P.cfg0, P.cell.ideal)
        S.run(40, exchange_every=10, log_every=10 ** 9)
        runs.append((S.move_log(40).copy(), S.species.copy(), S.chi2.copy()))
        assert S.nrow_e == P.sampler.nrow_e
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][2], runs[1][2])
    with pytest.raises(ValueError):
        RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma, nrep=6, swap_frac=1.0,
                   ms_library=type("NoT", (), {"lib": _Lib(), "nsp": 3, "T": None})())


def test_fit_driver_keywords_and_results(tern):
    """The exact constructor keywords and calls of the reference's fit driver
    (chimera/fitrun.py:302-333) through the drop-in: samples every collect_every steps
    after collect_after from the collection rungs, a finite exchange rate, best() whose
    chi^2 a from-scratch recompute confirms."""
    P = tern
    from chimera_amd.method_a_rmc import RMCSampler
    n = 300
    smp = RMCSampler(P.fm, P.ft, P.target, P.mask, P.sigma,
                     nrep=6, seed=int(7), swap_frac=1.0,
                     max_disp=0.30, disp_sigma=0.015,
                     collect_after=int(n * 0.6),
                     collect_every=50, profile_scale=True,
                     collect_beta=0.5, energy_prior=None,
                     ms_library=P.ms)
    smp.init_state(np.asarray(P.cfg0, np.int32), np.asarray(P.cell.ideal, float))
    c20 = float(smp.chi2.min())
    smp.run(n, log_every=10 ** 9)
    xr = float(np.nanmean(smp.exchange_rate))
    spb, posb, c2 = smp.best()
    assert np.isfinite(c20) and 0.0 <= xr <= 1.0
    assert c2 <= c20 + 1e-12
    np.testing.assert_allclose(smp.chi2_of(spb[None])[0], c2, rtol=1e-9)
    np.testing.assert_allclose(smp.best_verified, c2, rtol=1e-9)
    assert np.allclose(posb, P.cell.ideal)
    # (it + 1) % 50 == 0 for it >= 180: steps 200, 250, 300; four rungs each
    assert len(smp.samples) == 3 * 4
    for sp, c in smp.samples:
        assert sp.dtype == np.int32 and sp.shape == (P.cell.natoms,) and np.isfinite(c)
    mean, std, nsamp = smp.posterior_wc(float(P.radii[0]))
    assert nsamp == 12 and mean.shape == (6,)


def test_in_place_device_attribute_updates(tern):
    """Callers assign fm.d_shift and fm.d_sig2 in place (chimera/model.py:263, 278): the
    next path sum uses the new values (lattice scaled by 1.01, sigma^2 doubled)."""
    P = tern
    fm = P.fm
    s = 1.01
    shift0, sig0 = fm.d_shift.copy(), fm.d_sig2.copy()
    try:
        fm.d_shift = np.ascontiguousarray(shift0 * s, np.float64)
        fm.d_sig2 = np.ascontiguousarray(sig0 * 2.0, np.float64)
        pos = P.cell.ideal * s
        got = fm.chi(pos[None], P.cfg0[None], normalise=False)[0]
        want = OF.chi_pairs(P.k, P.tables.f, P.tables.delta, P.tables.lam, P.tables.s02, pos,
                            P.cfg0, P.i, P.j, shift0 * s, sig0 * 2.0, normalise=False)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    finally:
        fm.d_shift = shift0
        fm.d_sig2 = sig0


def test_chi_reference_and_device_attributes(tern):
    """ForwardModel.chi_reference (exafs_gpu/forward.py:265-293: normalised by absorber atom
    count, optional sigma^2 override) equals the oracle's normalised pair sum at both sigma^2
    values and leaves the model's own sigma^2 in place; the reference's device attribute
    names read back the model's arrays in the reference's shapes."""
    P = tern
    fm = P.fm
    rng = np.random.default_rng(4)
    pos = P.cell.ideal + rng.normal(0, 0.03, P.cell.ideal.shape)
    sp = P.cfg0
    sig_alt = np.full(fm.npairs, 0.009)
    for s2 in (None, sig_alt):
        got = fm.chi_reference(pos[None], sp[None], sigma2=s2)[0]
        want = OF.chi_pairs(P.k, P.tables.f, P.tables.delta, P.tables.lam, P.tables.s02, pos, sp,
                            P.i, P.j, P.shift, P.sig2 if s2 is None else s2, normalise=True)
        assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    assert np.array_equal(fm.d_sig2, P.sig2)
    n = fm.nsp
    assert np.array_equal(fm.d_i, P.i) and np.array_equal(fm.d_j, P.j)
    assert np.array_equal(fm.d_shell, P.shell) and np.array_equal(fm.d_k, P.k)
    assert fm.d_delta.shape == (n * n, fm.nk) and fm.d_f.shape == (n, fm.nk)
    assert fm.d_dr.shape == (fm.nshell * n * n,) and fm.d_s2tab.shape == (fm.nshell * n * n,)
    assert fm.d_c3.shape == (fm.nshell,) and fm.d_lam.shape == (fm.nk,)
