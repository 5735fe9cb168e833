"""The pair-count model (ShellModel) on the device against pair space, and its Cramer-Rao
bounds against the oracle's closed-form linear-Gaussian posterior."""
import sys

import numpy as np
import pytest

from oracle import forward as OF
from oracle import fourier as OT
from oracle import shellmodel as OS
from oracle import tables as OTab

pytestmark = pytest.mark.gpu

R_FIT = np.arange(1.4, 5.0001, 0.02)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def calibrated():
    from chimera_amd.scattering import k_grid
    from chimera_amd.shellmodel import ShellModel
    from chimera_amd.spectrum import Spectrum
    _say("[build] Spectrum fcc 3x3x3 FeNi, calibrated; ShellModel.from_spectrum")
    S = Spectrum(elements=("Fe", "Ni"), ncell=(3, 3, 3), rmax=5.2, ms_rmax=5.2, a0_ref=3.59,
                 sigma2=0.006, s02=0.9, kweight=(1, 2, 3), k_phys=k_grid(), R=R_FIT)
    S.set_a0(3.62)
    S.set_sigma_scale(1.15, per_shell=[1.0, 1.1, 0.95, 1.05])
    S.set_dr([[0.01, -0.015], [-0.015, 0.02]])
    S.set_c3(1.5e-4)
    S.set_window(dE0=[1.0, -1.5])
    return S, ShellModel.from_spectrum(S)


def test_counts_model_equals_pair_space(calibrated):
    """chi(k) from pair counts (device basis rows) equals the device pair-space path sum
    with the averaged MS term, at calibrated a0, sigma^2, dr, C3 (shellmodel.py:12)."""
    S, M = calibrated
    _say("[test] count space vs pair space")
    for seed in (1, 2):
        sp = S.random_config(seed)
        got = M.chi_k(M.counts_of(sp))
        ref = S.chi_k(sp)
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
        np.testing.assert_allclose(M.chi_R(M.counts_of(sp)), S.chi_R(sp, complex_out=True),
                                   rtol=0, atol=1e-11 * np.abs(S.chi_R(sp)).max())
    # random-alloy counts: every absorber keeps its coordination, alpha = 0
    N0 = M.random_counts()
    np.testing.assert_allclose(M.absorber_counts(N0), [54, 54])
    np.testing.assert_allclose(M.alpha(N0, 0), 0.0, atol=1e-14)


def test_crlb_linear_case_matches_closed_form(calibrated):
    """No nuisances, amplitude fixed: the residual is linear in the counts, so the
    finite-difference Jacobian and the posterior must equal the oracle's analytic
    Jacobian and QR posterior."""
    from chimera_amd.shellmodel import crlb
    S, M = calibrated
    _say("[test] crlb linear case")
    N = M.random_counts()
    mask = np.ones(R_FIT.size, bool)
    tgt = M.chi_R(N) * 1.03
    sig = np.full(tgt.shape, 0.05 * np.abs(tgt).max())
    res = crlb(M, N, tgt, sig, mask, None, shells=[0, 1], nuisance=(), profile_scale=False)
    B = M.basis()
    cnt = M.absorber_counts(N)
    trs = [OT.Transform(OTab.k_exp_axis(M.k, M.p["dE0"][e]), M.p["kmin"], M.p["kmax"], w, 1.0)
           for e in range(2) for w in M.p["kweight"]]
    Jo = OS.sro_jacobian(B, cnt, trs, M.chan_edge, R_FIT, mask, sig, [0, 1])
    assert np.abs(res["jacobian"] - Jo).max() <= 1e-6 * np.abs(Jo).max()
    da0 = OS.alpha_derivatives(cnt, M.z, 2, 0)
    da1 = OS.alpha_derivatives(cnt, M.z, 2, 1)
    prec = [np.abs(d).max() ** 2 for d in da0 + da1]
    cov = OS.posterior_cov(Jo, np.array(prec))
    iu = np.triu_indices(2)
    A = np.array([d[iu] for d in da0] + [np.zeros(3)] * len(da1))
    sd_o = np.sqrt(np.diag(A.T @ cov @ A))
    np.testing.assert_allclose(res["alpha_sd"], sd_o, rtol=1e-6)


def test_crlb_noise_scaling_and_nuisances(calibrated):
    """With a vanishing prior the bound is linear in the noise level; the default nuisance
    set adds a0, sigma^2 scale, MS amplitude, dE0 per edge, the free bond offsets and C3,
    and can only widen the bound on alpha."""
    from chimera_amd.shellmodel import crlb
    S, M = calibrated
    _say("[test] crlb scaling and nuisances")
    N = M.random_counts()
    mask = np.ones(R_FIT.size, bool)
    tgt = M.chi_R(N)
    sig = np.full(tgt.shape, 0.02 * np.abs(tgt).max())
    big = dict(alpha=1e6)
    r1 = crlb(M, N, tgt, sig, mask, None, shells=[0], nuisance=(), prior_sd=big)
    r2 = crlb(M, N, tgt, 3.0 * sig, mask, None, shells=[0], nuisance=(), prior_sd=big)
    np.testing.assert_allclose(np.array(r2["alpha_sd"]), 3.0 * np.array(r1["alpha_sd"]), rtol=1e-6)
    p0 = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in M.p.items()}
    rn = crlb(M, N, tgt, sig, mask, None, shells=[0], prior_sd=big)
    for k, v in p0.items():                        # parameters restored
        assert np.array_equal(np.asarray(M.p[k]), np.asarray(v)), k
    ncol = 1 + 1 + 1 + 2 + 2 + 1                   # a0, sigma, ms_amp, dE0 x2, dr x2, c3
    assert len(rn["columns"]) == r1["n_sro_directions"] + ncol
    assert np.all(np.array(rn["alpha_sd"]) >= np.array(r1["alpha_sd"]) * (1 - 1e-9))
