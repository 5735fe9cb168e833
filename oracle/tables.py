"""Synthetic scattering tables of the workloads, and the pair term (oracle).

Single scattering (the container is exafs_gpu/scattering.py:49-106; the synthetic
form replaces analytic_transition_metal_tables, scattering.py:238-278):

    f_b(k)      = Z_b/26 * exp(-(k - 4)^2 / 18)          by scatterer b
    delta_ab(k) = -0.9 k + 1.2 + 0.05 Z_b                 (the scatterer's Z)
    lambda(k)   = 6 + 0.3 k                               (A)
    S0^2 = 0.9, sigma^2 = 0.006 A^2 for every pair.

Pair term (exafs_gpu/kernels.py:22-29, forward.py:285-288):

    s02 f_b / (k r^2) exp(-2 r/lambda - 2 k^2 sigma^2) sin(2 k r + delta_ab - 4/3 k^3 C3)

defined as 0 at k = 0 (the reference's own grid never reaches k = 0).

Multiple scattering: one table row per library entry (absorber species, class,
species on the class's slots), the form of chimera/mslibrary.py:304-313

    T_e(k) = amp s02 F_e / (k r^2) exp(-2 r/lambda - 2 k^2 s2) sin(2 k r + delta_e),
    s2 = c_ms sigma2_nn r / (2 r_nn),  r = half path length,

where the reference reads F_e and delta_e from FEFF10 output. The synthetic stand-in
used here: F_e = prod over the path's scattering events of f_{s_m}(k), and
delta_e = -0.9 k + 1.2 + 0.05 * mean(Z_{s_m}) over the same events (the absorber
counts as a scatterer when the path revisits it).

Pinned by: closed forms evaluated directly (tests compare an independent scalar
evaluation), the k = 0 convention, and the count-space identity in forward.py.
"""
from __future__ import annotations

import numpy as np

Z = {"Cr": 24, "Mn": 25, "Fe": 26, "Co": 27, "Ni": 28, "Cu": 29, "Al": 13, "Ti": 22,
     "V": 23, "Pd": 46, "Pt": 78, "Au": 79}


def k_grid(kmax=16.0, dk=0.05):
    return np.linspace(0.0, kmax, int(round(kmax / dk)) + 1)


def synthetic_ss(k, zs, s02=0.9):
    """(k, f (nsp, nk), delta (nsp, nsp, nk), lam (nk), s02)."""
    k = np.asarray(k, float)
    zs = np.asarray(zs, float)
    nsp = zs.size
    f = np.array([z / 26.0 * np.exp(-(k - 4.0) ** 2 / 18.0) for z in zs])
    delta = np.empty((nsp, nsp, k.size))
    for a in range(nsp):
        for b in range(nsp):
            delta[a, b] = -0.9 * k + 1.2 + 0.05 * zs[b]
    lam = 6.0 + 0.3 * k
    return k, f, delta, lam, float(s02)


def pair_term(k, r, fb, dab, lam, sig2, s02, c3=0.0):
    """Pair term on the k grid for scalar r (kernels.py:22-29); 0 at k = 0."""
    k = np.asarray(k, float)
    out = np.zeros_like(k)
    nz = k > 0
    kk = k[nz]
    out[nz] = (s02 * fb[nz] / (kk * r * r)
               * np.exp(-2.0 * r / lam[nz] - 2.0 * kk * kk * sig2)
               * np.sin(2.0 * kk * r + dab[nz] - (4.0 / 3.0) * kk ** 3 * c3))
    return out


def ms_entry_table(k, zs, f, lam, entries, reff, s02=0.9, sigma2_nn=0.006, r_nn=1.0,
                   c_ms=1.3, amp=1.0):
    """(nsp * nblk, nk) table of one occurrence of every entry, absorber-major
    (mslibrary.py:304-313 with the synthetic F_e, delta_e above).

    entries: list of (class_index, scat_seq) where scat_seq is the function
    mapping (absorber species, slot species tuple) -> species sequence of the
    scattering events; here passed pre-expanded as a list per absorber of
    (class_index, species sequence) in library order.
    """
    k = np.asarray(k, float)
    zs = np.asarray(zs, float)
    rows = []
    for a_rows in entries:
        for g, seq in a_rows:
            r = reff[g]
            F = np.ones_like(k)
            for s in seq:
                F = F * f[s]
            dl = -0.9 * k + 1.2 + 0.05 * np.mean([zs[s] for s in seq])
            s2 = c_ms * sigma2_nn * r / (2.0 * r_nn)
            row = np.zeros_like(k)
            nz = k > 0
            kk = k[nz]
            row[nz] = (amp * s02 * F[nz] / (kk * r * r)
                       * np.exp(-2.0 * r / lam[nz] - 2.0 * kk * kk * s2)
                       * np.sin(2.0 * kk * r + dl[nz]))
            rows.append(row)
    return np.array(rows)


ETOK = 0.2624682843   # k^2 [A^-2] per eV (2 m_e / hbar^2), chimera/model.py:94


def k_exp_axis(k_phys, dE0):
    """The experimental k axis for an E0 error dE0 (chimera/model.py:120-130, k_axis
    "constant"): k_exp = sqrt(max(k^2 - ETOK dE0, 1e-4))."""
    k = np.asarray(k_phys, float)
    return np.sqrt(np.maximum(k ** 2 - ETOK * dE0, 1e-4))


def feff_columns_to_path(cols, reff):
    """feffNNNN.dat columns -> (p, f, delta, lambda) on the internal-momentum axis
    (exafs_gpu/scattering.py:158-178): rows with k > 0.3; amp = mag * red,
    pha = unwrap(2*phc + phase); f = amp p / k, delta = pha - 2 (p - k) R_eff."""
    c = np.asarray(cols, float)
    c = c[c[:, 0] > 0.3]
    k, phc, mag, pha, red, lam, p = c.T
    return p, mag * red * p / k, np.unwrap(phc + pha) - 2.0 * (p - k) * reff, lam


def read_feff_columns(fn):
    """(nleg, deg, reff, columns) of a feffNNNN.dat (exafs_gpu/scattering.py:181-190)."""
    lines = open(fn).read().splitlines()
    g = [i for i, ln in enumerate(lines) if "nleg, deg, reff" in ln][0]
    h = [i for i, ln in enumerate(lines) if ln.strip().startswith("k ") and "real[2*phc]" in ln][0]
    nleg, deg, reff = [float(x) for x in lines[g].split()[:3]]
    cols = np.array([[float(x) for x in ln.split()] for ln in lines[h + 1:] if ln.strip()])
    return int(nleg), deg, reff, cols


def tabulated_table(k, F, D, L, reff, s02, sigma2_nn, r_nn, c_ms=1.3, amp=1.0, a_scale=1.0):
    """(nrows, nk) table of one occurrence of every library entry from tabulated F, D, L
    on the model grid (chimera/mslibrary.py:304-313); 0 at k = 0."""
    k = np.asarray(k, float)
    out = np.zeros((len(reff), k.size))
    nz = k > 0
    kk = k[nz][None, :]
    r = (np.asarray(reff, float) * a_scale)[:, None]
    s2 = c_ms * sigma2_nn * r / (2.0 * r_nn)
    out[:, nz] = (amp * s02 * F[:, nz] / (kk * r * r)
                  * np.exp(-2.0 * r / np.maximum(L[:, nz], 1e-6) - 2.0 * kk * kk * s2)
                  * np.sin(2.0 * kk * r + D[:, nz]))
    return out
