"""ShellModel: the pair-count form of the swap-only forward model, and Cramer-Rao bounds on
the Warren-Cowley parameters.

Drop-in for chimera/shellmodel.py (ShellModel :116-441, crlb :450-593 and the helpers
dr_from_free / sro_directions / dN_direction :45-114).  On the ideal lattice chi depends on
a configuration only through N[a, b, s], the ordered pairs of absorber a and scatterer b in
shell s:

    chi_a(k) = 1/N_a sum_{b,s} N[a,b,s] B[a,b,s](k) + chi_MS,a(k)

with the same calibration parameters as Spectrum (a0, dE0, window, sigma^2 scale and
per-shell factors, bond offsets dr, C3, the multiple-scattering amplitude and Debye-Waller
scale, and the extensions C4, sigma_pair, ms_amp_edge).  The basis rows B are evaluated on
the device (the forward model's term-row kernel, `ForwardModel.term_rows`), the
composition-averaged multiple scattering by `MSLibrary.chi_fast`, and chi(R) by the device
transform; what remains on the host is the count algebra (nsp^2 nshell numbers per
configuration) and the small Fisher matrix of `crlb`.

Differences from the reference: the model is built on a device ForwardModel (`fm`) in place
of a host ScatteringTables, and `target` takes arrays (|chi(R)| or complex chi(R)) only.
"""
from __future__ import annotations

import numpy as np

from .fourier import FourierTransform
from .multift import MultiEdgeFT
from .spectrum import as_kweights, k_exp_axis
from .workloads import fit_scales, r_factor

PARAM_NAMES = ("a0", "dE0", "kmin", "kmax", "kweight", "sigma_scale", "sigma_shell_scale",
               "dr", "c3", "ms_amp", "ms_c", "c4", "sigma_pair", "ms_amp_edge")


def _pair_weights(comp):
    comp = np.asarray(comp, float)
    iu = np.triu_indices(comp.size)
    return iu, np.array([comp[a] * comp[b] * (1.0 if a == b else 2.0) for a, b in zip(*iu)])


def dr_from_free(x_free, comp):
    """Symmetric (nsp, nsp) bond offsets from the nsp(nsp+1)/2 - 1 free (a <= b) entries;
    the last diagonal entry makes the composition-weighted mean sum_ab c_a c_b dr_ab zero
    (shellmodel.py:45-62), so a0 keeps the mean bond length."""
    iu, w = _pair_weights(comp)
    nsp = len(comp)
    x = np.zeros(w.size)
    x[:-1] = np.asarray(x_free, float)
    x[-1] = -(w[:-1] @ x[:-1]) / w[-1]
    dr = np.zeros((nsp, nsp))
    dr[iu] = x
    return dr + dr.T - np.diag(np.diag(dr))


def sro_directions(nsp):
    """The nsp(nsp-1)/2 independent short-range-order directions of one shell
    (shellmodel.py:86-95): pairs (a, b), a < b."""
    return [(a, b) for a in range(nsp) for b in range(a + 1, nsp)]


def dN_direction(nsp, nshell, s, a, b):
    """+1 on N[a,b,s] and N[b,a,s], -1 on N[a,a,s] and N[b,b,s]: every absorber keeps its
    coordination (shellmodel.py:98-104)."""
    d = np.zeros((nsp, nsp, nshell))
    d[a, b, s] += 1.0
    d[b, a, s] += 1.0
    d[a, a, s] -= 1.0
    d[b, b, s] -= 1.0
    return d


class ShellModel:
    """chi(k) and chi(R) from pair counts, with cheap parameter updates."""

    def __init__(self, fm, k_phys, shells_ref, shell_mult, sigma2_shell_ref, natoms, a0_ref,
                 ms_paths=None, s02=1.0, dk_win=1.0, pairs=None, composition=None,
                 elements=None, R=None):
        self.fm = fm
        self.k = np.asarray(k_phys, float)
        if self.k.size != fm.nk or not np.allclose(self.k, fm.k):
            raise ValueError("k_phys must be the forward model's k grid")
        self.nsp = int(fm.nsp)
        comp = np.ones(self.nsp) if composition is None else np.asarray(composition, float)
        self.composition = tuple(float(x) for x in comp / comp.sum())
        self.elements = None if elements is None else tuple(elements)
        self.shells_ref = np.asarray(shells_ref, float)
        self.nshell = int(self.shells_ref.size)
        self.z = np.asarray(shell_mult, float)
        self.sigma2_shell_ref = np.asarray(sigma2_shell_ref, float)
        self.natoms = int(natoms)
        self.a0_ref = float(a0_ref)
        self.ms_paths = ms_paths
        self.s02 = float(s02)
        self.dk_win = float(dk_win)
        self.pairs = pairs
        self.R_out = None if R is None else np.asarray(R, float)
        self.p = dict(a0=self.a0_ref, dE0=np.zeros(self.nsp), kmin=3.0, kmax=13.0,
                      kweight=(3,), sigma_scale=1.0, sigma_shell_scale=np.ones(self.nshell),
                      dr=np.zeros((self.nsp, self.nsp)), c3=np.zeros(self.nshell),
                      ms_amp=1.0, ms_c=1.3, c4=np.zeros(self.nshell),
                      sigma_pair=np.ones((self.nsp, self.nsp)), ms_amp_edge=np.ones(self.nsp))
        self._B = self._ms = self._ft = None
        self._B_key = self._ms_key = self._ft_key = None

    # ---------------------------------------------------------- construction
    @classmethod
    def from_spectrum(cls, S):
        """The pair-count model of a Spectrum at its current parameters
        (shellmodel.py:186-205); with an exact path library attached its random-alloy
        expectation stands for the multiple scattering."""
        m = cls(S.fm, S.k_phys, S.shells_ref, S.shell_mult, S.sigma2_shell_ref, S.cell.natoms,
                a0_ref=S.a0_ref, ms_paths=S.ms_paths, s02=S.s02, dk_win=S.dk_win,
                pairs=(S.i, S.j, S.shell_idx), composition=S.composition, elements=S.elements,
                R=S._R_out)
        m.set_params(a0=S.a0, dE0=S.dE0, kmin=S.kmin, kmax=S.kmax, kweight=S.kweights,
                     sigma_scale=S.sigma_scale, sigma_shell_scale=S.sigma_shell_scale,
                     dr=S.dr, c3=S.c3, ms_amp=S.ms_amp, ms_c=S.ms_c)
        return m

    def apply_to(self, S):
        """Push the current parameters into a Spectrum (shellmodel.py:207-216)."""
        p = self.p
        S.set_a0(p["a0"])
        S.set_sigma_scale(p["sigma_scale"], p["sigma_shell_scale"])
        S.set_dr(p["dr"])
        S.set_c3(p["c3"])
        S.set_ms(p["ms_amp"], p["ms_c"])
        S.set_window(p["kmin"], p["kmax"], p["kweight"], p["dE0"])
        return S

    def params(self):
        p = self.p
        return dict(a0=p["a0"], dE0=p["dE0"].tolist(), kmin=p["kmin"], kmax=p["kmax"],
                    kweight=list(p["kweight"]), sigma_scale=p["sigma_scale"],
                    sigma_shell_scale=p["sigma_shell_scale"].tolist(), dr=p["dr"].tolist(),
                    c3=p["c3"].tolist(), c4=p["c4"].tolist(), sigma_pair=p["sigma_pair"].tolist(),
                    ms_amp_edge=p["ms_amp_edge"].tolist(), ms_amp=p["ms_amp"], ms_c=p["ms_c"],
                    multiple_scattering=self.ms_paths is not None)

    def set_params(self, **kw):
        """shellmodel.py:233-262: unknown names raise KeyError, None leaves a value."""
        for key, v in kw.items():
            if key not in self.p:
                raise KeyError(f"unknown parameter {key!r}; have {PARAM_NAMES}")
            if v is None:
                continue
            if key == "dE0":
                v = np.asarray(v, float) * np.ones(self.nsp)
            elif key == "sigma_shell_scale":
                v = np.asarray(v, float).ravel()
                if v.size != self.nshell:
                    raise ValueError(f"need {self.nshell} per-shell factors")
            elif key in ("dr", "sigma_pair"):
                v = np.asarray(v, float).reshape(self.nsp, self.nsp)
            elif key in ("c3", "c4"):
                c = np.atleast_1d(np.asarray(v, float))
                v = np.zeros(self.nshell)
                v[:min(c.size, self.nshell)] = c[:self.nshell]
            elif key == "ms_amp_edge":
                v = np.asarray(v, float).ravel() * np.ones(self.nsp)
            elif key == "kweight":
                v = as_kweights(v)
            elif key in ("kmin", "kmax"):
                v = float(v) if np.isscalar(v) else np.asarray(v, float)
            else:
                v = float(v)
            self.p[key] = v
        return self

    @property
    def chan_edge(self):
        return np.repeat(np.arange(self.nsp), len(self.p["kweight"])).astype(np.int32)

    @property
    def nchan(self):
        return self.nsp * len(self.p["kweight"])

    # ------------------------------------------------------------ geometry
    def radii(self):
        return self.shells_ref * (self.p["a0"] / self.a0_ref)

    def sigma2_shell(self):
        return self.sigma2_shell_ref * self.p["sigma_scale"] * self.p["sigma_shell_scale"]

    def sigma2_nn(self):
        return float(self.sigma2_shell()[0])

    # -------------------------------------------------------------- pieces
    def basis(self):
        """B[a, b, s](k) (nsp, nsp, nshell, nk) on the device (shellmodel.py:290-345):
        r = r_s + dr[a,b] in the first shell, sigma^2 = sigma2_shell[s] (x sigma_pair[a,b] in
        the first shell), phase - 4/3 k^3 C3_s, amplitude x exp(2/3 k^4 C4_s)."""
        p = self.p
        key = (p["a0"], p["sigma_scale"], p["sigma_shell_scale"].tobytes(), p["dr"].tobytes(),
               p["c3"].tobytes(), p["c4"].tobytes(), p["sigma_pair"].tobytes())
        if self._B_key != key:
            nsp, ns = self.nsp, self.nshell
            a, b, s = (x.ravel() for x in np.meshgrid(np.arange(nsp), np.arange(nsp),
                                                        np.arange(ns), indexing="ij"))
            first = s == 0
            r = self.radii()[s] + np.where(first, p["dr"][a, b], 0.0)
            s2 = self.sigma2_shell()[s] * np.where(first, p["sigma_pair"][a, b], 1.0)
            ev = np.zeros((a.size, 3), np.int32)
            ev[:, 0] = b
            rows = self.fm.term_rows(r, s2, p["c3"][s], a, np.ones(a.size, np.int32), ev)
            rows *= np.exp((2.0 / 3.0) * self.k[None, :] ** 4 * p["c4"][s][:, None])
            self._B = rows.reshape(nsp, nsp, ns, self.k.size)
            self._B_key = key
        return self._B

    def chi_ms(self):
        """Random-alloy multiple scattering per absorbing atom (shellmodel.py:347-365),
        times one amplitude per absorber."""
        if self.ms_paths is None:
            return 0.0
        p = self.p
        r_nn = float(self.radii()[0])
        key = (r_nn, self.sigma2_nn(), p["ms_amp"], p["ms_c"], p["ms_amp_edge"].tobytes())
        if self._ms_key != key:
            self.ms_paths.set_params(sigma2_nn=self.sigma2_nn(), r_nn=r_nn, c_ms=p["ms_c"],
                                     amp=p["ms_amp"], a_scale=r_nn / self.shells_ref[0])
            self._ms = self.ms_paths.chi_fast(self.fm, self.composition) * p["ms_amp_edge"][:, None]
            self._ms_key = key
        return self._ms

    def fts(self):
        """One transform per (edge, k weight) channel, edge-major, against k_exp(edge)."""
        p = self.p
        key = (p["dE0"].tobytes(), tuple(np.atleast_1d(p["kmin"])), tuple(np.atleast_1d(p["kmax"])),
               p["kweight"])
        if self._ft_key != key:
            lo, hi = np.atleast_1d(p["kmin"]), np.atleast_1d(p["kmax"])
            fts = [FourierTransform(k_exp_axis(self.k, p["dE0"][e]), float(lo[e % lo.size]),
                                    float(hi[e % hi.size]), w, dk_win=self.dk_win, R=self.R_out)
                   for e in range(self.nsp) for w in p["kweight"]]
            self._ft = MultiEdgeFT(fts, chan_edge=self.chan_edge)
            self._ft_key = key
        return self._ft

    @property
    def R(self):
        return self.fts().R

    # ---------------------------------------------------------- pair counts
    def counts_of(self, species):
        """Ordered pair counts N[a, b, s] of a configuration."""
        if self.pairs is None:
            raise ValueError("ShellModel built without pair lists")
        i, j, sh = self.pairs
        sp = np.asarray(species, int)
        N = np.zeros((self.nsp, self.nsp, self.nshell))
        np.add.at(N, (sp[i], sp[j], sh), 1.0)
        return N

    def absorber_counts(self, N):
        return np.asarray(N).sum(axis=(1, 2)) / self.z.sum()

    def random_counts(self, composition=None, counts=None):
        """Random-alloy expectation N[a,b,s] = N_a z_s c_b (shellmodel.py:393-403)."""
        if counts is None:
            comp = np.asarray(self.composition if composition is None else composition, float)
            cnt = np.round(comp / comp.sum() * self.natoms).astype(int)
            cnt[-1] = self.natoms - cnt[:-1].sum()
        else:
            cnt = np.asarray(counts, int)
        c = cnt / cnt.sum()
        return cnt[:, None, None] * self.z[None, None, :] * c[None, :, None]

    def alpha(self, N, s=0):
        """Warren-Cowley matrix of shell s: 1 - P(b|a)/c_b."""
        cnt = self.absorber_counts(N)
        c = cnt / cnt.sum()
        return 1.0 - N[:, :, s] / np.maximum(cnt[:, None] * self.z[s] * c[None, :], 1e-30)

    # ------------------------------------------------------------- forward
    def chi_k(self, N):
        cnt = np.maximum(self.absorber_counts(N), 1e-30)
        return np.einsum("abs,absk->ak", np.asarray(N, float), self.basis()) / cnt[:, None] + self.chi_ms()

    def chi_R(self, N, idx=None):
        """Complex chi(R) (nchan, nR), or on the data grid points idx."""
        out = self.fts().to_R(self.chi_k(N)[None])[0]
        return out if idx is None else out[:, idx]

    def mag_R(self, N, idx=None):
        return np.abs(self.chi_R(N, idx))

    def target(self, data):
        return np.asarray(data)

    def observe(self, N, tgt, idx=None):
        return self.chi_R(N, idx) if np.iscomplexobj(tgt) else self.mag_R(N, idx)

    def r_factor(self, N, data, mask, idx, return_scales=False):
        """R-factor with one amplitude per edge profiled out (shellmodel.py:427-437)."""
        d = self.target(data)
        m = self.observe(N, d, idx)
        sc = fit_scales(m, d, mask, self.chan_edge)
        rf = r_factor(m, d, mask, sc)
        if not return_scales:
            return rf
        ce = self.chan_edge
        return rf, np.array([sc[ce == e][0] for e in range(self.nsp)])


PRIOR_SD = dict(alpha=1.0, a0=0.02, sigma_scale=0.5, ms_amp=0.5, ms_c=0.5, c3=2e-4, dE0=5.0,
                dr=0.05)
NUISANCE_DEFAULT = ("a0", "sigma_scale", "ms_amp", "dE0", "dr", "c3")
FD_STEP = dict(a0=0.002, sigma_scale=0.05, ms_amp=0.05, ms_c=0.05, c3=2e-5, dE0=0.5, dr=0.005)
LOG_STEP = ("sigma_scale", "ms_amp", "ms_c")          # multiplicative parameters


def _nuisance_columns(model, name, step, comp):
    """(label, value at +step, value at -step) of every finite-difference column of one
    nuisance parameter: a0 and C3 (first shell) additive, dE0 one column per edge, the
    bond offsets along each free direction of dr_from_free, the rest in log steps."""
    v0 = model.p[name]
    if name == "a0":
        return [(name, v0 + step, v0 - step)]
    if name == "c3":
        e = np.zeros_like(v0)
        e[0] = step
        return [(name, v0 + e, v0 - e)]
    if name == "dE0":
        return [(f"dE0_{q}", v0 + step * np.eye(model.nsp)[q], v0 - step * np.eye(model.nsp)[q])
                for q in range(model.nsp)]
    if name == "dr":
        nfree = model.nsp * (model.nsp + 1) // 2 - 1
        return [(f"dr_{q}", v0 + dr_from_free(step * np.eye(nfree)[q], comp),
                 v0 - dr_from_free(step * np.eye(nfree)[q], comp)) for q in range(nfree)]
    if name in LOG_STEP:
        return [(name, v0 * np.exp(step), v0 * np.exp(-step))]
    raise KeyError(f"no finite-difference rule for {name!r}")


def crlb(model, N, data, sigma, mask, idx, shells=None, alpha_shell=0,
         nuisance=NUISANCE_DEFAULT, profile_scale=True, threshold=0.3, steps=None,
         prior_sd=None, whiten=None):
    """Cramer-Rao bounds on the Warren-Cowley parameters of shell `alpha_shell`
    (shellmodel.py:450-593).

    The normalised residual r = (model - target)/sigma on the masked points (Re and Im for a
    complex target, one amplitude per edge profiled out) is linearised around the counts N
    by central differences: one column per short-range-order direction of every shell in
    `shells` (a step of about 0.02 in alpha units) and per nuisance degree of freedom.  A
    weak Gaussian prior (`prior_sd`: 1 in alpha units on the SRO directions, PRIOR_SD on
    the nuisances) regularises directions the data do not see.  The posterior covariance
    C = (J^T J + P)^-1 comes from a Cholesky factorisation; the alpha covariance is
    A^T C_sro A with A = d alpha / d direction.  `whiten` maps r to L^-1 r for a
    correlated noise model C_noise = L L^T.  Returns the reference's dictionary (plus the
    Jacobian)."""
    nsp, nshell = model.nsp, model.nshell
    shells = list(range(nshell)) if shells is None else list(shells)
    data = model.target(data)
    cplx = np.iscomplexobj(data)
    if data.shape[0] != model.nchan:
        raise AssertionError(f"target has {data.shape[0]} channels, model {model.nchan}")
    sig = np.broadcast_to(np.asarray(sigma, float).reshape(
        (-1, 1) if np.ndim(sigma) == 1 else np.shape(sigma)), data.shape)
    cnt = model.absorber_counts(N)
    comp = cnt / cnt.sum()
    iu = np.triu_indices(nsp)
    step = dict(FD_STEP, **(steps or {}))
    prior = dict(PRIOR_SD, **(prior_sd or {}))

    def resid(Nx):
        m = model.observe(Nx, data, idx)
        sc = fit_scales(m, data, mask, model.chan_edge) if profile_scale else np.ones(model.nchan)
        r = (m * sc[:, None] - data)[:, mask] / sig[:, mask]
        r = np.concatenate([r.real.ravel(), r.imag.ravel()]) if cplx else r.ravel()
        return r if whiten is None else whiten(r)

    cols, names, prec, A = [], [], [], []
    for s in shells:
        h = 0.02 * cnt.min() * model.z[s] * comp.min()
        a_s = model.alpha(N, s)
        for a, b in sro_directions(nsp):
            d = h * dN_direction(nsp, nshell, s, a, b)
            cols.append((resid(N + d) - resid(N - d)) / (2.0 * h))
            names.append(f"sro_s{s + 1}_{a}{b}")
            dalpha = (model.alpha(N + d, s) - a_s) / h
            A.append(dalpha[iu] if s == alpha_shell else np.zeros(iu[0].size))
            prec.append((np.abs(dalpha).max() / prior["alpha"]) ** 2)
    nsro = len(cols)
    for name in nuisance or ():
        v0 = model.p[name]
        v0 = v0.copy() if isinstance(v0, np.ndarray) else v0
        try:
            for label, hi, lo in _nuisance_columns(model, name, step[name], comp):
                model.set_params(**{name: hi})
                rp = resid(N)
                model.set_params(**{name: lo})
                rm = resid(N)
                cols.append((rp - rm) / (2.0 * step[name]))
                names.append(label)
                prec.append(1.0 / prior[name] ** 2)
        finally:
            model.set_params(**{name: v0})
    J = np.array(cols).T
    F = J.T @ J
    L = np.linalg.cholesky(F + np.diag(prec))
    Linv = np.linalg.solve(L, np.eye(L.shape[0]))
    cov = Linv.T @ Linv
    A = np.array(A)
    cov_alpha = A.T @ cov[:nsro, :nsro] @ A
    sd = np.sqrt(np.maximum(np.diag(cov_alpha), 0.0))
    ev = np.linalg.eigvalsh(F[:nsro, :nsro])[::-1]
    with np.errstate(divide="ignore"):
        eig_sd = 1.0 / np.sqrt(np.maximum(ev, 1e-300))
    csd = np.sqrt(np.maximum(np.diag(cov), 0.0))
    nuis_sd = {}
    for j in range(nsro, len(names)):
        key = names[j].split("_")[0]
        nuis_sd.setdefault(key, []).append(float(csd[j]))
    dr_entry_sd = None
    if "dr" in nuis_sd:
        q = [j for j, n in enumerate(names) if n.startswith("dr_")]
        nfree = len(q)
        D = np.array([dr_from_free(np.eye(nfree)[i], comp)[iu] for i in range(nfree)])
        dr_entry_sd = np.sqrt(np.maximum(np.einsum("ie,ij,je->e", D, cov[np.ix_(q, q)], D),
                                         0.0)).tolist()
    return dict(alpha_sd=sd.tolist(), determined=(sd < threshold).tolist(), threshold=threshold,
                prior_alpha_sd=prior["alpha"], shells=[int(s) + 1 for s in shells],
                nuisance=list(nuisance or ()), n_sro_directions=int(nsro),
                eig_sd=eig_sd.tolist(), columns=names, alpha_cov=cov_alpha.tolist(),
                nuisance_sd={k: (v[0] if len(v) == 1 else v) for k, v in nuis_sd.items()},
                dr_entry_sd=dr_entry_sd, jacobian=J)
