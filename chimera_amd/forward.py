"""ForwardModel: configuration -> chi(k) on the gfx950 path-sum kernel.

Drop-in for exafs_gpu/forward.py:50-188 (ForwardModel).  Arguments, shapes, the
unnormalised/normalised convention and the exceptions are the reference's; the arrays
are NumPy on both sides (the library owns the device copies; `d_shift` and `d_sig2`
are properties whose setters re-upload, which is what the reference's callers rely on
when they assign them, chimera/model.py:263, 278).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from ._lib import check, f64, i32

KB = 8.617333262e-5   # Boltzmann constant, eV/K (CODATA 2018)


def thermal_sigma2(Phi, cell, i_idx, j_idx, shift, temperature=300.0):
    """Per-pair mean-square relative displacement along the bond from a harmonic
    force-constant matrix Phi (3N, 3N) (host helper, exafs_gpu/forward.py:31-47).

    The displacement covariance is C = kB T Phi^+ (pseudo-inverse over the modes with
    eigenvalue above 1e-6 of the largest, which drops the three translations);
    sigma^2_ij = e^T (C_ii + C_jj - C_ij - C_ji) e with e the unit bond vector."""
    Phi = np.asarray(Phi, float)
    w, V = np.linalg.eigh(Phi)
    keep = w > 1e-6 * max(1.0, float(w.max()))
    Vk = V[:, keep]
    i_idx, j_idx = np.asarray(i_idx), np.asarray(j_idx)
    d = cell.ideal[j_idx] + np.asarray(shift, float) - cell.ideal[i_idx]
    e = d / np.linalg.norm(d, axis=1)[:, None]
    # project each bond direction onto the modes at both ends: u_p = e . (V_i - V_j)
    Vr = Vk.reshape(-1, 3, Vk.shape[1])                    # (N, 3, modes)
    u = np.einsum("pa,pam->pm", e, Vr[i_idx] - Vr[j_idx])  # (npairs, modes)
    return (u * u * (KB * temperature / w[keep])[None, :]).sum(1)


class ForwardModel:
    def __init__(self, cell, tables, i_idx, j_idx, shift, sigma2=None, shell=None, device=0):
        self.cell = cell
        self.natoms = int(cell.natoms)
        self.npairs = int(np.asarray(i_idx).size)
        k, f, delta, lam = tables.as_float64()
        self.nk = int(k.size)
        self.nsp = int(f.shape[0])
        if self.nsp < 2:
            raise ValueError("need at least two species")
        self.s02 = float(tables.s02)
        if shell is None:
            shell = np.zeros(self.npairs, np.int32)
        shell = i32(shell)
        assert shell.shape == (self.npairs,), "shell must be one index per pair"
        self.nshell = int(shell.max()) + 1
        self.k, self.f, self.delta, self.lam = k, f, delta, lam
        self.i_idx, self.j_idx, self.shell = i32(i_idx), i32(j_idx), shell
        self._shift = f64(shift).reshape(self.npairs, 3)
        self._sig2 = f64(np.zeros(self.npairs) if sigma2 is None else sigma2).reshape(self.npairs)
        self.dr = np.zeros((self.nshell, self.nsp, self.nsp))
        self.c3 = np.zeros(self.nshell)
        self.s2types = np.zeros((self.nshell, self.nsp, self.nsp))
        self.ctx = _lib.Context(device)
        lib = _lib.load()
        check(lib.chm_fm_set_tables(self.ctx.h, self.nsp, self.nk, k, f,
                                    f64(delta.reshape(self.nsp * self.nsp, -1)), lam, self.s02))
        check(lib.chm_fm_set_topology(self.ctx.h, self.natoms, self.npairs, self.i_idx,
                                      self.j_idx, self._shift, self.shell, self._sig2))

    # ------------------------------------------------------- device attributes
    @property
    def d_shift(self):
        return self._shift

    @d_shift.setter
    def d_shift(self, v):
        self._shift = f64(v).reshape(self.npairs, 3)
        check(_lib.load().chm_fm_set_shift(self.ctx.h, self._shift))

    @property
    def d_sig2(self):
        return self._sig2

    @d_sig2.setter
    def d_sig2(self, v):
        self._sig2 = f64(v).reshape(self.npairs)
        check(_lib.load().chm_fm_set_sig2(self.ctx.h, self._sig2))

    # The reference's other device attributes (forward.py:70-92), which its callers read:
    # NumPy copies of what the library holds (shapes and flattening as the reference's).
    @property
    def d_i(self):
        return self.i_idx

    @property
    def d_j(self):
        return self.j_idx

    @property
    def d_shell(self):
        return self.shell

    @property
    def d_k(self):
        return self.k

    @property
    def d_f(self):
        return self.f

    @property
    def d_delta(self):
        return np.ascontiguousarray(self.delta.reshape(self.nsp * self.nsp, -1))

    @property
    def d_lam(self):
        return self.lam

    @property
    def d_dr(self):
        return self.dr.reshape(-1)

    @property
    def d_c3(self):
        return self.c3

    @property
    def d_s2tab(self):
        return self.s2types.reshape(-1)

    def _upload_cumulants(self):
        check(_lib.load().chm_fm_set_cumulants(self.ctx.h, self.nshell, f64(self.dr.reshape(-1)),
                                               f64(self.c3), f64(self.s2types.reshape(-1))))

    # ------------------------------------------------------ cumulant physics
    def set_dr(self, dr):
        """forward.py:104-116."""
        dr = np.asarray(dr, np.float64)
        full = np.zeros((self.nshell, self.nsp, self.nsp))
        if dr.ndim == 2:
            full[0] = dr
        else:
            assert dr.shape == full.shape, f"dr must be {full.shape} or (nsp, nsp)"
            full[:] = dr
        self.dr = full
        self._upload_cumulants()
        return self

    def set_sig2_types(self, s2):
        """forward.py:118-132."""
        s2 = np.asarray(s2, np.float64)
        full = np.zeros((self.nshell, self.nsp, self.nsp))
        if s2.ndim == 2:
            full[0] = s2
        else:
            n = min(self.nshell, s2.shape[0])
            full[:n] = s2[:n]
        self.s2types = full
        self._upload_cumulants()
        return self

    def set_c3(self, c3):
        """forward.py:134-145."""
        c3 = np.atleast_1d(np.asarray(c3, np.float64))
        full = np.zeros(self.nshell)
        if c3.size == 1:
            full[0] = c3[0]
        else:
            assert c3.shape == (self.nshell,), f"c3 must have {self.nshell} entries"
            full[:] = c3
        self.c3 = full
        self._upload_cumulants()
        return self

    # ---------------------------------------------------------------- helpers
    def distances(self, pos):
        """pos (nconf, natoms, 3) -> (r (nconf, npairs), evec (nconf, npairs, 3))."""
        pos = f64(pos).reshape(-1, self.natoms, 3)
        nconf = pos.shape[0]
        r = np.empty((nconf, self.npairs))
        ev = np.empty((nconf, self.npairs, 3))
        check(_lib.load().chm_fm_distances(self.ctx.h, nconf, pos, r, ev.ctypes.data))
        return r, ev

    def ptypes(self, species):
        species = i32(species).reshape(-1, self.natoms)
        pt = np.empty((species.shape[0], self.npairs), np.int32)
        check(_lib.load().chm_fm_ptypes(self.ctx.h, species.shape[0], species, pt))
        return pt

    # ---------------------------------------------------------------- forward
    def chi(self, pos, species, r=None, ptype=None, normalise=True, precision="f64"):
        """chi (nconf, nsp, nk) indexed by absorber species (forward.py:169-188).

        `precision` "f64" (default) or "f32" (FP32 terms, FP64 phase reduction and
        accumulation).  `r`/`ptype` are accepted for signature parity; the kernel
        recomputes both from pos/species on the device."""
        pos = f64(pos).reshape(-1, self.natoms, 3)
        species = i32(species).reshape(-1, self.natoms)
        nconf = pos.shape[0]
        if species.shape[0] != nconf:
            raise ValueError("pos and species disagree on nconf")
        prec = {"f64": _lib.PREC_F64, "f32": _lib.PREC_F32}[precision]
        chi = np.empty((nconf, self.nsp, self.nk))
        check(_lib.load().chm_fm_chi(self.ctx.h, nconf, pos, species, chi, prec))
        if normalise:
            cnt = np.stack([(species == s).sum(axis=1) for s in range(self.nsp)], axis=1)
            chi /= np.maximum(cnt, 1.0)[:, :, None]
        return chi

    def chi_reference(self, pos, species, sigma2=None):
        """Normalised chi (nconf, nsp, nk) with an optional per-pair sigma^2 override
        (forward.py:265-293, the reference's validation form: pair terms by absorber,
        divided by the absorber's atom count).  Evaluated by the device FP64 path sum,
        the kernel the tests validate against the NumPy oracle; the override is uploaded
        for the call and the model's own sigma^2 restored after it."""
        if sigma2 is None:
            return self.chi(pos, species, normalise=True)
        keep = self._sig2.copy()
        self.d_sig2 = np.broadcast_to(np.asarray(sigma2, np.float64), (self.npairs,))
        try:
            return self.chi(pos, species, normalise=True)
        finally:
            self.d_sig2 = keep

    def tab_rows(self, r, s2, F, D, L, amp=1.0):
        """Device evaluation of term rows from tabulated F, D, L (nrows, nk) on this
        model's k grid (include/chimera_hip.h, chm_fm_tab_rows)."""
        n = int(np.asarray(r).size)
        shp = (n, self.nk)
        out = np.empty(shp)
        check(_lib.load().chm_fm_tab_rows(self.ctx.h, n, f64(r), f64(s2), f64(F).reshape(shp),
                                          f64(D).reshape(shp), f64(L).reshape(shp), float(amp), out))
        return out

    # ---------------------------------------------------------------- adjoint
    def grad_pos(self, pos, species, gchi, r=None, evec=None, ptype=None, normalise=True):
        """dL/dpos (nconf, natoms, 3) given dL/dchi (nconf, nsp, nk) (forward.py:209-234):
        with `normalise`, gchi is dL/dchi of the normalised chi and is divided by the
        species counts first.  `r`/`evec`/`ptype` are accepted for signature parity; the
        device recomputes them."""
        pos = f64(pos).reshape(-1, self.natoms, 3)
        species = i32(species).reshape(-1, self.natoms)
        nconf = pos.shape[0]
        if species.shape[0] != nconf:
            raise ValueError("pos and species disagree on nconf")
        g = np.asarray(gchi, np.float64).reshape(nconf, self.nsp, self.nk)
        if normalise:
            cnt = np.stack([(species == s).sum(axis=1) for s in range(self.nsp)], axis=1)
            g = g / np.maximum(cnt, 1.0)[:, :, None]
        out = np.empty((nconf, self.natoms, 3))
        check(_lib.load().chm_fm_grad_pos(self.ctx.h, nconf, pos, species, f64(g), out))
        return out

    def chi_soft(self, pos, w, r=None, normalise=True):
        """Relaxed-occupancy chi (nconf, nsp, nk), w (nconf, natoms, nsp) on the simplex
        (forward.py:190-206); normalised by max(sum of w over sites, 1)."""
        pos = f64(pos).reshape(-1, self.natoms, 3)
        nconf = pos.shape[0]
        w = f64(w).reshape(nconf, self.natoms, self.nsp)
        chi = np.empty((nconf, self.nsp, self.nk))
        check(_lib.load().chm_fm_chi_soft(self.ctx.h, nconf, pos, w, chi))
        if normalise:
            chi /= np.maximum(w.sum(axis=1), 1.0)[:, :, None]
        return chi

    def grad_soft(self, pos, w, gchi, r=None, evec=None, normalise=True):
        """(dL/dpos, dL/dw) for the relaxed-occupancy model (forward.py:236-262), with the
        derivative of the normalisation by max(sum w, 1) applied as the reference does."""
        pos = f64(pos).reshape(-1, self.natoms, 3)
        nconf = pos.shape[0]
        w = f64(w).reshape(nconf, self.natoms, self.nsp)
        gchi = np.asarray(gchi, np.float64).reshape(nconf, self.nsp, self.nk)
        norm = np.maximum(w.sum(axis=1), 1.0)
        g = gchi / norm[:, :, None] if normalise else gchi
        gpos = np.empty((nconf, self.natoms, 3))
        gw = np.empty((nconf, self.natoms, self.nsp))
        check(_lib.load().chm_fm_grad_soft(self.ctx.h, nconf, pos, w, f64(g), gpos, gw))
        if normalise:
            chi_un = self.chi_soft(pos, w, normalise=False)
            gw -= ((gchi * chi_un).sum(axis=2) / norm ** 2)[:, None, :]
        return gpos, gw

    def profile(self, what, nconf, iters=20):
        """Average device ms of a forward-model launch sequence on the inputs of the
        last call of that kind (include/chimera_hip.h, chm_fm_profile): "chi" / "chi_f32"
        (path sum with distances and types), "adjoint" (dL/dr of grad_pos), "soft"
        (chi_soft), "soft_adjoint" (per-pair part of grad_soft)."""
        code = {"chi": 0, "chi_f32": 1, "adjoint": 2, "soft": 3, "soft_adjoint": 4}[what]
        ms = np.zeros(1)
        check(_lib.load().chm_fm_profile(self.ctx.h, code, int(nconf), int(iters), ms))
        return float(ms[0])

    def term_rows(self, r, s2, c3, absorber, nev, ev, amp=1.0):
        """Device evaluation of term-table rows (see include/chimera_hip.h)."""
        n = int(np.asarray(r).size)
        out = np.empty((n, self.nk))
        check(_lib.load().chm_fm_term_rows(self.ctx.h, n, f64(r), f64(s2), f64(c3), i32(absorber),
                                           i32(nev), i32(np.asarray(ev).reshape(n, 3)),
                                           float(amp), out))
        return out

    def ss_basis(self, radii, sig2_shell):
        """B[a, b, s](k) (nsp, nsp, nshell, nk): the term of one (absorber a, scatterer
        b, shell s) pair on the ideal lattice (chimera/shellmodel.py:290-336), with this
        model's dr / c3 / sigma^2-by-type tables."""
        nsp, ns = self.nsp, self.nshell
        a, b, s = np.meshgrid(np.arange(nsp), np.arange(nsp), np.arange(ns), indexing="ij")
        a, b, s = a.ravel(), b.ravel(), s.ravel()
        r = np.asarray(radii)[s] + self.dr[s, a, b]
        s2 = np.asarray(sig2_shell)[s] + self.s2types[s, a, b]
        ev = np.zeros((a.size, 3), np.int32)
        ev[:, 0] = b
        rows = self.term_rows(r, s2, self.c3[s], a, np.ones(a.size, np.int32), ev)
        return rows.reshape(nsp, nsp, ns, self.nk)
