"""Multiple-scattering path classes, library entries and supercell entry counts (oracle).

Follows chimera/mslibrary.py:
  * cubic group O_h and canonical form under O_h x reversal      :51-62, 86-97
  * PathClass (slots, directed images, degeneracy, half length)  :100-132
  * entry_list order (class-major, slot 0 fastest)                :190-198
  * cls_off                                                       :275
  * SupercellMS instances, incidence and entry counts             :350-407, 420-451
  * swap delta with paths through both sites counted once         :466-486

The reference reads the set of classes from a FEFF10 paths.dat (:135-149). Here the
classes are enumerated directly: every lattice path absorber -> s1 -> ... -> absorber
with 3 or 4 legs, no zero-length leg, and half path length <= rmax_ms. That rule
reproduces the reference's own figures for fcc (25 classes to 5.3 A, 327 ternary
entries per absorber: mslibrary.py:11,16, README.md:267), which the tests check.
Sites are integer points in units of a0/2 (fcc and bcc).
"""
from __future__ import annotations

import itertools

import numpy as np


def cubic_ops():
    out = []
    for perm in itertools.permutations(range(3)):
        for sg in itertools.product((1, -1), repeat=3):
            A = np.zeros((3, 3), int)
            for r, c in enumerate(perm):
                A[r, c] = sg[r]
            out.append(A)
    return out


OPS = cubic_ops()


def canonical(seq):
    """Smallest image of a site sequence under O_h and reversal (:86-97)."""
    best = None
    S0 = np.asarray(seq, int)
    for s in (S0, S0[::-1]):
        for A in OPS:
            t = tuple(map(tuple, s @ A.T))
            if best is None or t < best:
                best = t
    return best


class PathClass:
    """One geometric class (:100-132)."""

    def __init__(self, rep):
        self.rep = np.asarray(rep, int)
        self.nleg = self.rep.shape[0] + 1
        slot_of, seen = [], {}
        for q in map(tuple, self.rep):
            if q == (0, 0, 0):
                slot_of.append(-1)
            else:
                seen.setdefault(q, len(seen))
                slot_of.append(seen[q])
        self.slot_of = np.array(slot_of, int)
        self.nslot = len(seen)
        imgs = {}
        for rev in (False, True):
            s = self.rep[::-1] if rev else self.rep
            so = self.slot_of[::-1] if rev else self.slot_of
            for A in OPS:
                t = tuple(map(tuple, s @ A.T))
                if t not in imgs:
                    imgs[t] = so.copy()
        self.images = [(np.array(k, int), v) for k, v in imgs.items()]
        self.deg = len(self.images)

    def reff_units(self):
        pts = np.vstack([np.zeros((1, 3)), self.rep, np.zeros((1, 3))])
        return 0.5 * float(np.linalg.norm(np.diff(pts, axis=0), axis=1).sum())


def lattice_points(lattice, radius):
    """Integer points (units a0/2) of the lattice within `radius` (same units)."""
    R = int(np.ceil(radius))
    pts = np.array(list(itertools.product(range(-R, R + 1), repeat=3)))
    if lattice == "fcc":
        keep = pts.sum(1) % 2 == 0
    elif lattice == "bcc":
        keep = (pts % 2 == 0).all(1) | (pts % 2 == 1).all(1)
    elif lattice == "sc":
        keep = (pts % 2 == 0).all(1)
    else:
        raise ValueError(lattice)
    pts = pts[keep]
    return pts[np.linalg.norm(pts, axis=1) <= radius + 1e-9]


def enumerate_classes(lattice, a0, rmax_ms, legs=(3, 4)):
    """Classes of all 3-/4-leg lattice paths with half length <= rmax_ms (A),
    ordered by (number of legs, half length, canonical form) as :148."""
    unit = a0 / 2.0
    lim = 2.0 * rmax_ms / unit
    pts = lattice_points(lattice, lim / 2.0 + 1e-9)
    found = set()
    for n in legs:
        def rec(seq, L):
            prev = np.array(seq[-1]) if seq else np.zeros(3, int)
            if len(seq) == n - 1:
                if 0.5 * (L + np.linalg.norm(prev)) <= rmax_ms / unit + 1e-9:
                    found.add(canonical(seq))
                return
            for q in pts:
                if np.array_equal(q, prev):
                    continue
                if len(seq) == n - 2 and not q.any():
                    continue
                leg = np.linalg.norm(q - prev)
                if 0.5 * (L + leg + np.linalg.norm(q)) > rmax_ms / unit + 1e-9:
                    continue
                rec(seq + [tuple(int(x) for x in q)], L + leg)
        rec([], 0.0)
    keys = sorted(found, key=lambda k: (len(k), PathClass(k).reff_units(), k))
    return [PathClass(k) for k in keys]


def entry_list(classes, nsp):
    """[(class index, slot species tuple)] in library order (:190-198)."""
    out = []
    for g, cl in enumerate(classes):
        for code in range(nsp ** cl.nslot):
            out.append((g, tuple((code // nsp ** q) % nsp for q in range(cl.nslot))))
    return out


def expected_counts(classes, nsp, composition):
    """(nsp * nblk,) mean occurrences of every entry around one absorber in a random
    alloy (chimera/mslibrary.py:315-321): the class's directed image count times the
    product of the slot species' concentrations (slots are distinct sites)."""
    c = np.asarray(composition, float)
    c = c / c.sum()
    blk = np.array([classes[g].deg * np.prod([c[s] for s in sp]) for g, sp in entry_list(classes, nsp)])
    return np.tile(blk, nsp)


def class_offsets(classes, nsp):
    return np.cumsum([0] + [nsp ** c.nslot for c in classes])[:-1].astype(np.int64)


def scattering_sequences(classes, nsp):
    """Per absorber species a: [(class, species of each scattering event)] in library
    order, the input of tables.ms_entry_table."""
    ent = entry_list(classes, nsp)
    out = []
    for a in range(nsp):
        rows = []
        for g, sp in ent:
            so = classes[g].slot_of
            rows.append((g, [a if q < 0 else sp[q] for q in so]))
        out.append(rows)
    return out


class SupercellPaths:
    """Every directed path instance of every class on a periodic cubic supercell,
    and the per-site incidence (SupercellMS.__init__, :350-407)."""

    def __init__(self, classes, nsp, ideal, a0, ncell):
        self.classes = classes
        self.nsp = nsp
        self.nblk = int(sum(nsp ** c.nslot for c in classes))
        self.nent = nsp * self.nblk
        self.cls_off = class_offsets(classes, nsp)
        unit = a0 / 2.0
        X = np.asarray(ideal, float) / unit
        Xi = np.rint(X).astype(int)
        assert np.allclose(X, Xi, atol=1e-6)
        box = 2 * np.asarray(ncell, int)
        Xi %= box
        look = -np.ones(tuple(box), int)
        look[Xi[:, 0], Xi[:, 1], Xi[:, 2]] = np.arange(len(Xi))
        n = self.natoms = len(Xi)
        P_abs, P_cls, P_slot = [], [], []
        for g, cl in enumerate(classes):
            assert 2 * np.abs(cl.rep).max() < box.min(), "supercell too small for the paths"
            for img, so in cl.images:
                p = (Xi[:, None, :] + img[None]) % box
                sites = look[p[..., 0], p[..., 1], p[..., 2]]
                assert (sites >= 0).all()
                slot = -np.ones((n, 3), int)
                for q in range(cl.nslot):
                    slot[:, q] = sites[:, int(np.argmax(so == q))]
                P_abs.append(np.arange(n))
                P_cls.append(np.full(n, g))
                P_slot.append(slot)
        self.P_abs = np.concatenate(P_abs)
        self.P_cls = np.concatenate(P_cls)
        self.P_slot = np.concatenate(P_slot)
        self.npath = self.P_abs.size
        site = np.concatenate([self.P_abs] + [self.P_slot[:, q] for q in range(3)])
        pid = np.tile(np.arange(self.npath), 4)
        keep = site >= 0
        site, pid = site[keep], pid[keep]
        order = np.argsort(site, kind="stable")
        self.inc_site, self.inc_pid = site[order], pid[order]
        self.inc_start = np.r_[0, np.cumsum(np.bincount(self.inc_site, minlength=n))]

    def entries(self, species, pids=None, swap=None):
        """Entry index of each path (:420-434); `swap` = (a, b, new_a, new_b)."""
        pids = np.arange(self.npath) if pids is None else np.asarray(pids)
        sp = np.asarray(species, np.int64)
        ab = self.P_abs[pids]
        sl = self.P_slot[pids]
        s_ab = sp[ab]
        s_sl = sp[np.maximum(sl, 0)]
        if swap is not None:
            a, b, na, nb = swap
            s_ab = np.where(ab == a, na, np.where(ab == b, nb, s_ab))
            s_sl = np.where(sl == a, na, np.where(sl == b, nb, s_sl))
        pw = np.array([1, self.nsp, self.nsp ** 2])
        return (s_ab * self.nblk + self.cls_off[self.P_cls[pids]]
                + (s_sl * pw * (sl >= 0)).sum(-1))

    def counts(self, species):
        """(nent,) entry counts of one configuration (:436-451)."""
        return np.bincount(self.entries(species), minlength=self.nent).astype(np.int64)

    def incident(self, site):
        return self.inc_pid[self.inc_start[site]:self.inc_start[site + 1]]

    def delta_counts(self, species, a, b, na, nb):
        """Change of the entry counts for the swap a->na, b->nb, paths through
        both sites counted once (:466-486)."""
        pa = self.incident(a)
        pb = self.incident(b)
        if a != b:
            has_a = (self.P_abs[pb] == a) | (self.P_slot[pb] == a).any(-1)
            pb = pb[~has_a]
        else:
            pb = pb[:0]
        pid = np.concatenate([pa, pb])
        e_old = self.entries(species, pid)
        e_new = self.entries(species, pid, (a, b, na, nb))
        return (np.bincount(e_new, minlength=self.nent)
                - np.bincount(e_old, minlength=self.nent)).astype(np.int64)
