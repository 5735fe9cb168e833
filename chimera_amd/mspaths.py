"""Multiple-scattering path classes and the per-site incidence stencil (host setup).

Mirrors chimera/mslibrary.py: classes under the cubic group O_h and time reversal
(:51-62, :86-132), library entries class-major with slot 0 fastest (:190-198),
per-supercell entry counts (SupercellMS, :341-492).

The class set is found by a breadth-first walk over lattice offsets: every path
absorber -> s1 -> ... -> absorber with 3 or 4 legs, no zero-length leg and half path
length <= rmax_ms (the reference reads it from a FEFF10 paths.dat, :135-149).

The labelling of a class's directed images must be the reference's, because an
entry index is s_abs * nblk + cls_off + sum_q s_q nsp^q over the slot labels: the
canonical representative is the smallest image under O_h x reversal, slots are
numbered by first appearance along it, and images are generated forward over O_h,
then reversed over O_h, the first image found keeping its labels (:100-124).

For the device the directed instances are turned into a translation-invariant
stencil: images of one class with the same site -> slot map give the same entry for
every configuration and are merged with a weight (a path and its reverse); then for
every role a site can play (absorber, or slot q) one record holds the other sites as
offsets from it.  A swap visits only the records of the two swapped sites.
"""
from __future__ import annotations

import itertools
from collections import deque

import numpy as np

from .lattice import lattice_offsets


def _ops():
    out = []
    for perm in itertools.permutations(range(3)):
        for sg in itertools.product((1, -1), repeat=3):
            A = np.zeros((3, 3), np.int64)
            for r, c in enumerate(perm):
                A[r, c] = sg[r]
            out.append(A)
    return out


OPS = _ops()


def _canonical(seq):
    S0 = np.asarray(seq, np.int64)
    best = None
    for s in (S0, S0[::-1]):
        for A in OPS:
            t = tuple(map(tuple, s @ A.T))
            if best is None or t < best:
                best = t
    return best


class PathClass:
    def __init__(self, rep):
        self.rep = np.asarray(rep, np.int64)
        self.nleg = self.rep.shape[0] + 1
        slot_of, seen = [], {}
        for q in map(tuple, self.rep):
            if q == (0, 0, 0):
                slot_of.append(-1)
            else:
                seen.setdefault(q, len(seen))
                slot_of.append(seen[q])
        self.slot_of = np.array(slot_of, np.int64)
        self.nslot = len(seen)
        imgs = {}
        for rev in (False, True):
            s = self.rep[::-1] if rev else self.rep
            so = self.slot_of[::-1] if rev else self.slot_of
            for A in OPS:
                t = tuple(map(tuple, s @ A.T))
                if t not in imgs:
                    imgs[t] = so.copy()
        self.images = [(np.array(k, np.int64), v) for k, v in imgs.items()]
        self.deg = len(self.images)

    def reff_units(self):
        pts = np.vstack([np.zeros((1, 3)), self.rep, np.zeros((1, 3))])
        return 0.5 * float(np.linalg.norm(np.diff(pts, axis=0), axis=1).sum())


def find_classes(lattice, a0, rmax_ms, legs=(3, 4)):
    """All path classes with nleg in `legs` and half length <= rmax_ms (A)."""
    unit = a0 / 2.0
    lim = 2.0 * rmax_ms / unit + 1e-9                 # full path length budget (units)
    steps = lattice_offsets(lattice, lim / 2.0)       # one leg is at most half the path
    steplen = np.sqrt((steps ** 2).sum(1))
    found = set()
    nmax = max(legs)
    # state: (sites so far after the absorber, length so far)
    queue = deque([((), 0.0)])
    while queue:
        seq, L = queue.popleft()
        prev = np.array(seq[-1]) if seq else np.zeros(3, np.int64)
        nxt = prev[None, :] + steps
        back = np.sqrt((nxt ** 2).sum(1))
        ok = L + steplen + back <= lim
        for q, l in zip(nxt[ok], steplen[ok]):
            tq = tuple(int(v) for v in q)
            s2 = seq + (tq,)
            n_leg = len(s2) + 1
            if any(tq):  # a path may close only from a site other than the absorber
                if n_leg in legs:
                    found.add(_canonical(s2))
            if n_leg < nmax:
                queue.append((s2, L + l))
    keys = sorted(found, key=lambda k: (len(k), PathClass(k).reff_units(), k))
    return [PathClass(k) for k in keys]


LIB_FILE = "library.json"


class PathLibrary:
    """A multiple-scattering path library on disk (chimera/mslibrary.py:261-313):
    `library.json` (elements, nearest-neighbour distance d_nn, the class representatives
    in units of a0/2) and, per absorber element, one feffNNNN.dat per library entry in
    entry order (class-major, slot 0 fastest).  `prepare(k)` puts every entry's F, delta
    and lambda on the model grid; `params` gives the per-row inputs of the table at the
    current lattice and sigma^2."""

    def __init__(self, root, composition=None):
        import json
        import os
        from .scattering import read_feffdat
        with open(os.path.join(root, LIB_FILE)) as fh:
            js = json.load(fh)
        self.name = os.path.basename(os.path.normpath(root))
        self.elements = tuple(js["elements"])
        self.nsp = len(self.elements)
        self.d_nn = float(js["d_nn"])
        self.unit = float(js["unit"])
        self.lattice = js.get("lattice", "fcc")
        self.classes = [PathClass(c["rep"]) for c in js["classes"]]
        self.nblk = int(sum(self.nsp ** c.nslot for c in self.classes))
        comp = np.ones(self.nsp) if composition is None else np.asarray(composition, float)
        self.composition = comp / comp.sum()
        self.paths = {}
        for el in self.elements:
            d = os.path.join(root, el)
            files = sorted(f for f in os.listdir(d) if f.startswith("feff0") and f.endswith(".dat"))
            if len(files) != self.nblk:
                raise AssertionError(f"{el}: {len(files)} files for {self.nblk} entries")
            self.paths[el] = [read_feffdat(os.path.join(d, f)) for f in files]
        self._grid = None
        self.F = self.D = self.L = self.reff = None

    def summary(self):
        return [dict(absorber=a, n_classes=len(self.classes), n_entries=self.nblk,
                     total_degeneracy=int(sum(c.deg for c in self.classes)))
                for a in self.elements]

    def prepare(self, k_phys):
        """Interpolate every entry onto the model grid (absorber-major rows)."""
        p = np.asarray(k_phys, np.float64)
        recs = [d for a in self.elements for d in self.paths[a]]
        self._grid = p
        self.F = np.array([np.interp(p, d["p"], d["f"]) for d in recs])
        self.D = np.array([np.interp(p, d["p"], d["delta"]) for d in recs])
        self.L = np.array([np.interp(p, d["p"], d["lam"]) for d in recs])
        self.reff = np.array([d["reff"] for d in recs])
        return self

    def params(self, sigma2_nn, r_nn, c_ms=1.3, a_scale=1.0):
        """Per-row r and sigma^2 of the table (mslibrary.py:304-313): the file's half path
        length times the lattice scale, sigma^2 = c_ms sigma2_nn r / (2 r_nn)."""
        r = self.reff * float(a_scale)
        return r, c_ms * sigma2_nn * r / (2.0 * r_nn)


class MSLibrary:
    """Path classes, entries, device stencil and synthetic table rows of a supercell.

    Synthetic entry table (the reference reads FEFF10 output, mslibrary.py:294-313):
    F_e = product of f over the path's scattering events, delta_e = mean over the
    events of delta[absorber, scatterer], lambda the single-scattering one; row value
    amp s02 F_e/(k r^2) exp(-2r/lambda - 2k^2 c_ms sigma2_nn r/(2 r_nn)) sin(2kr + delta_e)
    with r the class's half path length.
    """

    def __init__(self, cell, nsp, rmax_ms, sigma2_nn=0.006, r_nn=None, c_ms=1.3, amp=1.0,
                 library=None, classes=None, fixed_table=None):
        """`library`: a PathLibrary (prepared on the forward model's k grid) whose
        classes and FEFF tables replace the enumerated classes and synthetic rows; the
        table's lattice scale is then cell.unit / library.unit.  `classes` (PathClass
        list) and `fixed_table` ((nsp * nblk, nk) rows) instead take both from a
        library made elsewhere (see from_protocol)."""
        self.cell = cell
        self.nsp = int(nsp)
        self.library = library
        self.fixed_table = None if fixed_table is None else np.ascontiguousarray(fixed_table, np.float64)
        if classes is not None:
            self.classes = list(classes)
        elif library is not None:
            if library.nsp != self.nsp:
                raise ValueError(f"the library holds {library.nsp} species, the cell {self.nsp}")
            if library.lattice != cell.lattice:
                raise ValueError(f"a {library.lattice} library on a {cell.lattice} cell")
            if getattr(library, "F", None) is None:
                raise ValueError("prepare() the path library on the model k grid first")
            self.classes = list(library.classes)
        else:
            self.classes = find_classes(cell.lattice, cell.a0, rmax_ms)
        self.unit = cell.unit
        self.nblk = int(sum(self.nsp ** c.nslot for c in self.classes))
        self.cls_off = np.cumsum([0] + [self.nsp ** c.nslot for c in self.classes])[:-1]
        for c in self.classes:
            if 2 * np.abs(c.rep).max() >= cell.box_units.min():
                raise ValueError("supercell too small for the paths")
        self.sigma2_nn = float(sigma2_nn)
        self.r_nn = float(r_nn) if r_nn is not None else float(
            np.sqrt((lattice_offsets(cell.lattice, 2.0) ** 2).sum(1)).min() * self.unit)
        self.d_nn = self.r_nn
        self.c_ms, self.amp = float(c_ms), float(amp)
        self.a_scale = 1.0 if library is None else self.unit / library.unit
        self._build_stencil()

    @classmethod
    def from_protocol(cls, lib_obj, cell):
        """A sampler-side library from an object of the reference's ms_library protocol
        (chimera/mslibrary.py:341-492, SupercellMS): its path classes (`lib.classes`, each
        with the representative `rep` in units of a0/2; the slot labelling, and so every
        entry index, follows from the representative) and its entry table `T`
        ((nsp * nblk, nk), absorber-major, one occurrence per entry, as set_params left
        it; a device array is copied to the host).  The per-step protocol calls
        (init / delta / commit / permute / full) are what this sampler does on the device
        in count space, so only the classes and the table are taken."""
        lib = getattr(lib_obj, "lib", None)
        if lib is None or not hasattr(lib, "classes"):
            raise ValueError("ms_library must be a chimera_amd MSLibrary or an object with "
                             "lib.classes and T (the reference's SupercellMS protocol)")
        T = getattr(lib_obj, "T", None)
        if T is None:
            raise ValueError("ms_library.T is not set: call its set_params(...) first")
        T = T.get() if hasattr(T, "get") else np.asarray(T)
        nsp = int(getattr(lib_obj, "nsp", getattr(lib, "nsp", 0)))
        classes = [PathClass(np.asarray(c.rep, np.int64)) for c in lib.classes]
        out = cls(cell, nsp, 0.0, classes=classes, fixed_table=T)
        if out.fixed_table.shape[0] != nsp * out.nblk:
            raise ValueError(f"T has {out.fixed_table.shape[0]} rows, the classes give "
                             f"{nsp} x {out.nblk} entries")
        if hasattr(lib_obj, "nblk") and int(lib_obj.nblk) != out.nblk:
            raise ValueError("the object's nblk does not match its classes")
        return out

    def table(self, fm):
        """(nsp * nblk, nk) term rows, one occurrence of every entry (absorber-major), on
        the device: synthetic rows from the forward model's tables, or the library's
        tabulated F, delta, lambda at the current parameters."""
        if self.fixed_table is not None:
            if self.fixed_table.shape[1] != fm.nk:
                raise ValueError("the library table is on another k grid")
            return self.fixed_table
        if self.library is not None:
            r, s2 = self.library.params(self.sigma2_nn, self.r_nn, self.c_ms, self.a_scale)
            lib = self.library
            if lib.F.shape[1] != fm.nk:
                raise ValueError("the library was prepared on another k grid")
            return fm.tab_rows(r, s2, lib.F, lib.D, lib.L, amp=self.amp)
        r, s2, ab, nev, ev = self.table_rows()
        return fm.term_rows(r, s2, np.zeros(r.size), ab, nev, ev, amp=self.amp)

    def set_params(self, sigma2_nn, r_nn, c_ms=1.3, amp=1.0, a_scale=1.0):
        """Table parameters at the current lattice (chimera/mslibrary.py:304-313, 410-412):
        nearest-neighbour sigma^2 and distance, Debye-Waller scale c_ms, amplitude, and
        the lattice scale applied to every path length."""
        self.sigma2_nn, self.r_nn = float(sigma2_nn), float(r_nn)
        self.c_ms, self.amp, self.a_scale = float(c_ms), float(amp), float(a_scale)
        return self

    def expected_counts(self, composition):
        """(nsp * nblk,) occurrences of every entry around one absorber of each species in
        a random alloy of `composition` (chimera/mslibrary.py:315-321): the class's
        directed image count times the probability of the slot species, the slots
        being distinct sites."""
        c = np.asarray(composition, np.float64)
        c = c / c.sum()
        blk = np.array([self.classes[g].deg * np.prod([c[s] for s in sp]) for g, sp in self.entries()])
        return np.tile(blk, self.nsp)

    def chi_fast(self, fm, composition):
        """Composition-averaged chi_MS (nsp, nk) per absorbing atom
        (chimera/mslibrary.py:329-333, the drop-in for multiscat.py:133-145): the entry
        rows evaluated on the device at the current parameters, weighted by the
        random-alloy occurrences."""
        T = self.table(fm)
        w = self.expected_counts(composition)
        nb = self.nblk
        return np.stack([w[a * nb:(a + 1) * nb] @ T[a * nb:(a + 1) * nb] for a in range(self.nsp)])

    # ------------------------------------------------------------- entries
    def entries(self):
        """[(class index, slot species)] in library order."""
        out = []
        for g, cl in enumerate(self.classes):
            for code in range(self.nsp ** cl.nslot):
                out.append((g, tuple((code // self.nsp ** q) % self.nsp for q in range(cl.nslot))))
        return out

    def table_rows(self):
        """Per-row parameters of the entry table (absorber-major): r, s2, absorber,
        nev, ev (scatterer species of each event)."""
        ent = self.entries()
        n = self.nsp * len(ent)
        r = np.empty(n)
        s2 = np.empty(n)
        ab = np.empty(n, np.int32)
        nev = np.empty(n, np.int32)
        ev = np.zeros((n, 3), np.int32)
        q = 0
        for a in range(self.nsp):
            for g, sp in ent:
                cl = self.classes[g]
                rr = cl.reff_units() * self.unit * self.a_scale
                seq = [a if t < 0 else sp[t] for t in cl.slot_of]
                r[q] = rr
                s2[q] = self.c_ms * self.sigma2_nn * rr / (2.0 * self.r_nn)
                ab[q] = a
                nev[q] = len(seq)
                ev[q, :len(seq)] = seq
                q += 1
        return r, s2, ab, nev, ev

    # ------------------------------------------------------------- stencil
    def _build_stencil(self):
        recs = []       # (abs offset, [slot offsets], nslot, w, cls_off), absorber role
        for g, cl in enumerate(self.classes):
            merged = {}
            for img, so in cl.images:
                M = tuple(tuple(int(v) for v in img[int(np.argmax(so == q))])
                          for q in range(cl.nslot))
                merged[M] = merged.get(M, 0) + 1
            for M, w in merged.items():
                recs.append((M, w, int(self.cls_off[g])))
        self.nabsrec = len(recs)
        stencil = []
        for M, w, off in recs:
            stencil.append(self._rec((0, 0, 0), M, w, off))
        for M, w, off in recs:
            for q in range(len(M)):
                o = np.array(M[q])
                stencil.append(self._rec(tuple(-o), tuple(tuple(np.array(m) - o) for m in M),
                                         w, off))
        self.stencil = np.ascontiguousarray(np.array(stencil, np.int64).astype(np.int8).reshape(-1, 16))
        self.nstencil = self.stencil.shape[0]
        self.directed_per_absorber = int(sum(c.deg for c in self.classes))
        # a path and its reverse are one term (chimera/mslibrary.py:29-31): directed images
        # pair up except the palindromic ones
        pal = sum(int(np.array_equal(img, img[::-1])) for c in self.classes for img, _ in c.images)
        self.merged_per_absorber = (self.directed_per_absorber + pal) // 2

    @staticmethod
    def _rec(abs_off, slots, w, cls_off):
        r = np.zeros(16, np.int64)
        r[0:3] = abs_off
        for q, s in enumerate(slots):
            r[3 + 3 * q:6 + 3 * q] = s
        r[12] = len(slots)
        r[13] = w
        r[14] = cls_off & 0xFF
        r[15] = cls_off >> 8
        if np.abs(r[:12]).max() > 127 or cls_off >= 1 << 15 or w > 127:
            raise ValueError("stencil record does not fit its int8 fields")
        return r
