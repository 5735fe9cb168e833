"""Writer of FEFF-format path libraries for the tests (library.json + one feffNNNN.dat per
entry and absorber).  The columns encode a chosen f(p), delta(p), lambda(p) through the
FEFF convention: k measured from the Fermi level (p^2 = k^2 + p0^2), amp = mag * red,
pha = 2*phc + phase, f = amp p / k, delta = pha - 2 (p - k) R_eff."""
import json
import os

import numpy as np


def write_feffdat(fn, reff, deg, nleg, atoms, k, p, f, delta, lam, red=0.95):
    phc = 0.3 + 0.01 * k                      # any split of the phase between the columns
    pha = delta + 2.0 * (p - k) * reff - phc
    mag = f * k / (p * red)
    with open(fn, "w") as fh:
        fh.write(" synthetic path written by the tests\n")
        fh.write(" -----------------------------------------------------------------------\n")
        fh.write(f"   {nleg:d}  {deg:.3f}  {reff:.8f}  0.0  0.0 nleg, deg, reff, rnrmav(bohr), edge\n")
        fh.write("        x           y           z     pot at#  sym\n")
        for xyz, ip, sym in atoms:
            fh.write(f" {xyz[0]:11.6f} {xyz[1]:11.6f} {xyz[2]:11.6f}  {ip:d}  0 {sym}\n")
        fh.write("    k   real[2*phc]   mag[feff]  phase[feff] red factor   lambda     real[p]@#\n")
        for row in zip(k, phc, mag, pha, np.full_like(k, red), lam, p):
            fh.write(" " + " ".join(f"{v:.12e}" for v in row) + "\n")


def write_library(root, elements, classes, unit, d_nn, fdl, p0=0.6, kmax=20.0, dk=0.05):
    """classes: objects with .rep (integer offsets, units of a0/2), .slot_of, .nslot and
    .reff_units(); fdl(absorber, scatterer sequence, p) -> (f, delta, lambda)."""
    os.makedirs(root, exist_ok=True)
    nsp = len(elements)
    k = np.arange(0.0, kmax + 1e-9, dk)
    p = np.sqrt(k * k + p0 * p0)
    with open(os.path.join(root, "library.json"), "w") as fh:
        json.dump(dict(elements=list(elements), d_nn=d_nn, unit=unit, lattice="fcc",
                       classes=[dict(rep=np.asarray(c.rep).tolist()) for c in classes]), fh)
    for a, el in enumerate(elements):
        d = os.path.join(root, el)
        os.makedirs(d, exist_ok=True)
        n = 0
        for cl in classes:
            for code in range(nsp ** cl.nslot):
                sp = [(code // nsp ** q) % nsp for q in range(cl.nslot)]
                seq = [a if t < 0 else sp[t] for t in cl.slot_of]
                f, delta, lam = fdl(a, seq, p)
                reff = cl.reff_units() * unit
                atoms = [((0.0, 0.0, 0.0), 0, el)] + [
                    (tuple(np.asarray(v, float) * unit), 1, elements[s])
                    for v, s in zip(np.asarray(cl.rep), seq)]
                n += 1
                write_feffdat(os.path.join(d, f"feff{n:04d}.dat"), reff, 1.0, len(seq) + 1,
                              atoms, k, p, f, delta, lam)
    return root
