"""Build the oracle's restatement of a workload from a built problem's host arrays
(cell, neighbour list, tables, window, target, sigma): the checker gets the same
inputs, and recomputes every derived quantity with its own algorithms."""
from __future__ import annotations

from . import fourier as OT
from . import mspaths as OM
from . import rmc as OR
from . import tables as OTab


def problem_from(P):
    t = P.tables
    zs = [OTab.Z.get(e, 26) if isinstance(e, str) else float(e) for e in P.w.elements]
    trs = [OT.Transform(P.k, P.w.kmin, P.w.kmax, kw, P.w.dk_win)
           for e in range(P.nsp) for kw in P.w.kweights]
    ms = T = None
    if P.ms is not None:
        cl = OM.enumerate_classes(P.w.lattice, P.w.a0, P.w.rmax)
        ms = OM.SupercellPaths(cl, P.nsp, P.cell.ideal, P.w.a0, P.w.ncell)
        unit = P.w.a0 / 2
        T = OTab.ms_entry_table(P.k, zs, t.f, t.lam, OM.scattering_sequences(cl, P.nsp),
                                [c.reff_units() * unit for c in cl], s02=t.s02,
                                sigma2_nn=P.w.sigma2, r_nn=float(P.radii[0]))
    return OR.Problem(P.k, t.f, t.delta, t.lam, t.s02, P.cell.ideal, P.i, P.j, P.shift, P.shell,
                      P.w.sigma2, trs, P.chan_edge, P.target, P.sigma, R=P.R, ms=ms, T=T)
