"""Host-side code of the product against the oracle (no GPU), and the C ABI surface."""
import ctypes
import os
import re

import numpy as np
import pytest

from chimera_amd import _lib
from chimera_amd.lattice import CrystalSupercell, alpha_from_counts, warren_cowley
from chimera_amd.mspaths import MSLibrary, find_classes
from chimera_amd.scattering import k_grid, synthetic
from oracle import lattice as OL
from oracle import mspaths as OM
from oracle import tables as OTab

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("lat,nc,a0,rmax", [("fcc", (3, 3, 3), 3.59, 5.2),
                                             ("fcc", (5, 5, 5), 3.59, 5.2),
                                             ("bcc", (4, 4, 4), 2.87, 4.1),
                                             ("bcc", (8, 4, 4), 3.0, 4.4)])
def test_neighbour_list_matches_oracle(lat, nc, a0, rmax):
    c = CrystalSupercell(lat, nc, a0)
    i, j, sh = c.neighbour_list(rmax)
    o = OL.Supercell(lat, nc, a0)
    i2, j2, sh2 = o.neighbour_list(rmax)
    assert np.array_equal(i, i2) and np.array_equal(j, j2)
    np.testing.assert_allclose(sh, sh2, atol=1e-12)
    s1, r1 = c.shells(i, j, sh)
    s2, r2 = OL.shell_index(o.ideal, i2, j2, sh2)
    assert np.array_equal(s1, s2)
    np.testing.assert_allclose(r1, r2, rtol=1e-14)


def test_too_small_box_is_rejected():
    with pytest.raises(ValueError):
        CrystalSupercell("fcc", (2, 2, 2), 3.59).neighbour_list(5.2)


def test_synthetic_tables_match_oracle():
    t = synthetic(("Cr", "Co", "Ni"), k_grid())
    k, f, d, lam, s02 = OTab.synthetic_ss(OTab.k_grid(), [24, 27, 28])
    np.testing.assert_array_equal(t.k, k)
    np.testing.assert_allclose(t.f, f, rtol=1e-15)
    np.testing.assert_allclose(t.delta, d, rtol=1e-15, atol=1e-15)
    np.testing.assert_allclose(t.lam, lam, rtol=1e-15)
    assert t.s02 == s02


@pytest.mark.parametrize("a0,rmax", [(3.56, 5.5), (3.59, 5.2)])
def test_ms_classes_match_oracle(a0, rmax):
    mine = find_classes("fcc", a0, rmax)
    ref = OM.enumerate_classes("fcc", a0, rmax)
    assert [c.rep.tolist() for c in mine] == [c.rep.tolist() for c in ref]
    for a, b in zip(mine, ref):
        assert len(a.images) == len(b.images)
        for (ia, sa), (ib, sb) in zip(a.images, b.images):
            assert np.array_equal(ia, ib) and np.array_equal(sa, sb)


def _emulate_counts(lib, cell, species, i, j, shell, nsp):
    """CPU emulation of the device's full-count kernel over the stencil records (the
    same index arithmetic), to check the stencil format before any GPU run."""
    nshell = int(shell.max()) + 1
    nss = nsp * nshell
    nrow_e = nss + lib.nblk
    out = np.zeros(nsp * nrow_e, np.int64)
    np.add.at(out, species[i] * nrow_e + species[j] * nshell + shell, 1)
    st = lib.stencil.astype(np.int64)
    xyz = cell.xyz % cell.box_units
    for rec in st[:lib.nabsrec]:
        nslot, w = rec[12], rec[13]
        off = (rec[14] & 0xFF) | (rec[15] << 8)
        e = species * nrow_e + nss + off
        pw = 1
        for q in range(1, nslot + 1):
            e = e + species[cell.site_of(xyz + rec[3 * q:3 * q + 3])] * pw
            pw *= nsp
        np.add.at(out, e, w)
    return out


def _emulate_delta(lib, cell, species, i, j, shell, nsp, a, b):
    """CPU emulation of the device's swap gather (SS neighbour items + MS stencil of
    both sites, paths through both counted once)."""
    nshell = int(shell.max()) + 1
    nss = nsp * nshell
    nrow_e = nss + lib.nblk
    ncoord = i.size // cell.natoms
    out = np.zeros(nsp * nrow_e, np.int64)
    na, nb = species[b], species[a]
    new = species.copy()
    new[a], new[b] = na, nb
    for x in (a, b):
        for o in range(ncoord):
            y = j[x * ncoord + o]
            if x == b and y == a:
                continue
            sh = shell[x * ncoord + o]
            for (u, v) in ((x, y), (y, x)):
                out[species[u] * nrow_e + species[v] * nshell + sh] -= 1
                out[new[u] * nrow_e + new[v] * nshell + sh] += 1
    xyz = cell.xyz % cell.box_units
    for x in (a, b):
        for rec in lib.stencil.astype(np.int64):
            nslot, w = rec[12], rec[13]
            off = (rec[14] & 0xFF) | (rec[15] << 8)
            sites = [int(cell.site_of(xyz[x] + rec[3 * q:3 * q + 3])) for q in range(nslot + 1)]
            if x == b and a in sites:
                continue
            eo = species[sites[0]] * nrow_e + nss + off
            en = new[sites[0]] * nrow_e + nss + off
            pw = 1
            for q in range(1, nslot + 1):
                eo += species[sites[q]] * pw
                en += new[sites[q]] * pw
                pw *= nsp
            out[eo] -= w
            out[en] += w
    return out


def test_stencil_counts_match_oracle():
    cell = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(5.2)
    shell, radii = cell.shells(i, j, sh)
    lib = MSLibrary(cell, 3, 5.2)
    ref_cl = OM.enumerate_classes("fcc", 3.59, 5.2)
    sc = OM.SupercellPaths(ref_cl, 3, cell.ideal, 3.59, (3, 3, 3))
    assert lib.nblk == sc.nblk
    sp = OL.random_config(108, [1, 1, 1], 11).astype(np.int64)
    got = _emulate_counts(lib, cell, sp, i, j, shell, 3)
    nss = 3 * 4
    nrow_e = nss + lib.nblk
    for a in range(3):
        ms = got[a * nrow_e + nss:(a + 1) * nrow_e]
        assert np.array_equal(ms, sc.counts(sp)[a * sc.nblk:(a + 1) * sc.nblk])
    N = np.zeros((3, 3, 4), np.int64)
    np.add.at(N, (sp[i], sp[j], shell), 1)
    assert np.array_equal(got.reshape(3, nrow_e)[:, :nss].reshape(3, 3, 4), N)
    # the swap gather gives exactly the count change of the swap
    rng = np.random.default_rng(1)
    done = 0
    while done < 6:
        a, b = (int(x) for x in rng.integers(0, 108, 2))
        if sp[a] == sp[b]:
            continue
        d = _emulate_delta(lib, cell, sp, i, j, shell, 3, a, b)
        sp2 = sp.copy()
        sp2[a], sp2[b] = sp[b], sp[a]
        assert np.array_equal(_emulate_counts(lib, cell, sp2, i, j, shell, 3) - got, d)
        done += 1


def test_stencil_sizes_workload_4000():
    cell = CrystalSupercell("fcc", (10, 10, 10), 3.56)
    lib = MSLibrary(cell, 3, 5.5)
    assert len(lib.classes) == 25 and lib.nblk == 327
    assert lib.directed_per_absorber == 984
    w = lib.stencil[:lib.nabsrec, 13].astype(int)
    assert w.sum() == 984                          # merged weights cover every directed path
    assert np.abs(lib.stencil[:, :12]).max() < 20  # offsets stay inside the 20-unit box


def test_alpha_from_counts_matches_warren_cowley():
    cell = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = cell.neighbour_list(5.2)
    shell, _ = cell.shells(i, j, sh)
    sp = OL.random_config(108, [1, 1, 1], 2)
    N = np.zeros((3, 3))
    m = shell == 0
    np.add.at(N, (sp[i[m]], sp[j[m]]), 1)
    np.testing.assert_allclose(alpha_from_counts(N, np.bincount(sp)),
                               OL.warren_cowley(sp, i, shell, j, 0, 3), rtol=1e-14, atol=1e-14)
    np.testing.assert_allclose(warren_cowley(sp, i, j, shell, 0, 3),
                               OL.warren_cowley(sp, i, shell, j, 0, 3), rtol=1e-14, atol=1e-14)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libchimera_hip.so not built (run __graft_entry__.build())")
    hdr = open(os.path.join(ROOT, "include", "chimera_hip.h")).read()
    declared = set(re.findall(r"\b(chm_[a-z0-9_]+)\s*\(", hdr))
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.EXPORTS)
    assert _lib.load().chm_version() >= 1


def test_thermal_sigma2_springs_in_series():
    """Harmonic chain 0-1-2 along x with springs k1, k2 and free ends: each bond's
    extension is an independent mode, so <du^2> = kB T / k for a bond and the sum of
    both for the 0-2 pair; motion across the bond contributes nothing."""
    from types import SimpleNamespace
    from chimera_amd.forward import KB, thermal_sigma2
    k1, k2, T = 2.0, 5.0, 300.0
    Phi = np.zeros((9, 9))
    for (a, b), k in (((0, 1), k1), ((1, 2), k2)):
        for p, q, sgn in ((a, a, 1), (b, b, 1), (a, b, -1), (b, a, -1)):
            Phi[3 * p, 3 * q] += sgn * k
    cell = SimpleNamespace(ideal=np.array([[0.0, 0, 0], [2.5, 0, 0], [5.0, 0, 0]]))
    i = np.array([0, 1, 0, 1])
    j = np.array([1, 2, 2, 0])
    s2 = thermal_sigma2(Phi, cell, i, j, np.zeros((4, 3)), temperature=T)
    kt = KB * T
    np.testing.assert_allclose(s2, [kt / k1, kt / k2, kt / k1 + kt / k2, kt / k1], rtol=1e-10)
    # a bond direction perpendicular to the springs sees no relative motion
    cell2 = SimpleNamespace(ideal=np.array([[0.0, 0, 0], [0, 2.5, 0], [0, 5.0, 0]]))
    assert np.allclose(thermal_sigma2(Phi, cell2, i[:1], j[:1], np.zeros((1, 3)), T), 0.0)


def test_force_constant_matrix_springs():
    """Nearest-neighbour springs (exafs_gpu/lattice.py:97-111): symmetric, a rigid
    translation costs nothing, and one bond's block is k_bond e e^T + k_angle (1 - e e^T);
    the thermal sigma^2 from it is one value per shell (cubic symmetry) and grows with
    distance."""
    from chimera_amd.forward import thermal_sigma2
    from chimera_amd.spectrum import force_constant_matrix
    c = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = c.neighbour_list(5.2)
    Phi = force_constant_matrix(c, i, j, sh, k_bond=2.6, k_angle=0.35)
    np.testing.assert_allclose(Phi, Phi.T, atol=1e-13)
    for ax in range(3):
        t = np.zeros((c.natoms, 3))
        t[:, ax] = 1.0
        assert np.abs(Phi @ t.reshape(-1)).max() < 1e-12
    d = np.linalg.norm(c.ideal[j] + sh - c.ideal[i], axis=1)
    for p in np.flatnonzero(i == 0)[:20]:
        e = (c.ideal[j[p]] + sh[p] - c.ideal[0]) / d[p]
        want = (-(2.6 * np.outer(e, e) + 0.35 * (np.eye(3) - np.outer(e, e)))
                if d[p] < 2.6 else np.zeros((3, 3)))
        np.testing.assert_allclose(Phi[0:3, 3 * j[p]:3 * j[p] + 3], want, atol=1e-12)
    np.testing.assert_allclose(Phi[0:3, 0:3], (2.6 + 2 * 0.35) * 4 * np.eye(3), atol=1e-12)
    shell, radii = c.shells(i, j, sh)
    s2 = thermal_sigma2(Phi, c, i, j, sh, 300.0)
    for s in range(radii.size):
        v = s2[shell == s]
        assert np.ptp(v) <= 1e-10 * v.mean()


def test_k_exp_axis_and_expected_counts_match_oracle():
    from chimera_amd.spectrum import k_exp_axis
    k = k_grid()
    for dE0 in (0.0, 2.0, -3.5):
        np.testing.assert_array_equal(k_exp_axis(k, dE0), OTab.k_exp_axis(k, dE0))
    c = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    lib = MSLibrary(c, 3, 5.2)
    cl = OM.enumerate_classes("fcc", 3.59, 5.2)
    for comp in ([1, 1, 1], [0.2, 0.3, 0.5]):
        np.testing.assert_allclose(lib.expected_counts(comp), OM.expected_counts(cl, 3, comp),
                                   rtol=1e-14)


def test_path_library_reader_matches_oracle(tmp_path):
    """PathLibrary (chimera/mslibrary.py:261-313) on a FEFF-format library: the same
    classes, entries and stencil as the enumerated library, and F / delta / lambda on the
    model grid equal to the oracle's own reading of the files."""
    from chimera_amd.mspaths import PathLibrary
    from tests.feffio import write_library
    from tests.test_oracle import _fdl
    c = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    ref = MSLibrary(c, 2, 5.2)
    root = write_library(str(tmp_path / "lib"), ("Fe", "Ni"), ref.classes, c.unit,
                         c.unit * np.sqrt(2), _fdl([26, 28]))
    k = k_grid()
    lib = PathLibrary(root).prepare(k)
    ms = MSLibrary(c, 2, 5.2, library=lib)
    assert ms.nblk == ref.nblk == lib.nblk
    assert np.array_equal(ms.stencil, ref.stencil)
    rows = 0
    for a, el in enumerate(("Fe", "Ni")):
        for n in range(lib.nblk):
            fn = os.path.join(root, el, f"feff{n + 1:04d}.dat")
            nleg, deg, reff, cols = OTab.read_feff_columns(fn)
            p, f, d, lam = OTab.feff_columns_to_path(cols, reff)
            np.testing.assert_allclose(lib.F[rows], np.interp(k, p, f), rtol=1e-15, atol=0)
            np.testing.assert_allclose(lib.D[rows], np.interp(k, p, d), rtol=1e-15, atol=0)
            np.testing.assert_allclose(lib.L[rows], np.interp(k, p, lam), rtol=1e-15, atol=0)
            assert lib.reff[rows] == reff
            rows += 1
    with pytest.raises(AssertionError):
        os.remove(os.path.join(root, "Ni", "feff0001.dat"))
        PathLibrary(root)


def test_shell_model_directions():
    """dN_direction keeps every absorber's coordination and the pair symmetry; dr_from_free
    gives symmetric offsets of zero composition-weighted mean (chimera/shellmodel.py:45-104)."""
    from chimera_amd.shellmodel import dN_direction, dr_from_free, sro_directions
    for nsp in (2, 3, 4):
        assert len(sro_directions(nsp)) == nsp * (nsp - 1) // 2
        for a, b in sro_directions(nsp):
            d = dN_direction(nsp, 4, 2, a, b)
            np.testing.assert_array_equal(d.sum(axis=(1, 2)), 0)      # coordination kept
            np.testing.assert_array_equal(d, d.transpose(1, 0, 2))
        comp = np.arange(1, nsp + 1, dtype=float)
        comp /= comp.sum()
        nfree = nsp * (nsp + 1) // 2 - 1
        dr = dr_from_free(np.random.default_rng(nsp).normal(0, 0.01, nfree), comp)
        np.testing.assert_allclose(dr, dr.T)
        assert abs(comp @ dr @ comp) < 1e-16


def test_warren_cowley_radius_matches_shell_index():
    """The reference's radius form of Warren-Cowley (exafs_gpu/lattice.py:167-185) selects
    exactly the pairs of one lattice shell when tol is below half the shell spacing."""
    from chimera_amd.lattice import random_config, warren_cowley_radius
    c = CrystalSupercell("fcc", (3, 3, 3), 3.59)
    i, j, sh = c.neighbour_list(5.2)
    shell, radii = c.shells(i, j, sh)
    sp = random_config(108, [1, 2, 1], 4)
    for s in range(4):
        np.testing.assert_allclose(warren_cowley_radius(sp, i, j, sh, c, radii[s], nsp=3),
                                   OL.warren_cowley(sp, i, shell, j, s, 3), rtol=0, atol=1e-14)
