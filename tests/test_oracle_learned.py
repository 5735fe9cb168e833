"""Pins the oracle of the learned swap proposal: bf16 rounding, MLP gradients by
finite differences, and detailed balance of the proposal + Hastings ratio."""
import itertools

import numpy as np
import pytest

from oracle import lattice as OL
from oracle import learned as LP


def test_bf16_rounding_matches_torch_and_ties_to_even():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(2000) * 10.0 ** rng.integers(-6, 6, 2000),
                        [1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 0.0]]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(LP.bf16(x), want)
    assert LP.bf16(1 + 2 ** -8) == 1.0                     # halfway: to the even mantissa
    assert LP.bf16(1 + 3 * 2 ** -8) == 1 + 2 ** -6


def test_mlp_gradient_matches_finite_differences():
    nsp, nocc, n = 3, 10, 7
    nfeat = nsp * nocc
    rng = np.random.default_rng(1)
    flat = LP.init_params(nfeat, nocc, seed=3, w3_scale=1.0).astype(np.float64)
    kp, off, _ = LP.layout(nfeat)
    for name in ("b1", "b2", "b3"):
        o, m = off[name]
        flat[o:o + m] = rng.standard_normal(m) * 0.3
    occ = rng.integers(0, nsp, (n, nocc))
    y = rng.standard_normal(n)
    w = rng.random(n) + 0.5
    loss, grad = LP.loss_and_grad(flat, occ, y, w, nsp, rounding=False)
    # check a sample of every block, including W1 rows that some row activates
    o1 = off["W1"][0]
    active = sorted({nsp * q + int(occ[r, q]) for r in range(n) for q in range(nocc)})
    idx = [o1 + f * LP.HID + c for f in active[:6] for c in (0, 17, 255)]
    for name in ("b1", "W2", "b2", "w3", "b3"):
        o, m = off[name]
        idx += list(o + rng.choice(m, min(m, 8), replace=False))
    h = 1e-6
    for i in idx:
        fp, fm = flat.copy(), flat.copy()
        fp[i] += h
        fm[i] -= h
        lp_, _ = LP.loss_and_grad(fp, occ, y, w, nsp, rounding=False)
        lm_, _ = LP.loss_and_grad(fm, occ, y, w, nsp, rounding=False)
        fd = (lp_ - lm_) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])
    # a W1 row that no candidate activates gets no gradient
    unused = [f for f in range(nfeat) if f not in active]
    for f in unused:
        assert np.all(grad[o1 + f * LP.HID:o1 + (f + 1) * LP.HID] == 0.0)


def _toy():
    cell = OL.Supercell("sc", (2, 2, 2), 3.0)
    i, j, sh = cell.neighbour_list(1.5 * 3.0)
    shell = np.round(np.linalg.norm(cell.ideal[j] + sh - cell.ideal[i], axis=1) / 3.0 * 10).astype(int)
    _, shell = np.unique(shell, return_inverse=True)
    st = LP.Stencil(cell.ideal, i, j, sh, shell, cell.box)
    return cell, st


def test_stencil_reproduces_the_neighbour_sets():
    cell = OL.Supercell("fcc", (3, 3, 3), 3.6)
    i, j, sh = cell.neighbour_list(3.6 * 1.01)
    shell = np.zeros(i.size, int)
    st = LP.Stencil(cell.ideal, i, j, sh, shell, cell.box)
    for q in (0, 5, 77, cell.natoms - 1):
        assert sorted(st.neighbours(q)) == sorted(j[i == q].tolist())


@pytest.mark.parametrize("lam,kappa", [(1.0, 0.0), (0.7, 0.4)])
def test_learned_proposal_detailed_balance(lam, kappa):
    """For a FIXED candidate set, the transition matrix of (softmax proposal, Hastings
    acceptance) over every state of a binary 8-site toy satisfies
    pi(x) T(x, y) = pi(y) T(y, x); T(y, x) uses y's own forward probabilities."""
    cell, st = _toy()
    nsp, n = 2, cell.natoms
    nfeat = nsp * (2 * st.ncoord + 2)
    flat = LP.init_params(nfeat, 2 * st.ncoord + 2, seed=5, w3_scale=3.0)
    # candidates with a repeat, a reversed repeat and a same-site pair
    cands = [(0, 3), (1, 6), (3, 0), (2, 5), (4, 7), (0, 3), (5, 5), (6, 1), (2, 7)]
    states = [np.array(c) for c in itertools.product((0, 1), repeat=n) if sum(c) == n // 2]
    index = {tuple(s): k for k, s in enumerate(states)}
    rng = np.random.default_rng(7)
    E = rng.standard_normal(len(states)) * 1.5
    beta = 0.8
    T = np.zeros((len(states), len(states)))
    for k, sp in enumerate(states):
        lg, _, _ = LP.logits(flat, sp, st, cands, nsp, beta, lam, kappa)
        if not np.isfinite(lg).any():
            continue
        p = np.exp(lg - lg.max())
        p /= p.sum()
        for jj, (a, b) in enumerate(cands):
            if not np.isfinite(lg[jj]):
                continue
            y = sp.copy()
            y[a], y[b] = sp[b], sp[a]
            ky = index[tuple(y)]
            lqr = LP.log_q_ratio(flat, sp, st, nsp, cands, lg, jj, beta, lam, kappa)
            A = min(1.0, np.exp(-beta * (E[ky] - E[k]) + lqr))
            T[k, ky] += p[jj] * A
    pi = np.exp(-beta * E)
    pi /= pi.sum()
    F = pi[:, None] * T
    assert np.allclose(F, F.T, rtol=1e-10, atol=1e-15)
    assert T.sum() > 1.0                                    # moves do happen


@pytest.mark.parametrize("lam,kappa", [(1.0, 0.7), (0.5, 2.0)])
def test_lifted_proposal_skew_balance_and_stationarity(lam, kappa):
    """The lifted proposal on a FIXED candidate set over every state x direction of a
    binary 8-site toy: skew detailed balance pi(x) T_d(x, y) = pi(y) T_-d(y, x) for every
    move, and with the direction flipped on every non-accepted step the product law
    pi x (1/2, 1/2) is stationary (while the unlifted kernel with the signed tilt, a
    mistake one could make, is not reversible)."""
    cell, st = _toy()
    nsp, n = 2, cell.natoms
    nfeat = nsp * (2 * st.ncoord + 2)
    flat = LP.init_params(nfeat, 2 * st.ncoord + 2, seed=5, w3_scale=3.0)
    cands = [(0, 3), (1, 6), (3, 0), (2, 5), (4, 7), (0, 3), (5, 5), (6, 1), (2, 7)]
    states = [np.array(c) for c in itertools.product((0, 1), repeat=n) if sum(c) == n // 2]
    index = {tuple(s): k for k, s in enumerate(states)}
    ns = len(states)
    rng = np.random.default_rng(11)
    E = rng.standard_normal(ns) * 1.5
    beta = 0.8
    pi = np.exp(-beta * E)
    pi /= pi.sum()
    # T[(x, d) -> (y, d')] with d index 0: +1, 1: -1
    T = np.zeros((2 * ns, 2 * ns))
    Td = {1: np.zeros((ns, ns)), -1: np.zeros((ns, ns))}
    for k, sp in enumerate(states):
        for di, d in enumerate((1, -1)):
            lg, _, _ = LP.logits(flat, sp, st, cands, nsp, beta, lam, kappa, d)
            moved = 0.0
            if np.isfinite(lg).any():
                p = np.exp(lg - lg.max())
                p /= p.sum()
                for jj, (a, b) in enumerate(cands):
                    if not np.isfinite(lg[jj]):
                        continue
                    y = sp.copy()
                    y[a], y[b] = sp[b], sp[a]
                    ky = index[tuple(y)]
                    lqr = LP.log_q_ratio(flat, sp, st, nsp, cands, lg, jj, beta, lam, kappa, d)
                    A = min(1.0, np.exp(-beta * (E[ky] - E[k]) + lqr))
                    T[di * ns + k, di * ns + ky] += p[jj] * A
                    Td[d][k, ky] += p[jj] * A
                    moved += p[jj] * A
            T[di * ns + k, (1 - di) * ns + k] += 1.0 - moved        # not accepted: flip d
    for d in (1, -1):
        F = pi[:, None] * Td[d]
        assert np.allclose(F, (pi[:, None] * Td[-d]).T, rtol=1e-10, atol=1e-15)
    assert np.allclose(T.sum(1), 1.0)
    mu = np.concatenate([pi, pi]) / 2.0
    assert np.allclose(mu @ T, mu, rtol=1e-10, atol=1e-15)
    assert Td[1].sum() > 1.0
    # a plausible mistake, the reverse move scored in the forward direction (d, not -d)
    # with the same flips, breaks the stationarity of pi x (1/2, 1/2)
    Tw = np.zeros((2 * ns, 2 * ns))
    for k, sp in enumerate(states):
        for di, d in enumerate((1, -1)):
            lg, _, _ = LP.logits(flat, sp, st, cands, nsp, beta, lam, kappa, d)
            moved = 0.0
            if np.isfinite(lg).any():
                p = np.exp(lg - lg.max())
                p /= p.sum()
                for jj, (a, b) in enumerate(cands):
                    if not np.isfinite(lg[jj]):
                        continue
                    y = sp.copy()
                    y[a], y[b] = sp[b], sp[a]
                    ky = index[tuple(y)]
                    same = np.array([{c, e} == {a, b} for c, e in cands])
                    lg2, _, _ = LP.logits(flat, y, st, cands, nsp, beta, lam, kappa, d)
                    A = min(1.0, np.exp(-beta * (E[ky] - E[k]) + LP.log_q(lg2, same) - LP.log_q(lg, same)))
                    Tw[di * ns + k, di * ns + ky] += p[jj] * A
                    moved += p[jj] * A
            Tw[di * ns + k, (1 - di) * ns + k] += 1.0 - moved
    assert np.abs(mu @ Tw - mu).max() > 1e-4


def test_candidate_stream_is_uniform_and_state_free():
    c1 = LP.candidates(11, 5, 3, 4000, 256)
    c2 = LP.candidates(11, 5, 3, 4000, 256)
    assert c1 == c2 and len(set(c1)) > 250
    a = np.array([x for x, _ in LP.candidates(1, 1, 0, 1000, 20000)])
    counts = np.bincount(a // 100, minlength=10)
    assert counts.min() > 1800 and counts.max() < 2200


def test_scorer_layer1_weights_fixed_point():
    """The scorer's layer-1 weights: bf16 values of magnitude >= 2^-9 are already on the
    2^-16 grid and stay unchanged; every value is a multiple of 2^-16 within [-2, 2]; a sum
    of 110 of them is exact in FP32 (below 2^24 grid units), so a per-site sum kept by
    increments equals one formed from scratch."""
    rng = np.random.default_rng(5)
    w = LP.bf16(rng.standard_normal(20000) * 10.0 ** rng.uniform(-5, 0.5, 20000))
    q = LP.w1_fixed(w)
    assert np.all(q * 65536.0 == np.rint(q * 65536.0))
    assert np.abs(q).max() <= 2.0
    big = (np.abs(w) >= 2.0 ** -9) & (np.abs(w) <= 2.0)
    np.testing.assert_array_equal(q[big], w[big])
    small = np.abs(w) < 2.0 ** -9
    assert np.all(np.abs(q[small] - w[small]) <= 2.0 ** -17)
    assert LP.w1_fixed(np.array([3.5, -7.0])).tolist() == [2.0, -2.0]
    # half a grid step rounds to even
    assert LP.w1_fixed(np.array([2.0 ** -17, 3 * 2.0 ** -17])).tolist() == [0.0, 2.0 ** -15]
    rows = q.reshape(-1, 100)[:, :110 if q.size >= 110 else 100]
    for r in rows[:50]:
        pick = rng.choice(r.size, min(110, r.size), replace=False)
        exact = r[pick].sum()
        f32 = np.float32(0.0)
        for v in r[pick]:
            f32 = np.float32(f32 + np.float32(v))
        assert float(f32) == exact


def test_weighted_bond_change_equals_recount():
    """dU of a swap with symmetric bond weights w equals the change of
    sum over first-shell bonds of w[s_i, s_j] recounted over the whole configuration
    (including swaps of two first-shell neighbours), and with the default weights it is
    the change of the number of unlike bonds."""
    cell = OL.Supercell("fcc", (3, 3, 3), 3.6)
    i, j, sh = cell.neighbour_list(3.6 * 1.01)
    d = np.linalg.norm(cell.ideal[j] + sh - cell.ideal[i], axis=1)
    shell = np.where(d < 3.6 * 0.75, 0, 1)
    st = LP.Stencil(cell.ideal, i, j, sh, shell, cell.box)
    rng = np.random.default_rng(7)
    nsp = 3
    first = [[s for s, q in zip(st.neighbours(x), range(st.ncoord)) if st.shell[q] == 0]
             for x in range(cell.natoms)]

    def total(sp, w):
        return sum(int(w[sp[x], sp[y]]) for x in range(cell.natoms) for y in first[x]) // 2

    for trial in range(30):
        sp = rng.integers(0, nsp, cell.natoms)
        W = rng.integers(-5, 6, (nsp, nsp))
        W = W + W.T
        a = int(rng.integers(cell.natoms))
        b = int(first[a][rng.integers(len(first[a]))]) if trial % 3 == 0 else int(rng.integers(cell.natoms))
        if sp[a] == sp[b]:
            continue
        _, na, nb = LP.occupancy(sp, st, a, b)
        sp2 = sp.copy()
        sp2[a], sp2[b] = sp[b], sp[a]
        assert LP.unlike_change(sp, st, a, b, na, nb, W) == total(sp2, W) - total(sp, W)
        U = LP.unlike_weights(nsp)
        assert LP.unlike_change(sp, st, a, b, na, nb) == total(sp2, U) - total(sp, U)
