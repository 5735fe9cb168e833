"""Learned swap proposal: the oracle's FP64 restatement (test infrastructure only).

The reference has no learned proposal; this restates the device algorithm so the
device can be checked step by step, and it is pinned independently of the device:
  * the MLP gradient by central finite differences in FP64 (tests/test_oracle.py);
  * the proposal by detailed balance: with the Hastings ratio below, every move's
    forward and reverse probabilities multiply to the posterior ratio (checked by
    enumerating a small candidate set, tests/test_oracle.py);
  * bf16 rounding by its definition (round to nearest even of the top 16 bits).

Network (per candidate swap (a, b) at state x): occupancy occ[0..nocc) with
  occ[q]            species of the q-th stencil neighbour of a   (q < ncoord)
  occ[ncoord + q]   species of the q-th stencil neighbour of b
  occ[2 ncoord]     species of a,   occ[2 ncoord + 1]  species of b
one-hot encoded as X[nsp*q + s] = [occ[q] == s]; then
  Z1 = b1 + X W1,  H1 = bf16(silu(Z1)),  Z2 = b2 + H1 W2,  H2 = silu(Z2),
  g = b3 + H2 w3,
with W1, W2, w3 used through bf16 rounding and every sum in FP32 on the device
(FP64 here).  g predicts the energy change nfit * (chi2(x') - chi2(x)) of the swap.

Proposal (per replica and step): 256 candidate pairs drawn from a counter-based
stream that does not depend on the state; logit_i = -lam * beta * g_i / 2 +
kappa * |dU_i| (dU: change of unlike first-shell bonds), -inf where the two sites
hold the same species; j ~ softmax.  Reverse probability: the same candidate set,
scored at x' = swap(x, a_j, b_j).  Acceptance: x = beta dE - log(q_rev / q_fwd),
where q sums the probabilities of every candidate naming the same unordered pair.
With probability eps, and on every rung colder than beta_max, a step uses the plain
uniform proposal instead (a state-independent mixture of two kernels that each keep
the posterior).

Lifted variant (lift=True): each chain carries a direction d = +-1; the tilt is
kappa * d * dU, the reverse probability is scored with -d, and d flips whenever a learned
step is not accepted (also when no candidate is valid).  Skew detailed balance
pi(x) T_d(x, y) = pi(y) T_-d(y, x) then keeps pi x (1/2, 1/2) stationary (Diaconis,
Holmes & Neal, Ann. Appl. Probab. 10 (2000) 726; Turitsyn, Chertkov & Vucelja, Physica D
240 (2011) 410); pinned by exact enumeration in tests/test_oracle_learned.py.
"""
from __future__ import annotations

import numpy as np

from .rmc import GOLDEN, M64, splitmix64, u01

HID = 256
LP_SALT = 0x2545F4914F6CDD1D
COIN_SALT = 0x5851F42D4C957F2D
SEL_SALT = 0x14057B7EF767814F
CAND_MUL = 0xD6E8FEB86659FD93


# ------------------------------------------------------------------ bf16
def bf16(x):
    """Round to bfloat16 (nearest even) through float32, returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


# ------------------------------------------------------------------ parameters
def layout(nfeat):
    """Flat FP32 parameter layout: W1 (KP, H) row f, b1 (H), W2 (H, H) row k, b2, w3, b3."""
    kp = (nfeat + 31) // 32 * 32
    sizes = [("W1", kp * HID), ("b1", HID), ("W2", HID * HID), ("b2", HID), ("w3", HID), ("b3", 1)]
    off, o = {}, 0
    for name, n in sizes:
        off[name] = (o, n)
        o += n
    return kp, off, o


def unpack(flat, nfeat):
    kp, off, _ = layout(nfeat)
    flat = np.asarray(flat, np.float64)
    g = {k: flat[o:o + n] for k, (o, n) in off.items()}
    g["W1"] = g["W1"].reshape(kp, HID)
    g["W2"] = g["W2"].reshape(HID, HID)
    return g


def init_params(nfeat, nocc, seed=0, w3_scale=0.0):
    """Seeded parameters for oracle-only tests: W1 ~ N(0, 1/nocc), W2 ~ N(0, 1/H),
    w3 = w3_scale N(0, 1/H), biases 0."""
    kp, off, n = layout(nfeat)
    rng = np.random.default_rng(seed)
    flat = np.zeros(n, np.float32)
    o, m = off["W1"]
    W1 = np.zeros((kp, HID), np.float32)
    W1[:nfeat] = rng.standard_normal((nfeat, HID)) / np.sqrt(nocc)
    flat[o:o + m] = W1.ravel()
    o, m = off["W2"]
    flat[o:o + m] = (rng.standard_normal(m) / np.sqrt(HID)).astype(np.float32)
    o, m = off["w3"]
    flat[o:o + m] = (w3_scale * rng.standard_normal(m) / np.sqrt(HID)).astype(np.float32)
    return flat


# ------------------------------------------------------------------ MLP
def one_hot(occ, nsp, kp):
    occ = np.asarray(occ, np.int64)
    n, nocc = occ.shape
    X = np.zeros((n, kp))
    cols = nsp * np.arange(nocc)[None, :] + occ
    X[np.repeat(np.arange(n), nocc), cols.ravel()] = 1.0
    return X


W1_SCALE = 65536.0          # the scorer's layer-1 weights are integers in units of 2^-16
W1_QMAX = 131072.0          # ... clamped to |w| <= 2, so a sum of <= 110 stays below 2^24


def w1_fixed(W1):
    """The scorer's layer-1 weights: bf16(W1) rounded (half to even) to a multiple of
    2^-16 and clamped to [-2, 2].  Sums of them over one-hot features are exact in FP64 and,
    as integers below 2^24, in FP32: the device keeps per-site layer-1 sums incrementally
    and still gets the same value as a sum from scratch."""
    return np.clip(np.rint(np.asarray(W1, np.float64) * W1_SCALE), -W1_QMAX, W1_QMAX) / W1_SCALE


def forward(flat, occ, nsp, rounding=True, scorer=False):
    """g (n,) and the intermediates, FP64 with the bf16 roundings of the device
    (rounding=False: the same network without them, for finite differences).
    scorer=True: layer 1 with the scorer's fixed-point weights (w1_fixed), the network
    that scores candidate swaps; the trainer differentiates the bf16 one."""
    occ = np.asarray(occ, np.int64)
    nfeat = occ.shape[1] * nsp
    p = unpack(flat, nfeat)
    kp = p["W1"].shape[0]
    X = one_hot(occ, nsp, kp)
    rnd = bf16 if rounding else (lambda v: np.asarray(v, np.float64))
    W1, W2, w3 = rnd(p["W1"]), rnd(p["W2"]), rnd(p["w3"])
    if scorer:
        W1 = w1_fixed(W1)
    Z1 = X @ W1 + p["b1"]
    H1 = rnd(silu(Z1))
    Z2 = H1 @ W2 + p["b2"]
    H2 = silu(Z2)
    g = H2 @ w3 + p["b3"][0]
    return g, dict(X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, W1=W1, W2=W2, w3=w3)


def loss_and_grad(flat, occ, y, w, nsp, rounding=True):
    """Weighted mean 0.5 (g - y)^2 and its gradient w.r.t. the flat parameters
    (bf16 roundings passed straight through, as the device trains its FP32 masters)."""
    g, c = forward(flat, occ, nsp, rounding)
    y = np.asarray(y, np.float64)
    w = np.asarray(w, np.float64)
    W = max(w.sum(), 1e-30)
    r = g - y
    loss = 0.5 * (w * r * r).sum() / W
    dg = w * r / W
    nfeat = occ.shape[1] * nsp
    kp, off, n = layout(nfeat)
    grad = np.zeros(n)
    dw3 = c["H2"].T @ dg
    db3 = dg.sum()
    dZ2 = dg[:, None] * c["w3"][None, :] * dsilu(c["Z2"])
    dW2 = c["H1"].T @ dZ2
    db2 = dZ2.sum(0)
    dH1 = dZ2 @ c["W2"].T
    dZ1 = dH1 * dsilu(c["Z1"])
    dW1 = c["X"].T @ dZ1
    db1 = dZ1.sum(0)
    for name, v in (("W1", dW1), ("b1", db1), ("W2", dW2), ("b2", db2), ("w3", dw3),
                    ("b3", np.array([db3]))):
        o, m = off[name]
        grad[o:o + m] = np.ravel(v)
    return loss, grad


def adam(flat, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected), t counted from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------------ lattice stencil
class Stencil:
    """The neighbour stencil of the sampler: offsets of site 0's neighbours in the
    neighbour-list order, applied to every site with periodic wrap."""

    def __init__(self, ideal, i_idx, j_idx, shift, shell, box_len):
        ideal = np.asarray(ideal, float)
        first = np.nonzero(np.asarray(i_idx) == 0)[0]
        self.off = ideal[np.asarray(j_idx)[first]] + np.asarray(shift)[first] - ideal[0]
        self.shell = np.asarray(shell)[first]
        self.ncoord = first.size
        self.box = np.asarray(box_len, float)
        self.ideal = ideal
        self._L = np.round(self.box * 1000).astype(np.int64)
        self._index = {self._key(p): q for q, p in enumerate(ideal)}

    def _key(self, pos):
        return tuple(np.round(np.asarray(pos) * 1000).astype(np.int64) % self._L)

    def site_at(self, pos):
        return self._index[self._key(pos)]

    def neighbours(self, q):
        return [self.site_at(self.ideal[q] + o) for o in self.off]


def occupancy(sp, st: Stencil, a, b):
    na = st.neighbours(a)
    nb = st.neighbours(b)
    occ = [int(sp[s]) for s in na] + [int(sp[s]) for s in nb] + [int(sp[a]), int(sp[b])]
    return np.array(occ, np.int64), na, nb


def unlike_weights(nsp):
    """The default bond weights of the tilt's count: 1 for an unlike pair, 0 for a like one."""
    return 1 - np.eye(nsp, dtype=np.int64)


def unlike_change(sp, st: Stencil, a, b, na, nb, w=None):
    """Change of sum_bonds w[s_i, s_j] over the first-shell bonds when a and b swap
    (w: symmetric bond weights; default: the number of unlike bonds).  The a-b bond keeps
    its pair and is skipped."""
    A, B = int(sp[a]), int(sp[b])
    if w is None:
        w = unlike_weights(max(int(sp.max()), A, B) + 1)
    w = np.asarray(w, np.int64)
    d = 0
    for q, s in enumerate(na):
        if st.shell[q] == 0 and s != b:
            d += int(w[B, sp[s]]) - int(w[A, sp[s]])
    for q, s in enumerate(nb):
        if st.shell[q] == 0 and s != a:
            d += int(w[A, sp[s]]) - int(w[B, sp[s]])
    return d


# ------------------------------------------------------------------ proposal
def stream(seed, stepc, key):
    return splitmix64((seed ^ ((stepc * GOLDEN) & M64) ^ ((key << 32) & M64) ^ LP_SALT) & M64)


def candidates(seed, stepc, key, natoms, ncand):
    h0 = stream(seed, stepc, key)
    out = []
    for i in range(ncand):
        h = splitmix64((h0 ^ (((i + 1) * CAND_MUL) & M64)) & M64)
        a = ((h & 0xFFFFFFFF) * natoms) >> 32
        b = ((h >> 32) * natoms) >> 32
        out.append((int(a), int(b)))
    return out


def uses_uniform(seed, stepc, key, eps):
    h0 = stream(seed, stepc, key)
    return u01(splitmix64((h0 ^ COIN_SALT) & M64)) < eps


def select_uniform(seed, stepc, key):
    h0 = stream(seed, stepc, key)
    return u01(splitmix64((h0 ^ SEL_SALT) & M64))


def tilt(kappa, du, d=None):
    """kappa |dU| (symmetric), or kappa d dU for a lifted proposal in direction d."""
    du = np.asarray(du, float)
    return kappa * np.abs(du) if d is None else kappa * d * du


def logits(flat, sp, st, cands, nsp, beta, lam, kappa, d=None, w=None):
    occs, du, valid = [], [], []
    w = unlike_weights(nsp) if w is None else w
    for a, b in cands:
        occ, na, nb = occupancy(sp, st, a, b)
        occs.append(occ)
        du.append(unlike_change(sp, st, a, b, na, nb, w))
        valid.append(sp[a] != sp[b])
    occs = np.array(occs)
    g, _ = forward(flat, occs, nsp, scorer=True)
    lg = -lam * beta * g / 2.0 + tilt(kappa, du, d)
    lg = np.where(np.array(valid), lg, -np.inf)
    return lg, g, occs


def select(lg, u, with_margin=False):
    """Index j with cumsum_{<j} <= u Z < cumsum_{<=j} (index order), or -1; with_margin
    also returns the distance of u Z to the nearest bin edge, relative to Z."""
    if not np.isfinite(lg).any():
        return (-1, 1.0) if with_margin else -1
    p = np.exp(lg - lg.max())
    c = np.cumsum(p)
    t = u * c[-1]
    j = min(int(np.searchsorted(c, t, side="right")), len(lg) - 1)
    if not with_margin:
        return j
    lo = c[j - 1] if j > 0 else 0.0
    return j, min(abs(t - lo), abs(c[j] - t)) / c[-1]


def log_q(lg, same):
    m = lg.max()
    p = np.exp(lg - m)
    return float(np.log(p[same].sum()) - np.log(p.sum()))


def log_q_ratio(flat, sp, st, nsp, cands, lg, j, beta, lam, kappa, d=None, w=None):
    """log q(x' -> x) - log q(x -> x') for candidate j of the set `cands` at state sp
    (lg: the logits at sp; a lifted proposal in direction d is reversed with -d)."""
    a, b = cands[j]
    same = np.array([{c, e} == {a, b} for c, e in cands])
    sp2 = sp.copy()
    sp2[a], sp2[b] = sp[b], sp[a]
    lg2, _, _ = logits(flat, sp2, st, cands, nsp, beta, lam, kappa, None if d is None else -d, w)
    return log_q(lg2, same) - log_q(lg, same)


def propose_learned(flat, sp, st, nsp, seed, stepc, key, beta, ncand, lam, kappa, d=None, w=None):
    """(a, b, na, nb, sw, logqr, occ_j, margin) of the learned proposal at state sp
    (margin: how far the selecting uniform fell from a bin edge, relative; d: the
    direction of a lifted proposal; w: the tilt's bond weights)."""
    natoms = sp.size
    cands = candidates(seed, stepc, key, natoms, ncand)
    lg, _, occs = logits(flat, sp, st, cands, nsp, beta, lam, kappa, d, w)
    j, margin = select(lg, select_uniform(seed, stepc, key), with_margin=True)
    if j < 0:
        return 0, 0, int(sp[0]), int(sp[0]), 0, 0.0, None, 1.0
    a, b = cands[j]
    logqr = log_q_ratio(flat, sp, st, nsp, cands, lg, j, beta, lam, kappa, d, w)
    return a, b, int(sp[b]), int(sp[a]), 1, logqr, occs[j], margin


# ------------------------------------------------------------------ sampler
class LearnedSampler:
    """The plain sampler (oracle.rmc.PlainSampler) with the learned proposal mixed in:
    the checker of the device sampler in learned mode."""

    def __init__(self, plain, flat, stencil, ncand=256, eps=0.0, lam=1.0, kappa=0.0, beta_max=1e300,
                 lift=False, tilt_w=None):
        self.S = plain
        self.flat = np.asarray(flat, np.float32)
        self.st = stencil
        self.ncand, self.eps, self.lam, self.kappa = ncand, eps, lam, kappa
        self.beta_max = beta_max
        self.margins = []
        self.lift = bool(lift)
        self.tilt_w = None if tilt_w is None else np.asarray(tilt_w, np.int64)
        # directions start alternating over the chains (d travels with its configuration)
        self.dir = np.where(np.arange(plain.nrep) % 2 == 1, -1, 1)

    def step(self):
        from .rmc import accept_uniform, propose
        S, P = self.S, self.S.P
        S.stepc += 1
        rec = []
        for c in range(S.nrep):
            sp = S.species[c]
            u = accept_uniform(S.seed, S.stepc, c)
            margin = 1.0
            learned = not (uses_uniform(S.seed, S.stepc, c, self.eps) or S.beta[c] > self.beta_max)
            if not learned:
                a, b, na, nb, sw = propose(sp, S.seed, S.stepc, c)
                lq = 0.0
            else:
                a, b, na, nb, sw, lq, _, margin = propose_learned(
                    self.flat, sp, self.st, P.nsp, S.seed, S.stepc, c, S.beta[c], self.ncand,
                    self.lam, self.kappa, int(self.dir[c]) if self.lift else None, self.tilt_w)
            self.margins.append((S.stepc, c, margin))
            if not sw:
                if self.lift and learned:
                    self.dir[c] = -self.dir[c]
                rec.append((a, b, 0, S.chi2[c], 0.0, u))
                continue
            dchi, _ = P.delta_chi(sp, a, b, na, nb)
            trial = S.chi[c] + dchi
            c2 = P.chi2(trial, S.cnt, S.profile)
            x = S.beta[c] * (c2 - S.chi2[c]) * P.nfit - lq
            acc = (x <= 0) or (u < np.exp(-min(x, 60.0)))
            if acc:
                sp[a], sp[b] = na, nb
                S.chi[c] = trial
                S.chi2[c] = c2
            elif self.lift and learned:
                self.dir[c] = -self.dir[c]
            rec.append((a, b, int(acc), c2, x, u))
        imp = S.chi2 < S.best_chi2
        S.best_chi2 = np.where(imp, S.chi2, S.best_chi2)
        S.best_sp[imp] = S.species[imp]
        S.log.append(rec)

    def run(self, nsteps, exchange_every=25):
        for it in range(nsteps):
            self.step()
            if exchange_every and (it + 1) % exchange_every == 0:
                order = self.S.replica_exchange()
                self.dir = self.dir[order]
