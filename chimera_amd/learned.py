"""Learned swap proposal of the device sampler (host side).

No counterpart in the reference, whose swap proposal is uniform
(exafs_gpu/kernels.py:362-398).  The network, its scoring of candidate swaps, the
Hastings ratio, the move records and the trainer all live on the device
(chimera_amd/csrc/learned.hip); this class owns the parameters' initialisation and
the calls:

    lp = LearnedProposal(sampler, seed=0)
    lp.record(2_000_000)                 # ring of move records written by every step
    sampler.run(20000)                   # plain uniform proposals produce the records
    lp.train(iters=3000)                 # Adam on minibatches of the ring
    lp.enable()                          # learned proposal from now on (uniform at eps)
    sampler.run(...)

Network: one-hot species occupancy of both candidate sites' stencil neighbours and of
the two sites (nocc = 2 ncoord + 2 entries) -> 256 -> 256 -> 1 with SiLU, bf16 weights,
FP32 accumulation; the scorer's layer-1 weights are bf16(W1) on a 2^-16 grid, so each
site's layer-1 sum is an exact integer the device keeps per replica and updates per move.  It predicts the swap's energy change nfit * dchi2; a candidate's
logit is -lam * beta * g / 2 + kappa * |dU| (dU: change of unlike first-shell bonds), so
with a perfect predictor the proposal is locally balanced and the acceptance tends to 1.
The posterior is exact for any parameters: the acceptance carries the Hastings ratio of
the candidate-set softmax.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np

from . import _lib
from ._lib import check

HID = 256


class LearnedProposal:
    def __init__(self, sampler, seed=0, ncand=256, eps=0.1, lam=1.0, kappa=0.0, beta_max=None,
                 lift=False):
        if sampler._S is None:
            raise RuntimeError("call init_state on the sampler first")
        self.S = sampler
        self.ncand, self.eps, self.lam, self.kappa = int(ncand), float(eps), float(lam), float(kappa)
        self.beta_max = 1e300 if beta_max is None else float(beta_max)
        self.lift = bool(lift)
        n = [C.c_int(0) for _ in range(4)]
        check(_lib.load().chm_lp_info(sampler._S, *[C.byref(v) for v in n]))
        self.nparam, self.nocc, self.nfeat, self.kp = (int(v.value) for v in n)
        self.mode = 0
        self.rec_cap = 0
        self.set_params(self.initial_params(seed))
        # a weak hook: the sampler must not keep its proposal (and itself) alive in a cycle
        sampler._build_hooks.append(weakref.WeakMethod(self._reapply))

    # ------------------------------------------------------------ parameters
    def offsets(self):
        kp = self.kp
        o = {"W1": (0, kp * HID)}
        o["b1"] = (kp * HID, HID)
        o["W2"] = (o["b1"][0] + HID, HID * HID)
        o["b2"] = (o["W2"][0] + HID * HID, HID)
        o["w3"] = (o["b2"][0] + HID, HID)
        o["b3"] = (o["w3"][0] + HID, 1)
        return o

    def initial_params(self, seed=0):
        """W1 ~ N(0, 1/nocc) (each row activates nocc features), W2 ~ N(0, 1/256), w3 = 0
        and zero biases: the untrained network scores every candidate alike, so the
        proposal starts as a uniform draw from the candidate set."""
        rng = np.random.default_rng(seed)
        flat = np.zeros(self.nparam, np.float32)
        o = self.offsets()
        W1 = np.zeros((self.kp, HID), np.float32)
        W1[:self.nfeat] = rng.standard_normal((self.nfeat, HID)) / np.sqrt(self.nocc)
        flat[o["W1"][0]:o["W1"][0] + o["W1"][1]] = W1.ravel()
        flat[o["W2"][0]:o["W2"][0] + o["W2"][1]] = rng.standard_normal(HID * HID) / np.sqrt(HID)
        return flat

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        if flat.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters")
        check(_lib.load().chm_lp_set_params(self.S._S, flat.ctypes.data))
        self._params = flat.copy()

    def get_params(self):
        out = np.empty(self.nparam, np.float32)
        check(_lib.load().chm_lp_get_params(self.S._S, out.ctypes.data))
        return out

    # ------------------------------------------------------------ proposal mode
    def enable(self, on=True, ncand=None, eps=None, lam=None, kappa=None, beta_max=None, lift=None):
        if ncand is not None:
            self.ncand = int(ncand)
        if eps is not None:
            self.eps = float(eps)
        if lam is not None:
            self.lam = float(lam)
        if kappa is not None:
            self.kappa = float(kappa)
        if beta_max is not None:
            self.beta_max = float(beta_max)
        if lift is not None:
            self.lift = bool(lift)
        self.mode = 1 if on else 0
        check(_lib.load().chm_lp_configure(self.S._S, self.mode, self.ncand, self.eps, self.lam,
                                           self.kappa, self.beta_max, int(self.lift)))

    def set_tilt_weights(self, w):
        """Bond weights (nsp, nsp) of the tilt's count: dU = the change of
        sum_bonds w[s_i, s_j] over the first-shell bonds a swap changes (integers in
        [-63, 63], symmetric).  The default, w[s, t] = [s != t], counts unlike bonds."""
        w = np.ascontiguousarray(w, np.int32)
        self.tilt_w = w.copy()
        check(_lib.load().chm_lp_set_tilt_weights(self.S._S, w.ctypes.data, int(w.shape[0])))

    def disable(self):
        self.enable(False)

    # ------------------------------------------------------------ records / training
    def record(self, capacity):
        self.rec_cap = int(capacity)
        check(_lib.load().chm_lp_set_records(self.S._S, self.rec_cap))

    def records(self, rows=None):
        """(nvalid, occ (rows, nocc) uint8, y (rows), meta (rows)) of the ring."""
        rows = self.rec_cap if rows is None else int(rows)
        occ = np.empty((rows, self.nocc), np.uint8)
        y = np.empty(rows, np.float32)
        meta = np.empty(rows, np.int32)
        nv = C.c_int64(0)
        check(_lib.load().chm_lp_read_records(self.S._S, C.byref(nv), occ.ctypes.data,
                                              y.ctypes.data, meta.ctypes.data, rows))
        return int(nv.value), occ, y, meta

    def train(self, iters=1000, batch=512, lr=1e-3, max_rung=None, clip=30.0, seed=1):
        """Adam steps on minibatches of the record ring; returns the last minibatch loss."""
        if max_rung is None:
            max_rung = self.S.nrep - 1
        loss = C.c_float(0.0)
        check(_lib.load().chm_lp_train(self.S._S, int(iters), int(batch), float(lr), int(max_rung),
                                       float(clip), int(seed) & ((1 << 64) - 1), C.byref(loss)))
        self._params = self.get_params()
        return float(loss.value)

    def eval(self, occ):
        """Network output for occupancy rows (n, nocc) through the device scorer."""
        occ = np.ascontiguousarray(occ, np.uint8)
        out = np.empty(occ.shape[0], np.float32)
        check(_lib.load().chm_lp_eval(self.S._S, occ.shape[0], occ.ctypes.data, out.ctypes.data))
        return out

    def grad(self, occ, y, w):
        """(loss, flat gradient) of one batch through the device trainer (no update)."""
        occ = np.ascontiguousarray(occ, np.uint8)
        y = np.ascontiguousarray(y, np.float32)
        w = np.ascontiguousarray(w, np.float32)
        g = np.empty(self.nparam, np.float32)
        loss = np.zeros(1, np.float32)
        check(_lib.load().chm_lp_grad(self.S._S, occ.shape[0], occ.ctypes.data, y.ctypes.data,
                                      w.ctypes.data, g.ctypes.data, loss.ctypes.data))
        return float(loss[0]), g

    def profile(self, iters=100):
        """Device ms per launch (HIP events) of one step's scoring workgroups (every slot
        learned), launched alone (in a run they ride in the transform's launch).  An accepted
        swap is folded into the per-site sums inside the per-replica kernel, so there is no
        separate folding time."""
        ms = np.zeros(2)
        check(_lib.load().chm_lp_profile(self.S._S, int(iters), ms.ctypes.data))
        return {"score_ms": float(ms[0])}

    def site_sums(self):
        """The per-site layer-1 sums of every slot's committed configuration (nrep, 2, natoms,
        256) int32 in units of 2^-16, slot order (role 0: the site as candidate a, 1: as b)."""
        S = self.S
        return S._read(15, (S.nrep, 2, S.natoms, 256), np.int32)

    def _reapply(self, sampler):
        """A new device sampler (new composition): restore parameters, records, mode."""
        self.set_params(self._params)
        if self.rec_cap:
            self.record(self.rec_cap)
        if getattr(self, "tilt_w", None) is not None:
            self.set_tilt_weights(self.tilt_w)
        if self.mode:
            self.enable()
