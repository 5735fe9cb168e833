// Host-side state of a sampler (shared by sampler.hip and learned.hip).
#pragma once

#include "common.h"
#include "sampler_dev.h"

// Learned swap proposal: network, per-step candidate scores, move records, trainer.
struct chm_lp {
    int mode = 0;            // 0: uniform proposals only; 1: learned proposal mixed with uniform
    int ncand = 256;         // candidates per replica and step (multiple of 32, <= 256)
    int nsplit = 8;          // commit / gather workgroups per slot of a learned step's launch
    double eps = 0.0;        // probability that a step uses the uniform proposal instead
    double lam = 1.0;        // logit = -lam * beta * g / 2 + kappa * |dU|
    double kappa = 0.0;
    double beta_max = 1e300;  // slots at a larger beta (colder) use the uniform proposal
    signed char wpair[64] = {};  // dU = sum of w[s][t] changes over first-shell bonds (8 x 8)
    int lift = 0;            // lifted: logit tilt kappa * d * dU, d = dir[slot] flips when a
    chm::DevBuf dir;         //   learned step is not accepted (2 x nrep ints, +-1: before /
                             //   after a step, alternating by step)
    int nocc = 0, noccp = 0; // occupancy entries per candidate (2 ncoord + 2), padded to 16
    int nfeat = 0, kp = 0;   // one-hot features nsp * nocc, padded to 32
    int nparam = 0;          // flat FP32 parameters
    int r2max = 0;           // largest squared stencil offset, units (a0/2)^2
    int pm_r = 0;            // half-side of the offset -> stencil position cube
    signed char fs[16] = {};     // first-shell stencil positions (n1 of them)
    int n1 = 0;
    bool have_params = false;
    chm::DevBuf params, adam_m, adam_v;          // FP32 masters and Adam moments
    chm::DevBuf w1, w2t, w2, w3q;                // bf16 copies (W2 in both orders); w3 rounded
    chm::DevBuf w1q;                             // scorer's layer 1: int32 in units of 2^-16
    chm::DevBuf posmap;                          // int8 stencil position of an offset, or -1
    // the per-site layer-1 sums (nrep x 2 x natoms x 256 int32, valid when cache_ok)
    chm::DevBuf cache;
    bool cache_ok = false;
    // speculative scoring: g / dU of the pending step's candidates (2 parities x nrep x 2 x
    // 256; [.][.][1] with the previous swap made) and their touched flags (2 x nrep x 256),
    // the outcome of the last step, per-tile partial sums of the log q ratio (nrep x 8 x 6)
    // and tile maxima of the next logits (nrep x 8 x 2)
    chm::DevBuf sg, sdu, st8, cs, lq, dmax;
    chm::DevBuf lpn;     // (nrep) u64: the step the next transform launch's scoring serves
    chm::DevBuf pend_occ;                        // (nrep, noccp) occupancy of each pending proposal
    chm::DevBuf rec_occ, rec_y, rec_meta;        // move-record ring
    int rec_cap = 0;
    long long rec_from = 0;                      // step count when recording started
    long long adam_t = 0;
    int tr_cap = 0;                              // training batch capacity
    chm::DevBuf tr_occ, tr_y, tr_w, tr_wsum, tr_Z1, tr_H1, tr_Z2, tr_H2, tr_g, tr_dZ2, tr_dZ1, grad,
        loss;
};

struct chm_sampler {
    chm_ctx *ctx = nullptr;
    chm_smp::Geo g{};
    int nrep = 0, nk = 0, nsplit = 1, profile = 1;
    int incremental = 0, reanchor_every = 512;   // 0 exact FP64, 1 split-BF16 delta, 2 FP32 delta
    int acc32 = 0;                               // running chi(R) rounded to FP32 (study only)
    uint64_t seed = 0;
    bool inited = false;
    long long hstep = 0;
    double nfit = 1.0;
    size_t lds_k1 = 0;
    chm::DevBuf site_xyz, nbr_j, nbr_sh, ss_sten, stencil, basisR, basisK, target, wcol, q0, Tb;
    chm::DevBuf nbr_st;   // (natoms, ncoord) neighbour sites in stencil order
    chm::DevBuf basisR32; // FP32 copy of basis_R (FP32-transform mode only)
    chm::DevBuf species, counts, dcounts, c2, best_c2, best_snap, best_sp, rung_of, slot_of, beta, beta_slot,
        stepc, iex,
        prop, mtrial, mcur, partials, anchor, nanchor, exch_att, exch_acc, acc_rung, log, series,
        tmp;
    int log_cap = 0, series_cap = 0, series_rung = -1;
    long long rec_base = 0;   // steps completed when recording (re)started
    hipGraph_t graph[2] = {nullptr, nullptr};     // exchange period starting at an even /
    hipGraphExec_t gexec[2] = {nullptr, nullptr}; // odd completed-step count
    int graph_steps = 0;
    chm_lp lp;
    // paired steps (bf16x3 mode, uniform proposals): two steps per transform + per-replica
    // launch; round k uses proposal / count-change rows (k & 1) and the chi^2 row (k & 1)
    int pairs = 1;
    int gr2 = 0;                 // largest squared gather offset, units (a0/2)^2
    size_t lds_pair = 0;
    chm::DevBuf pdc, pprop, pmt, ppart, pc2;
    chm::DevBuf draws;           // (nrep, 24) random numbers made by the transform's extra workgroup
    // run-history snapshots taken on the stream (no host sync): per mark the smallest best
    // chi^2, the accepted-move total and the device wall clock
    chm::DevBuf marks;
    int mark_cap = 0;
};

namespace chm_smp {

constexpr int kPropInts = 6;   // a, b, new a, new b, is-swap, log q ratio (float bits)

// per-step double buffers, indexed by step parity: the count-change rows and the
// proposal records of step n, and chi^2 after step n
inline int *drow(chm_sampler *S, long long n) {
    return S->dcounts.as<int>() + (size_t)(n & 1) * S->nrep * S->g.nrow_tot;
}
inline int *prow(chm_sampler *S, long long n) {
    return S->prop.as<int>() + (size_t)(n & 1) * S->nrep * kPropInts;
}
inline double *c2row(chm_sampler *S, long long n) {
    return S->c2.as<double>() + (size_t)(n & 1) * S->nrep;
}

// learned.hip: the per-site layer-1 sums (built when stale), a segment's first scoring
// launch after h steps, and the dynamic LDS of a scoring workgroup
int lp_ensure_cache(chm_sampler *S);
int lp_launch_tiles0(chm_sampler *S, long long h);
size_t lp_tile_lds(const chm_sampler *S);
int lp_prepare(chm_sampler *S);   // buffers for the current geometry (at create)
void drop_graph(chm_sampler *S);  // sampler.hip: forget the captured step graphs
// sampler.hip: lattice geometry from the host description (sites checked), and the check
// of multiple-scattering stencil records against it
int geo_from(const chm_geom *gm, const int32_t *site_xyz, Geo &g);
int check_stencil(const Geo &g, const int8_t *stencil);

}  // namespace chm_smp
