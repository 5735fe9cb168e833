/*
 * chimera_hip.h — C ABI of the MI355X (gfx950) chi(k) forward model, windowed chi(R)
 * transform and parallel-tempering RMC swap sampler.
 *
 * The reference (dluccioni/CHIMERA) has no C ABI: its boundary is a Python object API
 * over CuPy arrays. Each entry point below names the reference interface it backs
 * (file:line in the reference tree). Conventions:
 *   - every function returns 0 on success, a negative CHM_E_* code on failure, and
 *     chm_last_error() then holds a message (thread-local);
 *   - all pointer arguments are HOST pointers unless the name says otherwise; the
 *     library owns every device allocation;
 *   - one HIP stream per context; a context (and everything created from it) is not
 *     thread-safe: call it from one host thread at a time;
 *   - floating point is IEEE double unless stated; species are int32 indices.
 */
#ifndef CHIMERA_HIP_H
#define CHIMERA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHM_OK 0
#define CHM_E_ARG (-1)     /* bad argument (maps to ValueError) */
#define CHM_E_HIP (-2)     /* HIP runtime failure */
#define CHM_E_STATE (-3)   /* call out of order (maps to RuntimeError) */
#define CHM_E_ASSERT (-4)  /* a reference assertion failed (maps to AssertionError) */

#define CHM_PREC_F64 0     /* FP64 path sum */
#define CHM_PREC_F32 1     /* FP32 path sum: FP32 terms on the hardware sine / exponential
                              units, summed in FP32 per run of one absorber species and in
                              FP64 across runs and workgroups */

typedef struct chm_ctx chm_ctx;
typedef struct chm_ft chm_ft;
typedef struct chm_sampler chm_sampler;
typedef struct chm_batch chm_batch;
typedef struct chm_psampler chm_psampler;

const char *chm_last_error(void);
int chm_version(void);

/* Context: device + stream.  Replaces the CuPy device/null stream the reference
 * uses implicitly (exafs_gpu/forward.py:20, methods/method_a_rmc.py:38). */
int chm_ctx_create(int device, chm_ctx **out);
int chm_ctx_destroy(chm_ctx *ctx);
int chm_ctx_sync(chm_ctx *ctx);

/* ---------------------------------------------------------------- forward model */

/* Scattering tables: k (nk), f (nsp, nk), delta (nsp*nsp, nk), lambda (nk), s02.
 * ForwardModel.__init__ table upload (exafs_gpu/forward.py:61-86); nsp >= 2 as
 * kernels.module (exafs_gpu/kernels.py:674-675). */
int chm_fm_set_tables(chm_ctx *ctx, int nsp, int nk, const double *k, const double *f,
                      const double *delta, const double *lam, double s02);

/* Fixed topology: ordered pairs i, j (npairs), periodic shift (npairs, 3), shell
 * (npairs), sigma^2 (npairs).  ForwardModel.__init__ (forward.py:65-92) and the
 * in-place d_shift / d_sig2 updates its callers make (chimera/model.py:263, 278). */
int chm_fm_set_topology(chm_ctx *ctx, int natoms, int npairs, const int32_t *i_idx,
                        const int32_t *j_idx, const double *shift, const int32_t *shell,
                        const double *sig2);
int chm_fm_set_shift(chm_ctx *ctx, const double *shift);
int chm_fm_set_sig2(chm_ctx *ctx, const double *sig2);

/* Cumulant tables: dr and s2 (nshell*nsp*nsp), c3 (nshell).
 * ForwardModel.set_dr / set_sig2_types / set_c3 (forward.py:104-145). */
int chm_fm_set_cumulants(chm_ctx *ctx, int nshell, const double *drtab, const double *c3tab,
                         const double *s2tab);

/* Pair distances r (nconf, npairs) and unit vectors evec (nconf, npairs, 3) from
 * positions (nconf, natoms, 3).  ForwardModel.distances (forward.py:148-157),
 * kernel compute_r (kernels.py:53-78).  evec may be NULL. */
int chm_fm_distances(chm_ctx *ctx, int nconf, const double *pos, double *r, double *evec);

/* Pair types s_i*nsp + s_j (nconf, npairs).  ForwardModel.ptypes (forward.py:159-166),
 * kernel build_ptype (kernels.py:81-96). */
int chm_fm_ptypes(chm_ctx *ctx, int nconf, const int32_t *species, int32_t *ptype);

/* Full single-scattering path sum, UNNORMALISED, chi (nconf, nsp, nk).
 * ForwardModel.chi (forward.py:169-188) / kernel chi_forward (kernels.py:99-163).
 * The k = 0 term is defined as 0.  precision: CHM_PREC_F64 or CHM_PREC_F32. */
int chm_fm_chi(chm_ctx *ctx, int nconf, const double *pos, const int32_t *species,
               double *chi, int precision);

/* The path sum of chm_fm_chi (same kernels, same result) on DEVICE pointers, allocating
 * nothing: d_pos (nconf, natoms, 3), d_species (nconf, natoms), d_chi (nconf, nsp, nk)
 * unnormalised, all device memory owned by the caller; enqueued on `stream` (a
 * hipStream_t, or NULL for the context's stream) and not waited for.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (work captured into
 * a graph, a serving loop).  Its intermediate arrays live in one static device array of
 * 32 MiB per device (held until the process ends); a call whose arrays do not fit
 * (about 8 MB per configuration at 216,000 pairs) fails with CHM_E_ARG before anything
 * is launched.  Species are not checked on the host: a configuration with a species
 * outside [0, nsp) gets NaN rows.  Calls that overlap on one device (two streams or two
 * threads) share that array and race. */
int chm_fm_chi_noalloc(chm_ctx *ctx, int nconf, const double *d_pos, const int32_t *d_species,
                       double *d_chi, int precision, void *stream);

/* Term table rows: row q gets
 *   s02 * prod_{m<nev} f[ev_m](k) / (k r^2) * exp(-2 r/lambda(k) - 2 k^2 s2)
 *       * sin(2 k r + mean_m delta[a*nsp+ev_m](k) - 4/3 k^3 c3)
 * scaled by amp, with the row's (r, s2, c3, a, nev, ev[3]).  It builds the count-space
 * basis B[a,b,s](k) (chimera/shellmodel.py:290-336) and the multiple-scattering
 * entry table (chimera/mslibrary.py:304-313).  out is (nrows, nk) host. */
int chm_fm_term_rows(chm_ctx *ctx, int nrows, const double *r, const double *s2,
                     const double *c3, const int32_t *absorber, const int32_t *nev,
                     const int32_t *ev /* (nrows, 3) */, double amp, double *out);

/* Term table rows from TABULATED per-row F, D, L (nrows, nk, on the model grid):
 *   amp * s02 * F / (k r^2) * exp(-2 r / max(L, 1e-6) - 2 k^2 s2) * sin(2 k r + D)
 * with the row's (r, s2).  The exact multiple-scattering table of a path library read
 * from FEFF output, PathLibrary.table (chimera/mslibrary.py:304-313).  out (nrows, nk). */
int chm_fm_tab_rows(chm_ctx *ctx, int nrows, const double *r, const double *s2,
                    const double *F, const double *D, const double *L, double amp, double *out);

/* Adjoint of the path sum: dL/dpos (nconf, natoms, 3) given dL/dchi (nconf, nsp, nk) of
 * the UNNORMALISED chi of chm_fm_chi at positions pos and species (the caller folds the
 * per-species normalisation into dL/dchi).  Per pair dT/dr_eff summed over k with the
 * absorber's dL/dchi (dr / C3 / sigma^2-by-type tables applied), then -g e_ij to the
 * absorber and +g e_ij to the scatterer, gathered per atom in a fixed order
 * (deterministic).  ForwardModel.grad_pos (forward.py:209-234), kernels chi_adjoint_r
 * and scatter_force (kernels.py:166-237). */
int chm_fm_grad_pos(chm_ctx *ctx, int nconf, const double *pos, const int32_t *species,
                    const double *gchi, double *gpos);

/* Relaxed-occupancy path sum, UNNORMALISED chi (nconf, nsp, nk): every pair contributes
 * sum_ab w_i[a] w_j[b] T_ab(r_ij, k) with per-site weights w (nconf, natoms, nsp); no
 * cumulant tables.  ForwardModel.chi_soft (forward.py:190-206), kernel chi_forward_soft
 * (kernels.py:514-589). */
int chm_fm_chi_soft(chm_ctx *ctx, int nconf, const double *pos, const double *w, double *chi);

/* Adjoint of chm_fm_chi_soft: dL/dpos (nconf, natoms, 3) and dL/dw (nconf, natoms, nsp)
 * given dL/dchi of the UNNORMALISED soft chi.  ForwardModel.grad_soft (forward.py:236-262;
 * the normalisation term is the caller's), kernel chi_adjoint_soft (kernels.py:592-660). */
int chm_fm_grad_soft(chm_ctx *ctx, int nconf, const double *pos, const double *w,
                     const double *gchi, double *gpos, double *gw);

/* Average device time (ms) of one launch sequence over `iters`, HIP events on the
 * context stream, on the inputs of the last call of that kind with the same nconf:
 * what 0 = chm_fm_chi FP64 (distances + types + path sum + partial reduction), 1 = the
 * same in FP32, 2 = the hard adjoint's dL/dr kernel, 3 = the soft path sum (+ reduction),
 * 4 = the soft adjoint's per-pair kernel.  No counterpart in the reference (it has no
 * profiler hooks). */
int chm_fm_profile(chm_ctx *ctx, int what, int nconf, int iters, double *ms);

/* ------------------------------------------------------------------ transform */

/* One windowed transform per channel, folded with the model-grid resample into a
 * DFT matrix D_c (nk_model, nR) complex on the device.
 * FourierTransform.__init__ (exafs_gpu/fourier.py:39-67) + MultiEdgeFT
 * (chimera/multift.py:34-56).  idx/frac/weight are (nchan, nfft): the resample
 * index, clamped fraction (already times `inside`) and W(k) k^w; kuni (nfft);
 * R (nR) any grid; norm = dk/sqrt(pi). */
int chm_ft_create(chm_ctx *ctx, int nchan, int nk_model, int nfft, int nR,
                  const int32_t *chan_edge, const int32_t *idx, const double *frac,
                  const double *inside, const double *weight, const double *kuni,
                  const double *R, double norm, chm_ft **out);
int chm_ft_destroy(chm_ft *ft);

/* chi (nrows, nsp, nk_model) -> out (nrows, nchan, nR, 2) [re, im], chi_extra
 * (nsp, nk_model) or NULL added first.  FourierTransform.to_R (fourier.py:75-80) fanned
 * out per channel by MultiEdgeFT.to_R (multift.py:58-63). */
int chm_ft_apply(chm_ft *ft, int nrows, const double *chi, const double *chi_extra,
                 double *out);

/* Rows (nrows, nk_model) through the channels of one edge e only ->
 * out (nrows, nkw_e, nR, 2).  Used to transform count-space basis rows. */
int chm_ft_apply_edge(chm_ft *ft, int edge, int nrows, const double *rows, double *out);

/* Linear resample of nrows chi rows (nrows, nk) from the model grid onto channel
 * `channel`'s uniform grid (nrows, nfft): FourierTransform.resample
 * (exafs_gpu/fourier.py:69-73), the same arithmetic as the reference, bit for bit. */
int chm_ft_resample(chm_ft *ft, int channel, int nrows, const double *chi, double *out);

/* -------------------------------------------------------------------- sampler */

/* Problem geometry of a count-space swap sampler (every pair and path term depends
 * only on the species at fixed lattice sites: the reference's swap-only use,
 * chimera/fitrun.py:307-315 with swap_frac = 1). */
typedef struct {
    int32_t natoms, nsp, ncoord, nshell;
    int32_t nblk;        /* multiple-scattering entries per absorber, 0 = none      */
    int32_t nstencil;    /* multiple-scattering incidence records per site           */
    int32_t nabsrec;     /* of which the first nabsrec have the site as absorber     */
    int32_t nkw, nwin;   /* channels per edge; fit-window R points                  */
    int32_t box[3];      /* supercell in units of a0/2                              */
    int32_t ncell[3];    /* conventional cells                                      */
    int32_t nbasis;      /* sites per conventional cell                             */
    int32_t basis_of_parity[8]; /* basis index of parity (x&1)|(y&1)<<1|(z&1)<<2   */
    int32_t has_const;   /* one constant row per edge (MultiEdgeFT.chi_extra)       */
} chm_geom;

/* Create a sampler.
 *   nbr_j, nbr_shell (natoms, ncoord): the i-sorted pair list of the forward model
 *       (methods/method_a_rmc.py:106-117: nbr_pairs = arange reshaped);
 *   stencil (nstencil, 16) int8: per incidence record [abs dx,dy,dz, slot0 dx,dy,dz,
 *       slot1 .., slot2 .., nslot, weight, cls_off lo, cls_off hi] relative to the site;
 *   basis_R (nsp, nrow_e, nkw*nwin*2): the transformed count-space basis per edge,
 *       already divided by the edge's absorber count (row order: SS rows b*nshell+s,
 *       then nblk MS entries, then the constant row);
 *   basis_k (nsp, nrow_e, nk) untransformed, for reading chi(k) back;
 *   target, sigma (nsp*nkw, nwin, 2) and (nsp*nkw, nwin): the fit window only.
 * RMCSampler.__init__ (methods/method_a_rmc.py:64-163). */
int chm_sampler_create(chm_ctx *ctx, const chm_geom *g, int nrep, int nk,
                       const int32_t *site_xyz, const int32_t *nbr_j,
                       const int32_t *nbr_shell, const int8_t *stencil,
                       const double *basis_R, const double *basis_k,
                       const double *target, const double *sigma, int profile_scale,
                       chm_sampler **out);
int chm_sampler_destroy(chm_sampler *s);

/* Start every replica from species0 (natoms); betas per rung (nrep); seed.
 * RMCSampler.init_state (method_a_rmc.py:166-184). */
int chm_sampler_init(chm_sampler *s, const int32_t *species0, const double *beta,
                     uint64_t seed);

/* Run nsteps swap steps (one proposal per replica each), replica exchange every
 * exchange_every steps (0 = never).  Device-resident; whole exchange periods are
 * replayed from a captured hipGraph.  Returns once the work is enqueued on the
 * sampler's stream (every read and every state-changing call waits for it; kernel
 * faults surface there or at chm_ctx_sync).  RMCSampler.run/step/replica_exchange
 * (method_a_rmc.py:250-308, 332-373, 401-455). */
int chm_sampler_run(chm_sampler *s, int nsteps, int exchange_every);

/* Record options: move log (a, b, accepted, rung per replica per step) and the SS
 * count vector of the rung nearest collect_beta each step (for alpha series / ESS).
 * Capacity in steps; 0 disables. */
int chm_sampler_set_recording(chm_sampler *s, int log_capacity, int series_capacity,
                              int series_rung);

/* Read state back (host buffers).  what:
 *   0 species (nrep, natoms) int32, rung order    1 chi2 (nrep) double, rung order
 *   2 counts (nrep, nsp*nrow_e) int32, rung order  3 best chi2 (nrep) double, rung order
 *   4 best species (nrep, natoms) int32            5 rung_of_slot (nrep) int32
 *   6 move log (n, nrep, 4) int32 [a, b, acc, rung], slot order, n = steps logged
 *   7 series (n, nsp*nsp*nshell) int32             8 exchange attempts/accepts (2, nrep-1) int64
 *   9 accepted moves per rung (nrep) int64         10 step counter (1) int64
 *   11 chi(k) (nrep, nsp, nk) double, rung order   12 trial chi(R) of the last step
 *      (nrep, nsp*nkw*nwin*2) double, slot order     13 running chi(R) accumulator, same
 *      layout (incremental mode)                      14 FP64 re-anchors done (nrep) int64,
 *      slot order                                     15 learned proposal's per-site
 *      layer-1 sums (nrep, 2, natoms, 256) int32 (units 2^-16), slot order, which lack
 *      the swap in 16 (nrep, 8) int32 [valid, a, b, old species of a, of b]
 * RMCSampler attributes species/chi2/chi/best (method_a_rmc.py:166-184, 493-496). */
int chm_sampler_read(chm_sampler *s, int what, void *host, int64_t nbytes);

/* Entry counts (nconf, nsp*nrow_e) of arbitrary configurations (nconf, natoms), from
 * scratch: SupercellMS.counts (chimera/mslibrary.py:436-451) and the pair counts
 * N[a,b,s] of ShellModel.counts_of (chimera/shellmodel.py:379-386) in one pass. */
int chm_sampler_counts(chm_sampler *s, int nconf, const int32_t *species, int32_t *counts);

/* Reduced chi^2 (with this sampler's target, sigma, window and profiling) of nrows
 * window transforms chiR (nrows, nsp*nkw*nwin*2).  RMCSampler._chi2
 * (method_a_rmc.py:211-227) as used for best_verified (:439-448). */
int chm_sampler_chi2(chm_sampler *s, int nrows, const double *chiR, double *c2);

/* Precision of the trial transform.  incremental = 0: chi(R)_trial = counts @ basis_R
 * in FP64 every step (exact, no drift).  incremental = 1: delta(R) = delta_counts @
 * basis_R in split-BF16 (3 planes) on MFMA with FP32 accumulation, added to an FP64
 * running chi(R) that is re-anchored (recomputed in FP64 from the integer counts) every
 * reanchor_every accepted moves of a replica (0 = never).  incremental = 2: the same with
 * delta(R) in FP32 FMA against an FP32 copy of basis_R.  Adding 16 to 1 or 2 rounds the
 * running chi(R) to FP32 after every update (an FP32 accumulator, for precision studies).
 * chi^2 is FP64 in every mode.
 * Takes effect at the next chm_sampler_init.  No counterpart in the reference (its
 * transform is an FP64 cuFFT of the full trial chi every step, fourier.py:75-80). */
int chm_sampler_set_precision(chm_sampler *s, int incremental, int reanchor_every);

/* One replica-exchange sweep now (method_a_rmc.py:332-373); run() also exchanges on
 * its own every exchange_every steps. */
int chm_sampler_exchange(chm_sampler *s);

/* Replace the temperature ladder (nrep betas, rung order, each finite and > 0) of a
 * running chain; every slot takes the beta of the rung it currently holds.  The ladder
 * retuning of RMCSampler.adapt (method_a_rmc.py:376-398) computes the new values on the
 * host, as the reference does, and hands them over here. */
int chm_sampler_set_beta(chm_sampler *s, const double *beta);

/* Unnormalised chi(k) (nconf, nsp, nk) of count vectors (nconf, nsp*nrow_e):
 * counts @ basis_k per edge (SupercellMS.chi_of, chimera/mslibrary.py:453-457, and the
 * pair-count model ShellModel.chi_k, chimera/shellmodel.py:411-413). */
int chm_sampler_chik(chm_sampler *s, int nconf, const int32_t *counts, double *chi);

/* Run-history snapshots without a host round trip.  reserve_marks sizes the buffer for
 * n marks; mark(idx) enqueues on the sampler's stream a snapshot of the smallest best
 * chi^2 over the replicas, the total of accepted moves and the device wall clock;
 * read_marks copies the first n marks (n, 3): best chi^2, accepted moves, seconds since
 * mark 0.  RMCSampler.run's `history` tuples (methods/method_a_rmc.py:430-435), which
 * the reference fills with a host sync at every log point. */
int chm_sampler_reserve_marks(chm_sampler *s, int n);
int chm_sampler_mark(chm_sampler *s, int idx);
int chm_sampler_read_marks(chm_sampler *s, int n, double *out);

/* Checkpoint: the whole chain state between two chm_sampler_run calls (species, counts,
 * chi^2, running chi(R), best state, rung <-> slot maps, ladder, step / exchange
 * counters, acceptance and re-anchor counts, step count and seed) as one host blob of
 * state_size bytes.  Importing it into a sampler of the same shape and precision (made
 * and initialised as the saving one was) continues the chain bit for bit; recording
 * restarts at the imported step.  The learned proposal's weights are saved separately
 * (chm_lp_get_params).  No counterpart in the reference (init_state accepts a warm
 * start, chimera/fitrun.py:317-320, but nothing saves a running chain). */
int chm_sampler_state_size(chm_sampler *s, int64_t *nbytes);
int chm_sampler_export(chm_sampler *s, void *host, int64_t nbytes);
int chm_sampler_import(chm_sampler *s, const void *host, int64_t nbytes);

/* Launch shape of the per-replica kernel: workgroups per replica slot for the
 * accept/propose/gather kernel (1..64).  Chosen per problem by the autotuner
 * (chimera_amd/autotune.py); the default is ~3 gather items per lane.  No counterpart
 * in the reference (its launch shapes are fixed, method_a_rmc.py:253-297). */
int chm_sampler_set_tuning(chm_sampler *s, int nsplit);
int chm_sampler_get_tuning(chm_sampler *s, int *nsplit);

/* Paired steps (default on): in the split-BF16 mode with the uniform proposal, each
 * transform + per-replica launch pair makes two Metropolis steps.  The transform computes
 * the trial of step n+1 and of step n+2 for both outcomes of step n+1 (pre-fetching,
 * Brockwell 2006); the per-replica launch takes both decisions in order and proposes the
 * next pair.  Decisions, counts, chi^2 and records are those of single steps.  Off: one
 * step per launch pair (also used with the learned proposal, move records, the FP64 and
 * FP32 modes and segments shorter than 3 steps).  No counterpart in the reference (its
 * steps are sequential kernel chains, method_a_rmc.py:253-297). */
int chm_sampler_set_pairs(chm_sampler *s, int on);

/* Average device time (ms) per launch of the step's kernels, from HIP events on the
 * sampler's stream over `iters` launches each: ms[0] = 0 (kept for layout), ms[1] trial
 * transform GEMM, ms[2] per-replica accept + propose + gather kernel (step minus
 * transform), ms[3] one whole step.  The timed steps are real steps of the chain
 * (iters + 1 moves per replica are made, without replica exchange or FP64 re-anchor);
 * the state stays consistent.  No counterpart in the reference (it has no profiler
 * hooks). */
int chm_sampler_profile(chm_sampler *s, int iters, double *ms);

/* ----------------------------------------------------- learned swap proposal
 * No counterpart in the reference (its swap proposal is uniform, exafs_gpu/kernels.py:
 * 362-398).  Per replica and step, ncand state-independent candidate pairs are scored
 * by an MLP (one-hot species occupancy of both sites' stencil neighbours and of the
 * sites: nocc = 2 ncoord + 2 entries, nfeat = nsp * nocc features -> 256 -> 256 -> 1,
 * SiLU, bf16 weights, FP32 accumulation) that predicts the swap's energy change
 * nfit * dchi2; logit = -lam * beta * g / 2 + kappa * dU (dU: change of the weighted
 * first-shell bond count); one candidate is drawn from the softmax and accepted with the
 * exact Hastings ratio (the reverse probability from the same candidate set scored at
 * the proposed state).  With probability eps a step uses the uniform proposal.  The
 * scorer's layer 1 is W1 rounded to bf16 and to a multiple of 2^-16 (|w| <= 2), summed in
 * int32 from per-site rows that the sampler keeps per replica (read back with
 * chm_sampler_read: 15 the sums, nrep x 2 x natoms x 256 int32 in slot order, 16 the swap
 * they lack, nrep x 8 ints: valid, a, b, old species of a, of b). */

/* Sizes: flat FP32 parameter count, occupancy entries, features, padded features kp.
 * Flat layout: W1 (kp, 256) feature-major | b1 (256) | W2 (256, 256) input-major |
 * b2 (256) | w3 (256) | b3 (1). */
int chm_lp_info(chm_sampler *s, int *nparam, int *nocc, int *nfeat, int *kp);
/* Network parameters (FP32 masters; the scorer uses them rounded to bf16); resets Adam. */
int chm_lp_set_params(chm_sampler *s, const float *flat);
int chm_lp_get_params(chm_sampler *s, float *flat);
/* mode 0: uniform proposals (the plain sampler); mode 1: learned proposal with the
 * uniform one mixed in at probability eps and used on every slot whose beta exceeds
 * beta_max (cold rungs).  ncand: multiple of 32 in 32..256.  lift = 1: a lifted
 * (direction-persistent) proposal: each slot carries d = +-1, the tilt is kappa * d * dU
 * instead of kappa * |dU|, the reverse probability is scored with -d, and d flips when a
 * learned step is not accepted (skew detailed balance; the posterior of the
 * configurations is unchanged). */
int chm_lp_configure(chm_sampler *s, int mode, int ncand, double eps, double lam, double kappa,
                     double beta_max, int lift);
/* Weights of the tilt's bond count: dU = sum over the first-shell bonds a swap changes of
 * w[s][t] (new pair) - w[s][t] (old pair); w is nsp x nsp, symmetric, integers in
 * [-63, 63].  The default is the unlike-bond indicator (w[s][t] = 1 for s != t). */
int chm_lp_set_tilt_weights(chm_sampler *s, const int32_t *w, int nsp);
/* Move records: a ring of `capacity` (occupancy, nfit * dchi2, rung | accepted << 16)
 * entries written by every step (null moves: meta -1).  0 disables. */
int chm_lp_set_records(chm_sampler *s, int capacity);
/* Copy the first `rows` ring entries: occ (rows, nocc) uint8, y (rows), meta (rows);
 * nvalid = entries written so far (capped at the capacity).  Any output may be NULL. */
int chm_lp_read_records(chm_sampler *s, int64_t *nvalid, uint8_t *occ, float *y, int32_t *meta,
                        int64_t rows);
/* Network output g of n occupancy rows (n, nocc) through the scorer's MFMA path. */
int chm_lp_eval(chm_sampler *s, int n, const uint8_t *occ, float *g);
/* Loss 0.5 * mean_w (g - y)^2 and its gradient (flat layout) over n rows, weights 0/1,
 * through the trainer's FP32 forward and backward kernels (no update). */
int chm_lp_grad(chm_sampler *s, int n, const uint8_t *occ, const float *y, const float *w,
                float *grad, float *loss);
/* iters Adam steps (lr; betas 0.9, 0.999) on minibatches of `batch` records drawn from
 * the ring with a counter-based stream (seed), records of rungs <= max_rung only, targets
 * clipped to [-clip, clip]; loss of the last minibatch. */
int chm_lp_train(chm_sampler *s, int iters, int batch, float lr, int max_rung, float clip,
                 uint64_t seed, float *loss);
/* Average device time (ms, HIP events on the sampler's stream) of one step's scoring
 * workgroups launched alone (ms[0]: every slot learned, the current state, the last
 * proposal as the pending swap) and of the folding workgroups of one recorded swap per
 * slot (ms[1]), over `iters` launches each.  The chain does not advance; the per-site
 * sums are rebuilt before the next run. */
int chm_lp_profile(chm_sampler *s, int iters, double *ms);

/* ------------------------------------------------- pair-space sampler (all moves) */

/* The reference sampler with every move it makes (methods/method_a_rmc.py:63-496):
 * swap and single-atom displacement moves (swap_frac, disp_sigma, max_disp), any start
 * positions, complex or |chi(R)| targets, the Boltzmann energy prior and multiple
 * scattering from the configuration's species.  State per replica: species, positions,
 * the distance and type of every pair of the context's topology, unnormalised chi(k),
 * chi^2, best state with positions.  One step: proposal + affected pairs
 * (exafs_gpu/kernels.py:362-488), Delta chi over them (kernels.py:243-325), trial
 * chi(R), profiled chi^2, Metropolis, commit (kernels.py:490-510), best retention.
 *
 * Create on a context whose tables and topology are set (pairs i-sorted, ncoord per
 * site):
 *   ideal (natoms, 3): the ideal sites the displacement clamp is measured from;
 *   rev (npairs): index of the reverse pair (method_a_rmc.py:41-50);
 *   Te (nsp, nk, nkw*nwin*2): per edge, the window transform of a unit chi(k) at each
 *       k (resample + window + DFT at the window's R points, Re/Im interleaved), NOT
 *       divided by the absorber count (init divides by the start's composition);
 *   Rx (nsp, nkw*nwin*2) or NULL: the same transform of MultiEdgeFT.chi_extra;
 *   target: complex (nsp*nkw, nwin, 2) or, with magnitude, |chi(R)| (nsp*nkw, nwin);
 *   sigma (nsp*nkw, nwin) > 0.  RMCSampler.__init__ (method_a_rmc.py:64-163). */
int chm_psampler_create(chm_ctx *ctx, int nrep, const double *ideal, const int32_t *rev, int nkw,
                        int nwin, const double *Te, const double *Rx, const double *target,
                        const double *sigma, int magnitude, int profile_scale,
                        chm_psampler **out);
int chm_psampler_destroy(chm_psampler *s);
/* Move parameters: swap_frac in [0, 1], disp_sigma >= 0, max_disp > 0
 * (method_a_rmc.py:64-66, the displacement step retuned by adapt, :376-398). */
int chm_psampler_set_moves(chm_psampler *s, double swap_frac, double disp_sigma, double max_disp);
/* Energy prior: etab (nshell, nsp*nsp) V/2 per directed pair, ebeta = 1/(kB T); NULL
 * clears it (method_a_rmc.py:147-157, 310-326). */
int chm_psampler_set_prior(chm_psampler *s, const double *etab, double ebeta);
/* Multiple scattering from the species along every 3- and 4-leg path of the lattice
 * (SupercellMS, chimera/mslibrary.py:341-492): the lattice geometry and stencil as for
 * chm_sampler_create (nblk > 0), and the entry table tms (nsp*nblk, nk). */
int chm_psampler_set_ms(chm_psampler *s, const chm_geom *g, const int32_t *site_xyz,
                        const int8_t *stencil, const double *tms);
/* Every replica starts at (species0 (natoms), pos0 (natoms, 3)); chi_ms (nsp, nk) is the
 * start's multiple-scattering chi (NULL: none); betas per rung; seed.
 * RMCSampler.init_state (method_a_rmc.py:166-184). */
int chm_psampler_init(chm_psampler *s, const int32_t *species0, const double *pos0,
                      const double *chi_ms, const double *beta, uint64_t seed);
/* nsteps steps, replica exchange every exchange_every steps (0 = never); exchange periods
 * replayed from a captured hipGraph; asynchronous (method_a_rmc.py:250-308, 332-373, 401-455). */
int chm_psampler_run(chm_psampler *s, int nsteps, int exchange_every);
/* One exchange sweep now (method_a_rmc.py:332-373); synchronous. */
int chm_psampler_exchange(chm_psampler *s);
/* New ladder (rung order); each slot takes the beta of the rung it holds (:376-398). */
int chm_psampler_set_beta(chm_psampler *s, const double *beta);
/* Keep the last log_steps steps' moves (a, b, is-swap | accepted << 1, slot per rung)
 * and trial chi^2 on the device; 0 disables. */
int chm_psampler_set_recording(chm_psampler *s, int log_steps);
/* Read state back, rung order unless noted.  what: 0 species (nrep, natoms) int32,
 * 1 chi2 (nrep), 2 best chi2, 3 best species, 4 positions (nrep, natoms, 3), 5 best
 * positions, 6 chi (nrep, nsp, nk), 7 pair distances (nrep, npairs), 8 pair types int32,
 * 9 accepted moves per rung int64, 10 exchange attempts and accepts (2, nrep-1) int64,
 * 11 completed steps (1) int64, 12 move log (log_steps, nrep, 4) int32, 13 logged trial
 * chi^2 (log_steps, nrep), 14 rung of each slot int32, 15 energy-prior change of each
 * replica's last proposal (eV; method_a_rmc.py:310-319).  nbytes must match. */
int chm_psampler_read(chm_psampler *s, int what, void *host, int64_t nbytes);
/* chi^2 of nconf unnormalised chi (nconf, nsp, nk) with the sampler's target, window and
 * profiling (method_a_rmc.py:211-227, best_verified :439-448). */
int chm_psampler_chi2(chm_psampler *s, int nconf, const double *chi, double *c2);
/* Run-history snapshots on the stream: smallest best chi^2, accepted-move total, device
 * clock (as chm_sampler_mark).  read_marks converts the clock to seconds since mark 0. */
int chm_psampler_reserve_marks(chm_psampler *s, int n);
int chm_psampler_mark(chm_psampler *s, int idx);
int chm_psampler_read_marks(chm_psampler *s, int n, double *out);

/* ------------------------------------------------------------------- batches
 * Many independent problems stepped together: every kernel of a step is one launch for
 * all of them (descriptor tables of problem x tile and problem x replica slot), and an
 * exchange period of every problem is one captured hipGraph, launched on ctx's stream.
 * The samplers must live on ctx's device (any of its contexts), be initialised, share one
 * precision mode and one step count, and use the uniform
 * proposal; each keeps its own state and launch split (chm_sampler_set_tuning) and can
 * be read as usual.  Configure recording before creating the batch.  No counterpart in
 * the reference (its fit driver loops over problems and seeds on the host,
 * chimera/fitrun.py:302-333). */
int chm_batch_create(chm_ctx *ctx, int n, chm_sampler *const *samplers, chm_batch **out);
int chm_batch_run(chm_batch *b, int nsteps, int exchange_every);
int chm_batch_destroy(chm_batch *b);
/* Device time (ms, HIP events) of the batched transform (ms[0]), the batched per-replica
 * kernel (ms[1]) and one whole batched step (ms[2]), over `iters` steps that are real
 * steps of every chain (no exchange, no re-anchor). */
int chm_batch_profile(chm_batch *b, int iters, double *ms);

/* Tile shape of the batched split-BF16 transform: the K blocks (32 basis rows each) every
 * one of a tile's four wavefronts holds in registers.  1 (every problem's rows must fit
 * 4 K blocks: a quarter of the operand registers, more tiles in flight per CU), 4 (any
 * problem up to 16 K blocks), or 0 (1 when it fits, else 4).  Chosen by the batch
 * autotuner (chimera_amd/batch.py) by measurement.  get_shape returns the largest K-block
 * count over the problems. */
int chm_batch_set_transform(chm_batch *b, int kb_per_wave);
int chm_batch_get_shape(chm_batch *b, int *max_nkb);

#ifdef __cplusplus
}
#endif
#endif /* CHIMERA_HIP_H */
