// Context, tables, topology and the pair-space path sum (single scattering) for gfx950.
//
// Path sum layout: a workgroup owns one configuration and a contiguous run of pairs;
// its lanes own the k > 0 points, one each (five full wavefronts for 320 points).
// Pair records are staged through LDS one per lane per round (one coalesced load,
// k-independent parts precomputed by the loading lane), the per-k tables f[b](k),
// delta[t](k) sit in LDS indexed [row][k] so a wavefront's 64 lanes read 64 consecutive
// words.  Every pair record read from HBM is reused for all nk points.  Each workgroup
// writes one partial chi (nsp, nk); a second kernel sums the partials in a fixed order,
// so results are deterministic.
#include <cmath>
#include <cstring>
#include <string>

#include "common.h"
#include "fp64math.h"

namespace chm {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

}  // namespace chm

extern "C" const char *chm_last_error(void) { return chm::g_err.c_str(); }
extern "C" int chm_version(void) { return 1; }

extern "C" int chm_ctx_create(int device, chm_ctx **out) {
    CHM_REQUIRE(out != nullptr, CHM_E_ARG, "chm_ctx_create: out is NULL");
    int ndev = 0;
    CHM_HIP(hipGetDeviceCount(&ndev));
    CHM_REQUIRE(device >= 0 && device < ndev, CHM_E_ARG, "chm_ctx_create: no such device");
    CHM_HIP(hipSetDevice(device));
    int rc = chm::scratch_init(device);   // noalloc.hip: once per device, device current
    if (rc) return rc;
    auto *c = new chm_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        chm::set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
        return CHM_E_HIP;
    }
    *out = c;
    return CHM_OK;
}

extern "C" int chm_ctx_destroy(chm_ctx *ctx) {
    if (!ctx) return CHM_OK;
    chm::ctx_release(ctx);
    return CHM_OK;
}

extern "C" int chm_ctx_sync(chm_ctx *ctx) {
    CHM_REQUIRE(ctx, CHM_E_ARG, "null context");
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

namespace {

int upload(chm_ctx *ctx, chm::DevBuf &b, const void *h, size_t n) {
    int rc = b.ensure(n);
    if (rc) return rc;
    if (n) CHM_HIP(hipMemcpyAsync(b.p, h, n, hipMemcpyHostToDevice, ctx->stream));
    return CHM_OK;
}

int download(chm_ctx *ctx, void *h, const chm::DevBuf &b, size_t n) {
    if (n) CHM_HIP(hipMemcpyAsync(h, b.p, n, hipMemcpyDeviceToHost, ctx->stream));
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

}  // namespace

extern "C" int chm_fm_set_tables(chm_ctx *ctx, int nsp, int nk, const double *k,
                                 const double *f, const double *delta, const double *lam,
                                 double s02) {
    CHM_REQUIRE(ctx, CHM_E_ARG, "null context");
    CHM_REQUIRE(nsp >= 2, CHM_E_ARG, "need at least two species");
    CHM_REQUIRE(nsp <= chm::kMaxNsp, CHM_E_ARG, "at most 6 species are compiled in");
    CHM_REQUIRE(nk >= 2 && k && f && delta && lam, CHM_E_ARG, "bad tables");
    CHM_HIP(hipSetDevice(ctx->device));
    ctx->nsp = nsp;
    ctx->nk = nk;
    ctx->s02 = s02;
    ctx->h_k.assign(k, k + nk);
    int rc;
    if ((rc = upload(ctx, ctx->k, k, sizeof(double) * nk))) return rc;
    if ((rc = upload(ctx, ctx->f, f, sizeof(double) * nsp * nk))) return rc;
    if ((rc = upload(ctx, ctx->delta, delta, sizeof(double) * nsp * nsp * nk))) return rc;
    if ((rc = upload(ctx, ctx->lam, lam, sizeof(double) * nk))) return rc;
    // FP32 copies for the FP32 path sum: f, and delta in revolutions (delta / 2 pi)
    {
        std::vector<float> f32((size_t)nsp * nk), d32((size_t)nsp * nsp * nk);
        for (size_t q = 0; q < f32.size(); ++q) f32[q] = (float)f[q];
        for (size_t q = 0; q < d32.size(); ++q) d32[q] = (float)(delta[q] / 6.28318530717958647692);
        if ((rc = upload(ctx, ctx->f32, f32.data(), sizeof(float) * f32.size()))) return rc;
        if ((rc = upload(ctx, ctx->drev32, d32.data(), sizeof(float) * d32.size()))) return rc;
        CHM_HIP(hipStreamSynchronize(ctx->stream));   // the host vectors go out of scope
    }
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

extern "C" int chm_fm_set_cumulants(chm_ctx *ctx, int nshell, const double *drtab,
                                    const double *c3tab, const double *s2tab) {
    CHM_REQUIRE(ctx && ctx->nsp > 0, CHM_E_STATE, "set tables first");
    CHM_REQUIRE(nshell >= 1 && drtab && c3tab && s2tab, CHM_E_ARG, "bad cumulant tables");
    CHM_REQUIRE(ctx->npairs == 0 || nshell == ctx->nshell, CHM_E_ARG,
                "cumulant tables must have one entry per shell of the topology");
    CHM_HIP(hipSetDevice(ctx->device));
    size_t n3 = sizeof(double) * nshell * ctx->nsp * ctx->nsp;
    int rc;
    if ((rc = upload(ctx, ctx->drtab, drtab, n3))) return rc;
    if ((rc = upload(ctx, ctx->s2tab, s2tab, n3))) return rc;
    if ((rc = upload(ctx, ctx->c3tab, c3tab, sizeof(double) * nshell))) return rc;
    ctx->nshell = nshell;
    ctx->c3_zero = true;
    for (int q = 0; q < nshell; ++q) ctx->c3_zero = ctx->c3_zero && c3tab[q] == 0.0;
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

extern "C" int chm_fm_set_topology(chm_ctx *ctx, int natoms, int npairs, const int32_t *i_idx,
                                   const int32_t *j_idx, const double *shift,
                                   const int32_t *shell, const double *sig2) {
    CHM_REQUIRE(ctx && ctx->nsp > 0, CHM_E_STATE, "set tables first");
    CHM_REQUIRE(natoms > 0 && npairs > 0 && i_idx && j_idx && shift && shell && sig2,
                CHM_E_ARG, "bad topology");
    int nshell = 0;
    for (int p = 0; p < npairs; ++p) {
        CHM_REQUIRE(i_idx[p] >= 0 && i_idx[p] < natoms && j_idx[p] >= 0 && j_idx[p] < natoms,
                    CHM_E_ARG, "pair index out of range");
        CHM_REQUIRE(shell[p] >= 0, CHM_E_ARG, "negative shell index");
        if (shell[p] + 1 > nshell) nshell = shell[p] + 1;
    }
    CHM_HIP(hipSetDevice(ctx->device));
    ctx->natoms = natoms;
    ctx->npairs = npairs;
    int rc;
    if ((rc = upload(ctx, ctx->i_idx, i_idx, sizeof(int32_t) * npairs))) return rc;
    if ((rc = upload(ctx, ctx->j_idx, j_idx, sizeof(int32_t) * npairs))) return rc;
    if ((rc = upload(ctx, ctx->shift, shift, sizeof(double) * 3 * npairs))) return rc;
    if ((rc = upload(ctx, ctx->shell, shell, sizeof(int32_t) * npairs))) return rc;
    if ((rc = upload(ctx, ctx->sig2, sig2, sizeof(double) * npairs))) return rc;
    {
        std::vector<int> ioff(natoms + 1, 0), joff(natoms + 1, 0), il(npairs), jl(npairs);
        for (int p = 0; p < npairs; ++p) {
            ++ioff[i_idx[p] + 1];
            ++joff[j_idx[p] + 1];
        }
        for (int a = 0; a < natoms; ++a) {
            ioff[a + 1] += ioff[a];
            joff[a + 1] += joff[a];
        }
        std::vector<int> ci(ioff.begin(), ioff.end() - 1), cj(joff.begin(), joff.end() - 1);
        for (int p = 0; p < npairs; ++p) {
            il[ci[i_idx[p]]++] = p;
            jl[cj[j_idx[p]]++] = p;
        }
        if ((rc = upload(ctx, ctx->csr_ioff, ioff.data(), sizeof(int) * (natoms + 1)))) return rc;
        if ((rc = upload(ctx, ctx->csr_joff, joff.data(), sizeof(int) * (natoms + 1)))) return rc;
        if ((rc = upload(ctx, ctx->csr_i, il.data(), sizeof(int) * npairs))) return rc;
        if ((rc = upload(ctx, ctx->csr_j, jl.data(), sizeof(int) * npairs))) return rc;
        CHM_HIP(hipStreamSynchronize(ctx->stream));   // the host vectors go out of scope
    }
    for (int &v : ctx->prof_nconf) v = 0;
    // cumulant tables default to zero, sized for this topology's shells
    std::vector<double> z3((size_t)nshell * ctx->nsp * ctx->nsp, 0.0), z1(nshell, 0.0);
    ctx->nshell = nshell;
    if ((rc = upload(ctx, ctx->drtab, z3.data(), sizeof(double) * z3.size()))) return rc;
    if ((rc = upload(ctx, ctx->s2tab, z3.data(), sizeof(double) * z3.size()))) return rc;
    if ((rc = upload(ctx, ctx->c3tab, z1.data(), sizeof(double) * z1.size()))) return rc;
    ctx->c3_zero = true;
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

extern "C" int chm_fm_set_shift(chm_ctx *ctx, const double *shift) {
    CHM_REQUIRE(ctx && ctx->npairs > 0 && shift, CHM_E_STATE, "set topology first");
    CHM_HIP(hipSetDevice(ctx->device));
    int rc = upload(ctx, ctx->shift, shift, sizeof(double) * 3 * ctx->npairs);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

extern "C" int chm_fm_set_sig2(chm_ctx *ctx, const double *sig2) {
    CHM_REQUIRE(ctx && ctx->npairs > 0 && sig2, CHM_E_STATE, "set topology first");
    CHM_HIP(hipSetDevice(ctx->device));
    int rc = upload(ctx, ctx->sig2, sig2, sizeof(double) * ctx->npairs);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

// ---------------------------------------------------------------- distances, types

__global__ void k_distances(const double *__restrict__ pos, const int *__restrict__ iidx,
                            const int *__restrict__ jidx, const double *__restrict__ shift,
                            double *__restrict__ r, double *__restrict__ evec, int npairs,
                            int natoms, int nconf) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)npairs * nconf) return;
    int ic = (int)(tid / npairs), p = (int)(tid % npairs);
    const double *pi = pos + ((long)ic * natoms + iidx[p]) * 3;
    const double *pj = pos + ((long)ic * natoms + jidx[p]) * 3;
    double dx = pj[0] + shift[3 * p] - pi[0];
    double dy = pj[1] + shift[3 * p + 1] - pi[1];
    double dz = pj[2] + shift[3 * p + 2] - pi[2];
    double rr = sqrt(dx * dx + dy * dy + dz * dz);
    r[tid] = rr;
    if (evec) {
        double inv = 1.0 / rr;
        evec[3 * tid] = dx * inv;
        evec[3 * tid + 1] = dy * inv;
        evec[3 * tid + 2] = dz * inv;
    }
}

__global__ void k_ptypes(const int *__restrict__ species, const int *__restrict__ iidx,
                         const int *__restrict__ jidx, int *__restrict__ pt, int npairs,
                         int natoms, int nconf, int nsp) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)npairs * nconf) return;
    int ic = (int)(tid / npairs), p = (int)(tid % npairs);
    pt[tid] = species[(long)ic * natoms + iidx[p]] * nsp + species[(long)ic * natoms + jidx[p]];
}

static int run_distances(chm_ctx *ctx, int nconf, bool want_evec) {
    long tot = (long)nconf * ctx->npairs;
    int rc = ctx->r.ensure(sizeof(double) * tot);
    if (rc) return rc;
    if (want_evec && (rc = ctx->evec.ensure(sizeof(double) * 3 * tot))) return rc;
    dim3 grid((unsigned)((tot + 255) / 256));
    k_distances<<<grid, 256, 0, ctx->stream>>>(ctx->pos.as<double>(), ctx->i_idx.as<int>(),
                                                ctx->j_idx.as<int>(), ctx->shift.as<double>(),
                                                ctx->r.as<double>(),
                                                want_evec ? ctx->evec.as<double>() : nullptr,
                                                ctx->npairs, ctx->natoms, nconf);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

static int check_species(const chm_ctx *ctx, int nconf, const int32_t *species) {
    long n = (long)nconf * ctx->natoms;
    for (long q = 0; q < n; ++q)
        CHM_REQUIRE(species[q] >= 0 && species[q] < ctx->nsp, CHM_E_ARG,
                    "species index out of range");
    return CHM_OK;
}

extern "C" int chm_fm_distances(chm_ctx *ctx, int nconf, const double *pos, double *r,
                                double *evec) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set topology first");
    CHM_REQUIRE(nconf > 0 && pos && r, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(ctx->device));
    int rc = upload(ctx, ctx->pos, pos, sizeof(double) * 3 * (size_t)nconf * ctx->natoms);
    if (rc) return rc;
    if ((rc = run_distances(ctx, nconf, evec != nullptr))) return rc;
    long tot = (long)nconf * ctx->npairs;
    if (evec) {
        CHM_HIP(hipMemcpyAsync(evec, ctx->evec.p, sizeof(double) * 3 * tot,
                               hipMemcpyDeviceToHost, ctx->stream));
    }
    return download(ctx, r, ctx->r, sizeof(double) * tot);
}

extern "C" int chm_fm_ptypes(chm_ctx *ctx, int nconf, const int32_t *species, int32_t *pt) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set topology first");
    CHM_REQUIRE(nconf > 0 && species && pt, CHM_E_ARG, "bad arguments");
    int rc = check_species(ctx, nconf, species);
    if (rc) return rc;
    CHM_HIP(hipSetDevice(ctx->device));
    if ((rc = upload(ctx, ctx->species, species, sizeof(int32_t) * (size_t)nconf * ctx->natoms)))
        return rc;
    long tot = (long)nconf * ctx->npairs;
    if ((rc = ctx->ptype.ensure(sizeof(int32_t) * tot))) return rc;
    k_ptypes<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, ctx->stream>>>(
        ctx->species.as<int>(), ctx->i_idx.as<int>(), ctx->j_idx.as<int>(), ctx->ptype.as<int>(),
        ctx->npairs, ctx->natoms, nconf, ctx->nsp);
    CHM_LAUNCH_CHECK();
    return download(ctx, pt, ctx->ptype, sizeof(int32_t) * tot);
}

// ------------------------------------------------------------------ path sum

namespace {

constexpr int kMaxK = 512;    // k points with k > 0 (one lane each, at most 8 waves)
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kTwoPi = 6.28318530717958647692;

// FP64 path sum.  Lane l of a workgroup of roundup(nkz, 64) lanes owns k index k0 + l
// (k0 = first k > 0; the k = 0 term is 0), so the 0..16 A^-1 grid at 0.05 runs five full
// wavefronts.  Each round stages blockDim pairs into LDS (one per lane, with its
// k-independent parts: r_eff, ln(1/r_eff^2), sigma^2, C3) and every lane then sums all
// of them at its k.  The phase is carried in revolutions, (k/pi) r + delta/2pi - C3 term,
// formed by two fused multiply-adds, and sin(2 pi x) / e^x are the range-specific
// polynomials of fp64math.h (the sine as cos2pi of the phase less 1/4, the 1/4 folded
// into the delta table); two pairs are evaluated per step so that their dependent
// polynomial chains interleave.  Pairs of one absorber are contiguous, so terms go into
// one running sum that moves to the species' sum when the absorber species changes (a
// uniform branch); the pair type is uniform across the wave, so table offsets are
// scalar.
struct PairRec {
    double r;      // r_eff = r + dr[shell][type]
    double lir2;   // ln(1 / r_eff^2), folded into the exponent
    double s2;     // sigma^2 + s2tab[shell][type]
    double c3;     // c3tab[shell]
    int t, pad;    // pair type
};

// f_b(k) e^(...) and sin(phase) of one term at this lane's k: (k/pi) r + (delta/2pi - 1/4)
// - C3 term is the phase less a quarter revolution, so the sine is cos2pi of it.
template <int NSP>
__device__ __forceinline__ void pair_term(const PairRec &rec, int t, const double *sf,
                                          const double *sd, int nk, int ikc, double kp,
                                          double c1, double c2, double c3k,
                                          const chm::PolyCoefs &pc, double &fe, double &sn) {
    const int sb = t % NSP;
    const double ph = fma(kp, rec.r, fma(-c3k, rec.c3, sd[t * nk + ikc]));
    fe = sf[sb * nk + ikc] * chm::exp_fin(fma(c1, rec.r, fma(c2, rec.s2, rec.lir2)), pc);
    sn = chm::cos2pi(ph, pc);
}

template <int NSP>
__global__ __launch_bounds__(512) void k_chi_pairs(
    const double *__restrict__ r_all, const int *__restrict__ pt_all,
    const int *__restrict__ shell, const double *__restrict__ sig2,
    const double *__restrict__ ftab, const double *__restrict__ dtab,
    const double *__restrict__ lamtab, const double *__restrict__ ktab,
    const double *__restrict__ drtab, const double *__restrict__ c3tab,
    const double *__restrict__ s2tab, double *__restrict__ partial, int npairs, int nk, int k0,
    int pairs_per_wg, int kspan, double s02, const chm::PolyCoefs pc) {
    extern __shared__ double smem[];
    double *sf = smem;                                       // f (NSP, kspan)
    double *sd = sf + NSP * kspan;                           // delta/2pi - 1/4 (NSP*NSP, kspan)
    PairRec *tile = reinterpret_cast<PairRec *>(sd + NSP * NSP * kspan);
    const int ic = blockIdx.y;
    const int w = blockIdx.x;
    const int nwg = gridDim.x;
    const int nthr = blockDim.x;
    constexpr double kInv2Pi = 1.0 / kTwoPi;
    // this workgroup's k columns kb .. kb + kc - 1: chunk blockIdx.z of kspan columns
    const int kb = k0 + (int)blockIdx.z * kspan;
    const int kc = min(kspan, nk - kb);
    for (int q = threadIdx.x; q < NSP * kc; q += nthr) {
        const int row = q / kc, c = q - row * kc;
        sf[row * kspan + c] = ftab[row * nk + kb + c];
    }
    for (int q = threadIdx.x; q < NSP * NSP * kc; q += nthr) {
        const int row = q / kc, c = q - row * kc;
        sd[row * kspan + c] = fma(dtab[row * nk + kb + c], kInv2Pi, -0.25);
    }

    const int ik = kb + (int)threadIdx.x;
    const bool live = (int)threadIdx.x < kc;
    const int ikc = live ? (int)threadIdx.x : kc - 1;      // LDS column (valid for dead lanes)
    const double kv = live ? ktab[ik] : 0.0;
    const double lv = live ? lamtab[ik] : 1.0;
    const double kp = kv * kInvPi;                          // revolutions of 2 k r per unit r
    const double c1 = -2.0 / lv;
    const double c2 = -2.0 * kv * kv;
    const double c3k = (4.0 / 3.0) * kv * kv * kv * kInv2Pi;
    const double pref = (live && kv > 0.0) ? s02 / kv : 0.0;
    double acc[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) acc[s] = 0.0;

    const double *rc = r_all + (long)ic * npairs;
    const int *tc = pt_all + (long)ic * npairs;
    const int pbeg = w * pairs_per_wg;
    const int pend = min(npairs, pbeg + pairs_per_wg);

    for (int p0 = pbeg; p0 < pend; p0 += nthr) {
        __syncthreads();
        {
            const int p = p0 + (int)threadIdx.x;
            PairRec rec;
            if (p < pend) {
                const int t = tc[p];
                const int sh = shell[p];
                const double re = rc[p] + drtab[sh * NSP * NSP + t];
                rec.r = re;
                rec.lir2 = -2.0 * log(re);
                rec.s2 = sig2[p] + s2tab[sh * NSP * NSP + t];
                rec.c3 = c3tab[sh];
                rec.t = t;
            } else {
                rec.r = 1.0;
                rec.lir2 = -2000.0;   // e^-2000 = 0: contributes exactly 0
                rec.s2 = 0.0;
                rec.c3 = 0.0;
                rec.t = 0;
            }
            rec.pad = 0;
            tile[threadIdx.x] = rec;
        }
        __syncthreads();
        const int nt = __builtin_amdgcn_readfirstlane(min(nthr, pend - p0));
        double run = 0.0;
        int run_sa = __builtin_amdgcn_readfirstlane(tile[0].t) / NSP;
        auto add = [&](int sa, double fe, double sn) {
            if (sa != run_sa) {
#pragma unroll
                for (int s = 0; s < NSP; ++s)
                    if (run_sa == s) acc[s] += run;
                run = 0.0;
                run_sa = sa;
            }
            run = fma(fe, sn, run);
        };
        // two pairs per step: their two polynomial chains interleave
        int q = 0;
        for (; q + 1 < nt; q += 2) {
            const PairRec ra = tile[q], rb = tile[q + 1];
            const int ta = __builtin_amdgcn_readfirstlane(ra.t);
            const int tb = __builtin_amdgcn_readfirstlane(rb.t);
            double fea, sna, feb, snb;
            pair_term<NSP>(ra, ta, sf, sd, kspan, ikc, kp, c1, c2, c3k, pc, fea, sna);
            pair_term<NSP>(rb, tb, sf, sd, kspan, ikc, kp, c1, c2, c3k, pc, feb, snb);
            if (ta / NSP == run_sa && tb / NSP == run_sa) {    // both terms feed the run:
                run = fma(fea, sna, run);                      // no branch between the two
                run = fma(feb, snb, run);                      // evaluations
            } else {
                add(ta / NSP, fea, sna);
                add(tb / NSP, feb, snb);
            }
        }
        if (q < nt) {
            const PairRec ra = tile[q];
            const int ta = __builtin_amdgcn_readfirstlane(ra.t);
            double fea, sna;
            pair_term<NSP>(ra, ta, sf, sd, kspan, ikc, kp, c1, c2, c3k, pc, fea, sna);
            add(ta / NSP, fea, sna);
        }
#pragma unroll
        for (int s = 0; s < NSP; ++s)
            if (run_sa == s) acc[s] += run;
    }
    double *out = partial + ((long)ic * nwg + w) * NSP * nk;
    if (blockIdx.z == 0)
        for (int q = threadIdx.x; q < k0; q += nthr)
#pragma unroll
            for (int s = 0; s < NSP; ++s) out[s * nk + q] = 0.0;
    if (live) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) out[s * nk + ik] = pref * acc[s];
    }
}

// FP32 path sum.  Same layout as k_chi_pairs (lanes own k, pair tiles in LDS), with
// every per-term operation in FP32 on the hardware units: the phase in revolutions,
// sin on v_sin_f32 (sin(2 pi x)), the exponential on v_exp_f32 (2^x).  The phase
// k r / pi is formed exactly enough without FP64: k/pi and r are split into an 11-bit
// head and a tail, head x head is exact in FP32 and its integer part is dropped exactly,
// the three small cross products carry the rest (error ~1e-7 revolution at k r ~ 90).
// Terms are summed in FP32 over each run of one absorber species within a staging round,
// and the run sums in FP64.
struct PairRecF {
    float rhi, rlo;   // r_eff = rhi + rlo, rhi with an 11-bit significand
    float r;          // r_eff
    float l2ir2;      // log2(1 / r_eff^2), folded into the exponential
    float ea;         // -2 log2(e) sigma^2 (times k^2 gives the Debye-Waller exponent)
    float c3;         // c3tab[shell]
    int t, pad;
};

__device__ __forceinline__ float head11(float x) {
    return __int_as_float(__float_as_int(x) & (int)0xFFFFE000);
}

// One lane per k point with k > 0 (the k = 0 term is 0): the workgroup has
// roundup(nkz, 64) lanes (320 for the 0..16 A^-1 grid at 0.05: five full wavefronts),
// lane l owns k index k0 + l, and a tile stages blockDim pairs.
template <int NSP>
__global__ __launch_bounds__(512) void k_chi_pairs_f32(
    const double *__restrict__ r_all, const int *__restrict__ pt_all,
    const int *__restrict__ shell, const double *__restrict__ sig2,
    const float *__restrict__ ftab32, const float *__restrict__ drev32,
    const double *__restrict__ lamtab, const double *__restrict__ ktab,
    const double *__restrict__ drtab, const double *__restrict__ c3tab,
    const double *__restrict__ s2tab, double *__restrict__ partial, int npairs, int nk, int k0,
    int pairs_per_wg, int kspan, double s02) {
    extern __shared__ double smem[];
    float *sf = reinterpret_cast<float *>(smem);            // f (NSP, kspan)
    float *sd = sf + NSP * kspan;                           // delta / 2 pi (NSP*NSP, kspan)
    const int words = (NSP + NSP * NSP) * kspan;
    PairRecF *tile = reinterpret_cast<PairRecF *>(sf + ((words + 7) & ~7));
    const int ic = blockIdx.y;
    const int w = blockIdx.x;
    const int nwg = gridDim.x;
    const int nthr = blockDim.x;
    // this workgroup's k columns kb .. kb + kc - 1: chunk blockIdx.z of kspan columns
    const int kb = k0 + (int)blockIdx.z * kspan;
    const int kc = min(kspan, nk - kb);
    for (int q = threadIdx.x; q < NSP * kc; q += nthr) {
        const int row = q / kc, c = q - row * kc;
        sf[row * kspan + c] = ftab32[row * nk + kb + c];
    }
    for (int q = threadIdx.x; q < NSP * NSP * kc; q += nthr) {
        const int row = q / kc, c = q - row * kc;
        sd[row * kspan + c] = drev32[row * nk + kb + c];
    }

    constexpr double kLog2e = 1.44269504088896340736;
    const int ik = kb + (int)threadIdx.x;
    const bool live = (int)threadIdx.x < kc;
    const int ikc = live ? (int)threadIdx.x : kc - 1;      // LDS column (valid for dead lanes)
    const double kv = live ? ktab[ik] : 0.0;
    const double lv = live ? lamtab[ik] : 1.0;
    const double kp = kv * kInvPi;                          // revolutions of 2 k r per unit r
    const float khi = head11((float)kp);
    const float klo = (float)(kp - (double)khi);
    const float c1 = (float)(-2.0 * kLog2e / lv);
    const float k2 = (float)(kv * kv);
    const float c3k = (float)((4.0 / 3.0) * kv * kv * kv / kTwoPi);
    const double pref = (live && kv > 0.0) ? s02 / kv : 0.0;  // applied to the FP64 sums
    double acc[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) acc[s] = 0.0;

    const double *rc = r_all + (long)ic * npairs;
    const int *tc = pt_all + (long)ic * npairs;
    const int pbeg = w * pairs_per_wg;
    const int pend = min(npairs, pbeg + pairs_per_wg);

    for (int p0 = pbeg; p0 < pend; p0 += nthr) {
        __syncthreads();
        {
            const int p = p0 + (int)threadIdx.x;
            PairRecF rec;
            if (p < pend) {
                const int t = tc[p];
                const int sh = shell[p];
                const double re = rc[p] + drtab[sh * NSP * NSP + t];
                rec.r = (float)re;
                rec.rhi = head11(rec.r);
                rec.rlo = (float)(re - (double)rec.rhi);
                rec.l2ir2 = (float)(-2.0 * log2(re));
                rec.ea = (float)(-2.0 * kLog2e * (sig2[p] + s2tab[sh * NSP * NSP + t]));
                rec.c3 = (float)c3tab[sh];
                rec.t = t;
            } else {
                rec.r = rec.rhi = 1.f;
                rec.rlo = 0.f;
                rec.l2ir2 = -1000.f;   // 2^-1000 = 0: contributes exactly 0
                rec.ea = 0.f;
                rec.c3 = 0.f;
                rec.t = 0;
            }
            rec.pad = 0;
            tile[threadIdx.x] = rec;
        }
        __syncthreads();
        const int nt = __builtin_amdgcn_readfirstlane(min(nthr, pend - p0));
        // pairs are absorber-sorted in runs (every pair of atom i has i's species), so
        // terms go into one running sum, moved to the species' sum when the absorber
        // species changes (a uniform branch).  The pair type is uniform: its table
        // offsets are scalar arithmetic.
        float run = 0.f;
        int run_sa = __builtin_amdgcn_readfirstlane(tile[0].t) / NSP;
        for (int q = 0; q < nt; ++q) {
            const PairRecF rec = tile[q];
            const int t = __builtin_amdgcn_readfirstlane(rec.t);
            const int sa = t / NSP;
            if (sa != run_sa) {
#pragma unroll
                for (int s = 0; s < NSP; ++s)
                    if (run_sa == s) acc[s] += (double)run;
                run = 0.f;
                run_sa = sa;
            }
            const int sb = t - sa * NSP;
            const float x = khi * rec.rhi;                      // exact
            float ph = x - rintf(x);
            ph += fmaf(khi, rec.rlo, klo * rec.r);
            ph += sd[t * kspan + ikc] - c3k * rec.c3;
            const float sn = __builtin_amdgcn_sinf(ph);
            const float e = __builtin_amdgcn_exp2f(fmaf(c1, rec.r, fmaf(k2, rec.ea, rec.l2ir2)));
            run = fmaf(sf[sb * kspan + ikc] * e, sn, run);
        }
#pragma unroll
        for (int s = 0; s < NSP; ++s)
            if (run_sa == s) acc[s] += (double)run;
    }
    double *out = partial + ((long)ic * nwg + w) * NSP * nk;
    // columns below k0 (k <= 0) are zero, written by the first chunk
    if (blockIdx.z == 0)
        for (int q = threadIdx.x; q < k0; q += nthr)
#pragma unroll
            for (int s = 0; s < NSP; ++s) out[s * nk + q] = 0.0;
    if (live) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) out[s * nk + ik] = pref * acc[s];
    }
}

// chi[ic][q] = sum over workgroups w of partial[ic][w][q], q < n_per_conf, in a fixed
// order: a 1024-lane workgroup takes 64 consecutive q (one per lane, coalesced rows) of one
// configuration; wave v sums w = v, v + 16, ... in turn and the 16 wave sums are added in
// wave order.  (One lane per q walking every w serially left a handful of CUs busy for
// ~0.2 ms at one configuration.)
constexpr int kRedWaves = 16;
__global__ __launch_bounds__(1024) void k_reduce_partials(const double *__restrict__ partial,
                                                          double *__restrict__ chi, int nwg,
                                                          int n_per_conf, int nconf) {
    __shared__ double red[kRedWaves][64];
    const int lane = threadIdx.x & 63, v = threadIdx.x >> 6;
    const int q = blockIdx.x * 64 + lane, ic = blockIdx.y;
    double s = 0.0;
    if (q < n_per_conf) {
        const double *p = partial + (long)ic * nwg * n_per_conf + q;
        for (int w = v; w < nwg; w += kRedWaves) s += p[(long)w * n_per_conf];
    }
    red[v][lane] = s;
    __syncthreads();
    if (v == 0 && q < n_per_conf) {
        double t = red[0][lane];
#pragma unroll
        for (int z = 1; z < kRedWaves; ++z) t += red[z][lane];
        chi[(long)ic * n_per_conf + q] = t;
    }
}

int launch_reduce(hipStream_t st, const double *partial, double *out, int nwg, int n_per_conf,
                  int nconf) {
    const dim3 grid((unsigned)((n_per_conf + 63) / 64), (unsigned)nconf);
    k_reduce_partials<<<grid, 64 * kRedWaves, 0, st>>>(partial, out, nwg, n_per_conf, nconf);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

}  // namespace

namespace chm {
// Launch shape of the path sum for nconf configurations: lanes over the k > 0 points only
// (the grid is ascending), ~2048 workgroups over the device, whole staging rounds each.
int chi_shape(const chm_ctx *ctx, int nconf, ChiShape &sh) {
    int k0 = 0;
    while (k0 < ctx->nk && !(ctx->h_k[k0] > 0.0)) ++k0;
    CHM_REQUIRE(k0 < ctx->nk, CHM_E_ARG, "no k point above 0");
    const int nkz = ctx->nk - k0;
    // k chunks (grid z): at most kMaxK lanes each, and the chunk's FP64 table columns plus
    // one pair tile within 160 KB of LDS (six species at 512 points take two chunks)
    const size_t nsp = (size_t)ctx->nsp;
    int nkc = (nkz + kMaxK - 1) / kMaxK;
    for (;; ++nkc) {
        const int span = (nkz + nkc - 1) / nkc;
        const int thr = (span + 63) / 64 * 64;
        if (sizeof(double) * (nsp + nsp * nsp) * (size_t)span + sizeof(PairRec) * thr <= 160 * 1024) break;
    }
    sh.k0 = k0;
    sh.nkc = nkc;
    sh.kspan = (nkz + nkc - 1) / nkc;
    sh.nthr = (sh.kspan + 63) / 64 * 64;
    const int rounds = (ctx->npairs + sh.nthr - 1) / sh.nthr;
    int nwg = std::max(1, std::min(rounds, (2048 + nconf - 1) / nconf));
    sh.ppw = ((rounds + nwg - 1) / nwg) * sh.nthr;
    sh.nwg = (ctx->npairs + sh.ppw - 1) / sh.ppw;
    return CHM_OK;
}
}  // namespace chm

namespace {

// The path sum of nconf configurations from their distances r and pair types pt (both
// (nconf, npairs)) into chi (nconf, NSP, nk), with partial (nconf, nwg, NSP, nk) as the
// per-workgroup sums, on stream st.
template <int NSP>
int launch_chi_into(chm_ctx *ctx, hipStream_t st, int nconf, int precision, const chm::ChiShape &sh,
                    const double *r, const int *pt, double *partial, double *chi) {
    const int nk = ctx->nk;
    const dim3 grid(sh.nwg, nconf, sh.nkc);
    if (precision == CHM_PREC_F32) {
        const size_t words = (size_t)(NSP + NSP * NSP) * sh.kspan;
        const size_t lds = sizeof(float) * ((words + 7) & ~(size_t)7) + sizeof(PairRecF) * sh.nthr;
        CHM_REQUIRE(lds <= 160 * 1024, CHM_E_ARG, "tables too large for LDS");
        hipFuncSetAttribute((const void *)k_chi_pairs_f32<NSP>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        k_chi_pairs_f32<NSP><<<grid, sh.nthr, lds, st>>>(
            r, pt, ctx->shell.as<int>(), ctx->sig2.as<double>(), ctx->f32.as<float>(),
            ctx->drev32.as<float>(), ctx->lam.as<double>(), ctx->k.as<double>(),
            ctx->drtab.as<double>(), ctx->c3tab.as<double>(), ctx->s2tab.as<double>(), partial,
            ctx->npairs, nk, sh.k0, sh.ppw, sh.kspan, ctx->s02);
    } else {
        const size_t lds = sizeof(double) * (NSP + NSP * NSP) * sh.kspan + sizeof(PairRec) * sh.nthr;
        CHM_REQUIRE(lds <= 160 * 1024, CHM_E_ARG, "tables too large for LDS");
        hipFuncSetAttribute((const void *)k_chi_pairs<NSP>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        k_chi_pairs<NSP><<<grid, sh.nthr, lds, st>>>(
            r, pt, ctx->shell.as<int>(), ctx->sig2.as<double>(), ctx->f.as<double>(),
            ctx->delta.as<double>(), ctx->lam.as<double>(), ctx->k.as<double>(),
            ctx->drtab.as<double>(), ctx->c3tab.as<double>(), ctx->s2tab.as<double>(), partial,
            ctx->npairs, nk, sh.k0, sh.ppw, sh.kspan, ctx->s02, chm::kPolyCoefs);
    }
    CHM_LAUNCH_CHECK();
    return launch_reduce(st, partial, chi, sh.nwg, NSP * nk, nconf);
}

}  // namespace

namespace chm {
int chi_into(chm_ctx *ctx, hipStream_t st, int nconf, int precision, const ChiShape &sh,
             const double *r, const int *pt, double *partial, double *chi) {
    switch (ctx->nsp) {
        case 2: return launch_chi_into<2>(ctx, st, nconf, precision, sh, r, pt, partial, chi);
        case 3: return launch_chi_into<3>(ctx, st, nconf, precision, sh, r, pt, partial, chi);
        case 4: return launch_chi_into<4>(ctx, st, nconf, precision, sh, r, pt, partial, chi);
        case 5: return launch_chi_into<5>(ctx, st, nconf, precision, sh, r, pt, partial, chi);
        case 6: return launch_chi_into<6>(ctx, st, nconf, precision, sh, r, pt, partial, chi);
    }
    chm::set_error("unsupported species count");
    return CHM_E_ARG;
}

// Distances of nconf configurations (positions pos, device) into r (device), on stream st.
int distances_into(chm_ctx *ctx, hipStream_t st, int nconf, const double *pos, double *r) {
    const long tot = (long)nconf * ctx->npairs;
    k_distances<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, st>>>(
        pos, ctx->i_idx.as<int>(), ctx->j_idx.as<int>(), ctx->shift.as<double>(), r, nullptr,
        ctx->npairs, ctx->natoms, nconf);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int distances_device(chm_ctx *ctx, int nconf, bool want_evec) {
    return run_distances(ctx, nconf, want_evec);
}

// Pair types of the context's species buffer (nconf configurations) into ctx->ptype.
int ptypes_device(chm_ctx *ctx, int nconf) {
    long tot = (long)nconf * ctx->npairs;
    int rc = ctx->ptype.ensure(sizeof(int32_t) * tot);
    if (rc) return rc;
    k_ptypes<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, ctx->stream>>>(
        ctx->species.as<int>(), ctx->i_idx.as<int>(), ctx->j_idx.as<int>(), ctx->ptype.as<int>(),
        ctx->npairs, ctx->natoms, nconf, ctx->nsp);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int reduce_partials(chm_ctx *ctx, const double *partial, double *out, int nwg, int n_per_conf,
                    int nconf) {
    return launch_reduce(ctx->stream, partial, out, nwg, n_per_conf, nconf);
}

// Path sum over the context's pos/species buffers already on the device.
int chi_device(chm_ctx *ctx, int nconf, int precision) {
    int rc = run_distances(ctx, nconf, false);
    if (rc) return rc;
    if ((rc = ptypes_device(ctx, nconf))) return rc;
    ChiShape sh;
    if ((rc = chi_shape(ctx, nconf, sh))) return rc;
    if ((rc = ctx->partial.ensure(sizeof(double) * (size_t)nconf * sh.nwg * ctx->nsp * ctx->nk))) return rc;
    if ((rc = ctx->chi.ensure(sizeof(double) * (size_t)nconf * ctx->nsp * ctx->nk))) return rc;
    return chi_into(ctx, ctx->stream, nconf, precision, sh, ctx->r.as<double>(), ctx->ptype.as<int>(),
                    ctx->partial.as<double>(), ctx->chi.as<double>());
}
}  // namespace chm

extern "C" int chm_fm_chi(chm_ctx *ctx, int nconf, const double *pos, const int32_t *species,
                          double *chi, int precision) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(nconf > 0 && pos && species && chi, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(precision == CHM_PREC_F64 || precision == CHM_PREC_F32, CHM_E_ARG,
                "precision must be CHM_PREC_F64 or CHM_PREC_F32");
    int rc = check_species(ctx, nconf, species);
    if (rc) return rc;
    CHM_HIP(hipSetDevice(ctx->device));
    if ((rc = upload(ctx, ctx->pos, pos, sizeof(double) * 3 * (size_t)nconf * ctx->natoms)))
        return rc;
    if ((rc = upload(ctx, ctx->species, species, sizeof(int32_t) * (size_t)nconf * ctx->natoms)))
        return rc;
    if ((rc = chm::chi_device(ctx, nconf, precision))) return rc;
    ctx->prof_nconf[0] = ctx->prof_nconf[1] = nconf;
    return download(ctx, chi, ctx->chi, sizeof(double) * (size_t)nconf * ctx->nsp * ctx->nk);
}

// ------------------------------------------------------------------ term rows

__global__ void k_term_rows(const double *__restrict__ r, const double *__restrict__ s2,
                            const double *__restrict__ c3, const int *__restrict__ absorber,
                            const int *__restrict__ nev, const int *__restrict__ ev,
                            const double *__restrict__ ftab, const double *__restrict__ dtab,
                            const double *__restrict__ lamtab, const double *__restrict__ ktab,
                            double *__restrict__ out, int nrows, int nk, int nsp, double s02,
                            double amp) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)nrows * nk) return;
    int q = (int)(tid / nk), ik = (int)(tid % nk);
    double kv = ktab[ik];
    if (!(kv > 0.0)) {
        out[tid] = 0.0;
        return;
    }
    int a = absorber[q], n = nev[q];
    double F = 1.0, D = 0.0;
    for (int m = 0; m < n; ++m) {
        int b = ev[3 * q + m];
        F *= ftab[b * nk + ik];
        D += dtab[(a * nsp + b) * nk + ik];
    }
    D /= (double)n;
    double rr = r[q];
    double v = amp * s02 * F / (kv * rr * rr) *
               exp(-2.0 * rr / lamtab[ik] - 2.0 * kv * kv * s2[q]) *
               sin(2.0 * kv * rr + D - (4.0 / 3.0) * kv * kv * kv * c3[q]);
    out[tid] = v;
}

extern "C" int chm_fm_term_rows(chm_ctx *ctx, int nrows, const double *r, const double *s2,
                                const double *c3, const int32_t *absorber, const int32_t *nev,
                                const int32_t *ev, double amp, double *out) {
    CHM_REQUIRE(ctx && ctx->nsp > 0, CHM_E_STATE, "set tables first");
    CHM_REQUIRE(nrows > 0 && r && s2 && c3 && absorber && nev && ev && out, CHM_E_ARG,
                "bad arguments");
    for (int q = 0; q < nrows; ++q) {
        CHM_REQUIRE(absorber[q] >= 0 && absorber[q] < ctx->nsp, CHM_E_ARG, "bad absorber");
        CHM_REQUIRE(nev[q] >= 1 && nev[q] <= 3, CHM_E_ARG, "1..3 scattering events per row");
        for (int m = 0; m < nev[q]; ++m)
            CHM_REQUIRE(ev[3 * q + m] >= 0 && ev[3 * q + m] < ctx->nsp, CHM_E_ARG,
                        "bad scatterer species");
        CHM_REQUIRE(r[q] > 0.0, CHM_E_ARG, "row radius must be positive");
    }
    CHM_HIP(hipSetDevice(ctx->device));
    chm::DevBuf dr, ds2, dc3, da, dn, dev;
    int rc;
    if ((rc = upload(ctx, dr, r, sizeof(double) * nrows))) return rc;
    if ((rc = upload(ctx, ds2, s2, sizeof(double) * nrows))) return rc;
    if ((rc = upload(ctx, dc3, c3, sizeof(double) * nrows))) return rc;
    if ((rc = upload(ctx, da, absorber, sizeof(int32_t) * nrows))) return rc;
    if ((rc = upload(ctx, dn, nev, sizeof(int32_t) * nrows))) return rc;
    if ((rc = upload(ctx, dev, ev, sizeof(int32_t) * 3 * nrows))) return rc;
    long tot = (long)nrows * ctx->nk;
    if ((rc = ctx->rows_out.ensure(sizeof(double) * tot))) return rc;
    k_term_rows<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, ctx->stream>>>(
        dr.as<double>(), ds2.as<double>(), dc3.as<double>(), da.as<int>(), dn.as<int>(),
        dev.as<int>(), ctx->f.as<double>(), ctx->delta.as<double>(), ctx->lam.as<double>(),
        ctx->k.as<double>(), ctx->rows_out.as<double>(), nrows, ctx->nk, ctx->nsp, ctx->s02, amp);
    CHM_LAUNCH_CHECK();
    return download(ctx, out, ctx->rows_out, sizeof(double) * tot);
}

// ------------------------------------------------------- tabulated term rows

// Row q from tabulated F, D, L (nrows, nk) on the model grid:
//   amp s02 F / (k r^2) exp(-2 r / max(L, 1e-6) - 2 k^2 s2) sin(2 k r + D)
__global__ void k_tab_rows(const double *__restrict__ r, const double *__restrict__ s2,
                           const double *__restrict__ F, const double *__restrict__ D,
                           const double *__restrict__ L, const double *__restrict__ ktab,
                           double *__restrict__ out, int nrows, int nk, double s02, double amp) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)nrows * nk) return;
    int q = (int)(tid / nk), ik = (int)(tid % nk);
    double kv = ktab[ik];
    if (!(kv > 0.0)) {
        out[tid] = 0.0;
        return;
    }
    double rr = r[q];
    out[tid] = amp * s02 * F[tid] / (kv * rr * rr) *
               exp(-2.0 * rr / fmax(L[tid], 1e-6) - 2.0 * kv * kv * s2[q]) *
               sin(2.0 * kv * rr + D[tid]);
}

extern "C" int chm_fm_tab_rows(chm_ctx *ctx, int nrows, const double *r, const double *s2,
                               const double *F, const double *D, const double *L, double amp,
                               double *out) {
    CHM_REQUIRE(ctx && ctx->nsp > 0, CHM_E_STATE, "set tables first");
    CHM_REQUIRE(nrows > 0 && r && s2 && F && D && L && out, CHM_E_ARG, "bad arguments");
    for (int q = 0; q < nrows; ++q) CHM_REQUIRE(r[q] > 0.0, CHM_E_ARG, "row radius must be positive");
    CHM_HIP(hipSetDevice(ctx->device));
    const size_t n = (size_t)nrows * ctx->nk;
    chm::DevBuf dr, ds2, dF, dD, dL;
    int rc;
    if ((rc = upload(ctx, dr, r, sizeof(double) * nrows))) return rc;
    if ((rc = upload(ctx, ds2, s2, sizeof(double) * nrows))) return rc;
    if ((rc = upload(ctx, dF, F, sizeof(double) * n))) return rc;
    if ((rc = upload(ctx, dD, D, sizeof(double) * n))) return rc;
    if ((rc = upload(ctx, dL, L, sizeof(double) * n))) return rc;
    if ((rc = ctx->rows_out.ensure(sizeof(double) * n))) return rc;
    k_tab_rows<<<dim3((unsigned)((n + 255) / 256)), 256, 0, ctx->stream>>>(
        dr.as<double>(), ds2.as<double>(), dF.as<double>(), dD.as<double>(), dL.as<double>(),
        ctx->k.as<double>(), ctx->rows_out.as<double>(), nrows, ctx->nk, ctx->s02, amp);
    CHM_LAUNCH_CHECK();
    return download(ctx, out, ctx->rows_out, sizeof(double) * n);
}
