// Device-resident parallel-tempering RMC sampler in pair space (gfx950): every move the
// reference makes, on any positions.
//
// The count-space sampler (sampler.hip) covers the swap-only chain on the ideal lattice.
// This one keeps what the reference keeps per replica (methods/method_a_rmc.py:165-185):
// species, positions, the distance and type of every ordered pair, the unnormalised
// chi(k), chi^2 and the best state with its positions.  One step, for all replicas:
//
//   P1 propose + affected pairs (workgroup 0 of a slot; the others gather MS paths): the
//      splitmix64 proposal with the swap flag, up to 8 partners and the displacement
//      fallback (Box-Muller step clamped to max_disp around the ideal site;
//      exafs_gpu/kernels.py:362-423); the directed pairs
//      touching the moved sites with dedup, their new distance and type (kernels.py:
//      425-488), compacted in pair order; the energy-prior change over them
//      (method_a_rmc.py:310-319); for a swap with multiple scattering, the entry-count
//      change of the paths through the two sites (chimera/mslibrary.py:466-486, the gather
//      of the count-space sampler), compacted to (entry, change) rows.
//   P2 Delta chi (grid: 64-point k chunks x slots, 1024 lanes): lanes own k points,
//      each wavefront a share of the new and old terms of the affected pairs
//      (kernels.py:243-325, records staged in LDS, the forward path sum's term
//      arithmetic) and of the MS rows (change x table row); the 16 shares are added in
//      order in LDS into the slot's trial chi = chi + Delta chi.  Deterministic.
//   P3 trial chi(R) (grid: 16-column blocks x edges x 32-slot groups): trial chi of 32
//      slots over the window's nonzero k rows staged in LDS, times the window's DFT
//      matrix pre-divided by the edge's absorber count (chimera/multift.py:58-63), FP64.
//   P4 accept (8 workgroups per slot, each taking the same decision and committing a
//      share of the copies): the model on the target's footing (complex or
//      |chi(R)|), profiled per-edge scale, chi^2 (method_a_rmc.py:180-227), Metropolis on
//      beta dchi2 nfit + beta_E dE with the counter-based uniform (method_a_rmc.py:
//      278-285), commit of pair distances / types, species, position, chi and chi^2
//      (kernels.py:490-510), best retention with positions (method_a_rmc.py:301-307).
//   Exchange: the count-space sampler's sweep (temperatures move between slots).
//
// A step's cost is the affected terms: 2 x (2 ncoord) pairs for a displacement, up to
// 2 x (4 ncoord) for a swap, each at every k; the full configuration is never summed.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"
#include "fp64math.h"
#include "sampler_dev.h"
#include "sampler_state.h"

using chm_smp::Geo;

namespace {

using namespace chm_smp;

// The move of a slot's pending step, written by P1.
struct PsMove {
    int a, b, na, nb;
    int sw, naff, nms, pad;
    double np[3];   // new position of a (its current position for a swap)
    double de;      // energy-prior change, eV
    // the step's number and the slot's chi^2 and best chi^2 before it, as P1 read them: the
    // accept workgroups decide from these, not from the counters and chi^2 that the lead
    // accept workgroup overwrites during the same launch
    uint64_t n;
    double c2c, bestc;
};

struct PsArgs {
    int nrep, natoms, natoms_pad, ncoord, npairs, nsp, nk, nshell, maxaff, msmax;
    uint64_t seed;
    double swap_frac, disp_sigma, max_disp;
    const int *rung_of;
    const uint64_t *stepc;
    const double *c2, *best_c2;   // the slot's chi^2 and best chi^2 (copied into the move)
    const uint8_t *sp;
    const double *pos, *ideal;
    const int *rev, *iidx, *jidx, *shell;
    const double *shift, *sig2, *r;
    const int *pt;
    const double *etab;   // (nshell, nsp*nsp) or null
    PsMove *mv;
    struct AffRec *aff;   // (nrep, maxaff) the move's pairs, compacted in pair order
    int *dcount;          // (nrep, g.nrow_tot) MS entry-count change of a swap, summed over
                          //   the split workgroups with integer atomics (zeroed by P4)
    int has_ms;
    Geo g;
    const int4 *ms_sten;
};

// One pair a move touches: its index, shell and sigma^2, its distance and type before and
// after the move (everything the Delta chi and the commit read, in one record).
struct AffRec {
    double rnew, rold, sig2;
    int p, tnew, told, sh;
};

// The Gaussian step of a displacement (kernels.py:405-421): two Box-Muller pairs from the
// continuing stream h, the offset from the ideal site clamped to max_disp.  Each operation
// rounded on its own, as the reference's (and NumPy's) arithmetic.
__device__ void displaced(uint64_t h, const double *P, const double *Id, double sigma,
                          double max_disp, double *out) {
#pragma clang fp contract(off)
    double g[3];
    for (int t = 0; t < 3; t += 2) {
        h = splitmix64(h);
        double u1 = u01(h);
        h = splitmix64(h);
        const double u2 = u01(h);
        u1 = fmax(u1, 1e-300);
        const double rad = sqrt(-2.0 * log(u1));
        g[t] = rad * cos(6.283185307179586 * u2);
        if (t + 1 < 3) g[t + 1] = rad * sin(6.283185307179586 * u2);
    }
    const double ox = (P[0] + g[0] * sigma) - Id[0];
    const double oy = (P[1] + g[1] * sigma) - Id[1];
    const double oz = (P[2] + g[2] * sigma) - Id[2];
    const double m = sqrt(ox * ox + oy * oy + oz * oz);
    const double sc = (m > max_disp) ? (max_disp / m) : 1.0;
    out[0] = Id[0] + ox * sc;
    out[1] = Id[1] + oy * sc;
    out[2] = Id[2] + oz * sc;
}

__device__ double pair_distance(const double *pi, const double *pj, const double *sh) {
#pragma clang fp contract(off)
    const double dx = (pj[0] + sh[0]) - pi[0];
    const double dy = (pj[1] + sh[1]) - pi[1];
    const double dz = (pj[2] + sh[2]) - pi[2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Order-preserving compaction of one round of the workgroup's lanes: every lane with
// `keep` gets its index among the kept lanes of this round plus `base` (all lanes).
__device__ __forceinline__ int compact_index(bool keep, int *wcnt, int &base) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    const unsigned long long m = __ballot(keep);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
    if (lane == 0) wcnt[w] = __popcll(m);
    __syncthreads();
    int off = base;
    int tot = 0;
    for (int q = 0; q < nw; ++q) {
        const int c = wcnt[q];
        if (q < w) off += c;
        tot += c;
    }
    base += tot;
    return off + before;
}

// Sum over a wavefront by butterfly exchanges (every lane gets the same, order fixed).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}

// Sum over the workgroup: wavefront sums, then the wavefronts' sums in wave order.
// Deterministic; every lane gets the result.  `red` holds at least 16 doubles.
__device__ double block_sum_d(double v, double *red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0.0;
    for (int q = 0; q < nw; ++q) t += red[q];
    return t;
}

constexpr int kPropThreads = 256;

// ------------------------------------------------------------------------------ P1
// grid (nsplit, slots), 256 lanes.  Every workgroup of a slot stages the slot's species in
// LDS and draws the same proposal; workgroup 0 builds the affected pairs and the energy
// change; with multiple scattering every workgroup gathers its share of the paths through
// the two swapped sites into an LDS histogram and adds its nonzero bins to the slot's
// count-change row with integer atomics (order-independent, so deterministic).
__global__ __launch_bounds__(kPropThreads) void k_ps_propose(PsArgs A) {
    extern __shared__ int lds[];
    // [MS entry-count change (g.nrow_tot ints), MS only][species (natoms_pad bytes)]
    int *hist = lds;
    uint8_t *lsp = reinterpret_cast<uint8_t *>(lds + (A.has_ms ? ((A.g.nrow_tot + 3) & ~3) : 0));
    __shared__ PsMove M;
    __shared__ int wcnt[16];
    __shared__ double red[16];
    const int s = blockIdx.y, split = blockIdx.x, nsplit = gridDim.x;
    const int rung = A.rung_of[s];
    const uint64_t n = A.stepc[s] + 1;
    const uint8_t *sp = A.sp + (size_t)s * A.natoms_pad;
    const double *pos = A.pos + (size_t)s * A.natoms * 3;
    load_species(lsp, sp, A.natoms_pad);
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t h = splitmix64(A.seed ^ (n * kGolden) ^ ((uint64_t)rung << 32));
        h = splitmix64(h);
        const bool swf = u01(h) < A.swap_frac;
        h = splitmix64(h);
        const int a = min((int)(u01(h) * A.natoms), A.natoms - 1);
        const int sa = lsp[a];
        int b = a, sb = sa;
        if (swf) {
            // the 8 candidate draws do not depend on the configuration: draw all, then
            // take the first of another species (the stream after a failed search is the
            // one after the 8th draw, as in the reference's loop)
            int cand[8];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                h = splitmix64(h);
                cand[t] = min((int)(u01(h) * A.natoms), A.natoms - 1);
            }
#pragma unroll
            for (int t = 7; t >= 0; --t) {
                const int sc = lsp[cand[t]];
                if (sc != sa) {
                    b = cand[t];
                    sb = sc;
                }
            }
        }
        const bool real = swf && b != a;
        M.a = a;
        M.b = b;
        M.na = real ? sb : sa;
        M.nb = real ? sa : sb;
        M.sw = real ? 1 : 0;
        if (real) {
            M.np[0] = pos[3 * a];
            M.np[1] = pos[3 * a + 1];
            M.np[2] = pos[3 * a + 2];
        } else {
            displaced(h, pos + 3 * a, A.ideal + 3 * a, A.disp_sigma, A.max_disp, M.np);
        }
    }
    __syncthreads();
    const int a = M.a, b = M.b, na = M.na, nb = M.nb, sw = M.sw;
    if (A.has_ms && sw) {
        const Geo &g = A.g;
        for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x) hist[q] = 0;
        __syncthreads();
        SwapSite W;
        W.p = Prop{a, b, na, nb, 1};
        site_xyz_of(g, a, W.ax, W.ay, W.az);
        site_xyz_of(g, b, W.bx, W.by, W.bz);
        // this workgroup's items, each lane's stencil records requested together first
        // with several workgroups the first builds the affected pairs meanwhile and
        // gathers nothing
        constexpr int kItems = 8;
        const int ng = nsplit > 1 ? nsplit - 1 : 1, gi = nsplit > 1 ? split - 1 : 0;
        const int stride = ng * (int)blockDim.x;
        for (int it0 = (gi < 0 ? 2 * g.nstencil : gi * (int)blockDim.x); it0 < 2 * g.nstencil; it0 += kItems * stride) {
            int4 rec[kItems];
#pragma unroll
            for (int u = 0; u < kItems; ++u) {
                const int it = it0 + u * stride + (int)threadIdx.x;
                const int m = it < g.nstencil ? it : it - g.nstencil;
                rec[u] = A.ms_sten[min(max(m, 0), g.nstencil - 1)];
            }
#pragma unroll
            for (int u = 0; u < kItems; ++u) {
                const int it = it0 + u * stride + (int)threadIdx.x;
                if (it < 2 * g.nstencil) gather_item_sp(g, W, SpPlain{lsp}, hist, it, 0, rec[u]);
            }
        }
        __syncthreads();
        int *dc = A.dcount + (size_t)s * g.nrow_tot;
        for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x) {
            const int v = hist[q];
            if (v != 0) atomicAdd(&dc[q], v);
        }
    }
    if (split != 0) return;
    const int C = A.ncoord, nn = A.nsp * A.nsp;
    const double *rs = A.r + (size_t)s * A.npairs;
    const int *ts = A.pt + (size_t)s * A.npairs;
    AffRec *aff = A.aff + (size_t)s * A.maxaff;
    double de = 0.0;
    int base = 0;
    for (int q0 = 0; q0 < A.maxaff; q0 += blockDim.x) {
        const int q = q0 + (int)threadIdx.x;
        // item q: block q / C (0: out-pairs of a, 1: their reverses, 2, 3: the same of b),
        // the o-th pair of the site's i-sorted run; the pair's ends follow from the block
        // without reading the pair list (a reverse pair runs other -> site).  Every load
        // below has an in-range index whether the item is kept or not (no branch around it),
        // so each round of loads is issued together.
        const bool in = q < A.maxaff;
        const int qq = in ? q : 0;
        const int blk = qq / C, o = qq - blk * C;
        const int site = blk < 2 ? a : b;
        const int p0 = site * C + o;
        const int other = A.jidx[p0];
        const int rv = A.rev[p0];
        const bool rvs = (blk & 1) != 0;
        const int p = rvs ? rv : p0;
        const int i = rvs ? other : site, j = rvs ? site : other;
        // dedup (kernels.py:448-468): b's blocks drop everything for a displacement and
        // every pair whose other end is a
        const bool keep = in && !(blk >= 2 && (b == a || other == a));
        const double rold = rs[p];
        const int told = ts[p];
        const int sh = A.shell[p];
        const double s2p = A.sig2[p];
        double rn = rold;
        if (!sw) {
            const double *pi = (i == a) ? M.np : pos + 3 * i;
            const double *pj = (j == a) ? M.np : pos + 3 * j;
            rn = pair_distance(pi, pj, A.shift + 3 * p);
        }
        const int si = (i == a) ? na : ((i == b) ? nb : (int)lsp[i]);
        const int sj = (j == a) ? na : ((j == b) ? nb : (int)lsp[j]);
        const int tn = si * A.nsp + sj;
        if (A.etab && keep) de += A.etab[sh * nn + tn] - A.etab[sh * nn + told];
        const int at = compact_index(keep, wcnt, base);
        if (keep) {
            AffRec rec;
            rec.rnew = rn;
            rec.rold = rold;
            rec.sig2 = s2p;
            rec.p = p;
            rec.tnew = tn;
            rec.told = told;
            rec.sh = sh;
            aff[at] = rec;
        }
    }
    if (A.etab) de = block_sum_d(de, red);
    if (threadIdx.x == 0) {
        PsMove o = M;
        o.naff = base;
        o.nms = 0;
        o.pad = 0;
        o.de = de;
        o.n = n;
        o.c2c = A.c2[s];
        o.bestc = A.best_c2[s];
        A.mv[s] = o;
    }
}

// ------------------------------------------------------------------------------ P2
struct TermRec {
    double r;      // r_eff
    double lir2;   // ln(1 / r_eff^2)
    double s2;     // sigma^2 + s2tab
    double c3;     // c3tab[shell]
    int t;         // pair type
    int sign;      // +1 new term, -1 old term
};

constexpr int kDeltaWaves = 16;   // term groups per workgroup (one wavefront each)

struct PsDeltaArgs {
    int nrep, npairs, nk, k0, maxaff, msmax, nblk;
    double s02;
    const PsMove *mv;
    const AffRec *aff;
    const double *ftab, *dtab, *lamtab, *ktab, *drtab, *c3tab, *s2tab;
    const double *dms;     // (ngrp, nrep, nsp, nk) MS chi change of the move in entry groups, or null
    int ngrp;
    const double *chi;     // (nrep, nsp, nk) current unnormalised chi
    double *trial;         // (nrep, nsp, nk) chi + Delta chi
    double *trialT;        // (nsp, nk, nrep) the same, k-major for the transform
};

// grid (chunks of 64 k points from k0, slots), 1024 lanes: wavefront w evaluates term
// group w (a contiguous share of the move's 2 naff terms, then of its MS rows) at the
// chunk's 64 k points, one per lane; the 16 groups' sums are added in group order in LDS.
// The first chunk also covers the k points below k0 (no pair term there, MS rows only).
template <int NSP>
__global__ __launch_bounds__(1024) void k_ps_delta(PsDeltaArgs A, const chm::PolyCoefs pc) {
    extern __shared__ double smem[];
    double *sf = smem;                              // f (NSP, 64)
    double *sd = sf + NSP * 64;                     // delta/2pi - 1/4 (NSP*NSP, 64)
    double *red = sd + NSP * NSP * 64;              // (16, NSP, 64) group sums
    TermRec *tile = reinterpret_cast<TermRec *>(red + kDeltaWaves * NSP * 64);
    constexpr double kInv2Pi = 0.15915494309189533577;
    constexpr double kInvPi = 0.31830988618379067154;
    const int s = blockIdx.y, nk = A.nk;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int kb = A.k0 + (int)blockIdx.x * 64;
    const int kc = min(64, nk - kb);
    for (int q = threadIdx.x; q < NSP * 64; q += blockDim.x) {
        const int row = q >> 6, c = q & 63;
        sf[q] = c < kc ? A.ftab[row * nk + kb + c] : 0.0;
    }
    for (int q = threadIdx.x; q < NSP * NSP * 64; q += blockDim.x) {
        const int row = q >> 6, c = q & 63;
        sd[q] = c < kc ? fma(A.dtab[row * nk + kb + c], kInv2Pi, -0.25) : 0.0;
    }
    const PsMove mv = A.mv[s];
    const int nterm = 2 * mv.naff;
    const int nn = NSP * NSP;
    // stage every term record of the move (2 naff <= 8 ncoord <= blockDim)
    for (int t = threadIdx.x; t < nterm; t += blockDim.x) {
        const AffRec ar = A.aff[(size_t)s * A.maxaff + (t >> 1)];
        const bool nw = (t & 1) == 0;
        const double rr = nw ? ar.rnew : ar.rold;
        const int tt = nw ? ar.tnew : ar.told;
        const double re = rr + A.drtab[ar.sh * nn + tt];
        TermRec rec;
        rec.r = re;
        rec.lir2 = -2.0 * log(re);
        rec.s2 = ar.sig2 + A.s2tab[ar.sh * nn + tt];
        rec.c3 = A.c3tab[ar.sh];
        rec.t = tt;
        rec.sign = nw ? 1 : -1;
        tile[t] = rec;
    }
    __syncthreads();
    const int ik = kb + lane;
    const bool live = lane < kc;
    const double kv = live ? A.ktab[ik] : 0.0;
    const double lv = live ? A.lamtab[ik] : 1.0;
    const double kp = kv * kInvPi;
    const double c1 = -2.0 / lv, c2 = -2.0 * kv * kv;
    const double c3k = (4.0 / 3.0) * kv * kv * kv * kInv2Pi;
    const double pref = (live && kv > 0.0) ? A.s02 / kv : 0.0;
    double acc[NSP];
#pragma unroll
    for (int e = 0; e < NSP; ++e) acc[e] = 0.0;
    const int per = (nterm + kDeltaWaves - 1) / kDeltaWaves;
    const int tb = min(nterm, w * per), te = min(nterm, tb + per);
    for (int q = tb; q < te; ++q) {
        const TermRec rc = tile[q];
        const int tt = __builtin_amdgcn_readfirstlane(rc.t);
        const int sa = tt / NSP, sb = tt - sa * NSP;
        const double ph = fma(kp, rc.r, fma(-c3k, rc.c3, sd[tt * 64 + lane]));
        const double fe = sf[sb * 64 + lane] * chm::exp_fin(fma(c1, rc.r, fma(c2, rc.s2, rc.lir2)), pc);
        const double v = (rc.sign > 0 ? fe : -fe) * chm::cos2pi(ph, pc);
#pragma unroll
        for (int e = 0; e < NSP; ++e)
            if (sa == e) acc[e] += v;
    }
#pragma unroll
    for (int e = 0; e < NSP; ++e) red[(w * NSP + e) * 64 + lane] = pref * acc[e];
    __syncthreads();
    // wavefront e < NSP adds the groups of absorber e in group order
    if (w < NSP && live) {
        double d = 0.0;
        for (int z = 0; z < kDeltaWaves; ++z) d += red[(z * NSP + w) * 64 + lane];
        const size_t at = ((size_t)s * NSP + w) * nk + ik;
        double m = 0.0;
        if (A.dms)
            for (int g = 0; g < A.ngrp; ++g) m += A.dms[(size_t)g * A.nrep * NSP * nk + at];
        const double t = A.chi[at] + d + m;
        A.trial[at] = t;
        A.trialT[((size_t)w * nk + ik) * A.nrep + s] = t;
    }
    if (blockIdx.x == 0 && w < NSP) {
        // k points below k0: no pair term, the MS change only
        for (int q = lane; q < A.k0; q += 64) {
            const size_t at = ((size_t)s * NSP + w) * nk + q;
            double m = 0.0;
            if (A.dms)
                for (int g = 0; g < A.ngrp; ++g) m += A.dms[(size_t)g * A.nrep * NSP * nk + at];
            const double t = A.chi[at] + m;
            A.trial[at] = t;
            A.trialT[((size_t)w * nk + q) * A.nrep + s] = t;
        }
    }
}

// MS chi change of every slot's swap: grid (32-point chunks of all k, edges, 32-slot
// groups x entry groups), 256 lanes, lane = (k, four slots).  Almost every entry of a
// swap's two sites changes, so this is a dense product over the entries,
// dms[g][s][e][k] = sum_{j in group g} dC[s][e][j] T[e][j][k]: the group's count changes
// ([entry][slot]) and table rows ([entry][k]) are staged in LDS with every lane's loads
// requested together, then each lane runs over the entries with one 16-byte count read
// and one table read per four multiply-adds.  The delta kernel adds the groups' partials.
constexpr int kMsRowsMax = 128;   // entries per group
__global__ __launch_bounds__(256) void k_ps_msgemm(int nrep, int nk, int nsp, int nblk, int rstride, int nss_rows,
                                                  int nrow_tot, int ngrp, const int *__restrict__ dcount,
                                                  const double *__restrict__ tms, double *__restrict__ dms) {
    __shared__ int4 dcl[kMsRowsMax * 8];           // (entry, slot quad)
    __shared__ double tt[kMsRowsMax * 32];         // (entry, k)
    const int e = blockIdx.y, k0 = blockIdx.x * 32;
    const int sg = blockIdx.z / ngrp, grp = blockIdx.z - sg * ngrp, s0 = sg * 32;
    const int rpg = (nblk + ngrp - 1) / ngrp;
    const int jb = min(nblk, grp * rpg), jn = min(nblk, jb + rpg) - jb;   // jn <= kMsRowsMax
    const int kk = threadIdx.x & 31, sq = threadIdx.x >> 5;
    {
        const Buf bd = make_buf(dcount, (uint32_t)(sizeof(int) * (size_t)nrep * nrow_tot));
        const Buf bt = make_buf(tms, (uint32_t)(sizeof(double) * (size_t)nsp * nblk * nk));
        constexpr int U = kMsRowsMax * 32 / 256;
        int vd[U];
        double vt[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = u * 256 + (int)threadIdx.x;
            const int j = q >> 5, c = q & 31;
            const int s = s0 + c, k = k0 + c;
            vd[u] = buf_i32(bd, j < jn && s < nrep,
                            4u * ((uint32_t)s * (uint32_t)nrow_tot + (uint32_t)(e * rstride + nss_rows + jb + j)));
            vt[u] = buf_f64(bt, j < jn && k < nk, (uint32_t)(sizeof(double) * (((size_t)e * nblk + jb + j) * nk + k)));
        }
        int *dci = reinterpret_cast<int *>(dcl);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int q = u * 256 + (int)threadIdx.x;
            dci[q] = vd[u];
            tt[q] = vt[u];
        }
    }
    __syncthreads();
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
#pragma unroll 4
    for (int j = 0; j < jn; ++j) {
        const int4 d = dcl[j * 8 + sq];
        const double t = tt[j * 32 + kk];
        a0 = fma((double)d.x, t, a0);
        a1 = fma((double)d.y, t, a1);
        a2 = fma((double)d.z, t, a2);
        a3 = fma((double)d.w, t, a3);
    }
    const int k = k0 + kk;
    if (k < nk) {
        const double a[4] = {a0, a1, a2, a3};
#pragma unroll
        for (int h = 0; h < 4; ++h) {
            const int s = s0 + 4 * sq + h;
            if (s < nrep) dms[(((size_t)grp * nrep + s) * nsp + e) * nk + k] = a[h];
        }
    }
}

// ------------------------------------------------------------------------------ P3
// Trial chi(R) and its residual sums: grid (column blocks of 16 of one edge, edges, slot
// groups of 32), 256 lanes.  Over the window's nonzero k rows [kb0, kb1) in tiles of 128:
// the edge's trial chi of the group's 32 slots (k-major) and the tile of the DFT matrix
// Te[e][k][cols] are staged in LDS, every lane's loads requested together (bounds-checked
// buffer loads, no branch); lane (column c, slot pair) accumulates two slots in FP64.
// The epilogue forms the model m (the column, or |z| of a point for an |chi(R)| target)
// and its residual d = m - t against the target, and reduces over the block's columns
//   U1 = sum d t,  U2 = sum d^2,  W1 = sum w d t,  W2 = sum w d^2    (w = 1/sigma^2)
// per slot: with T0 = sum t^2 and Q0 = sum w t^2 of the edge, the profiled scale is
// s = (T0 + U1) / (T0 + 2 U1 + U2) and the edge's share of chi^2 nfit is
// s^2 W2 + 2 s (s - 1) W1 + (s - 1)^2 Q0 (method_a_rmc.py:190-227); every term is small
// near a fit, so nothing cancels.
constexpr int kTfCols = 16, kTfK = 128;
struct TfArgs {
    int nrep, nk, nsp, kb0, kb1, ncol_e, nbc, magnitude;
    const double *trialT;  // (nsp, nk, nrep)
    const double *Te;      // (nsp, nk, ncol_e) DFT matrix / absorber count
    const double *Rx;      // (nsp * ncol_e) transform of the constant term
    const double *tgt;     // complex: (nsp * ncol_e); magnitude: (nsp * ncol_e / 2)
    const double *wgt;     // (nsp * ncol_e / 2) 1 / sigma^2 per point
    double *sums;          // (nrep, nsp, nbc, 4)
};

__global__ __launch_bounds__(256) void k_ps_transform(TfArgs A) {
    __shared__ double sa[kTfK * 32];
    __shared__ double sb[kTfK * kTfCols];
    const int e = blockIdx.y, s0 = blockIdx.z * 32, c0 = blockIdx.x * kTfCols;
    const int nrep = A.nrep, nk = A.nk, ncol_e = A.ncol_e;
    const Buf bA = make_buf(A.trialT, (uint32_t)(sizeof(double) * (size_t)A.nsp * nk * nrep));
    const Buf bB = make_buf(A.Te, (uint32_t)(sizeof(double) * (size_t)A.nsp * nk * ncol_e));
    const int c = threadIdx.x & (kTfCols - 1), sp = threadIdx.x / kTfCols;
    double a0 = 0.0, a1 = 0.0;
    double va[kTfK * 32 / 256], vb[kTfK * kTfCols / 256];
    auto load_tile = [&](int k0) {
        const int kn = min(kTfK, A.kb1 - k0);
#pragma unroll
        for (int u = 0; u < kTfK * 32 / 256; ++u) {
            const int q = u * 256 + threadIdx.x, k = q >> 5, cc = q & 31;
            const bool ok = k < kn && s0 + cc < nrep;
            va[u] = buf_f64(bA, ok, (uint32_t)(sizeof(double) * (((size_t)e * nk + k0 + k) * nrep + s0 + cc)));
        }
#pragma unroll
        for (int u = 0; u < kTfK * kTfCols / 256; ++u) {
            const int q = u * 256 + threadIdx.x, k = q / kTfCols, cc = q % kTfCols;
            const bool ok = k < kn && c0 + cc < ncol_e;
            vb[u] = buf_f64(bB, ok, (uint32_t)(sizeof(double) * (((size_t)e * nk + k0 + k) * ncol_e + c0 + cc)));
        }
    };
    load_tile(A.kb0);
    for (int k0 = A.kb0; k0 < A.kb1; k0 += kTfK) {
        const int kn = min(kTfK, A.kb1 - k0);
        __syncthreads();   // the previous tile's reads are done
#pragma unroll
        for (int u = 0; u < kTfK * 32 / 256; ++u) sa[u * 256 + threadIdx.x] = va[u];
#pragma unroll
        for (int u = 0; u < kTfK * kTfCols / 256; ++u) sb[u * 256 + threadIdx.x] = vb[u];
        __syncthreads();
        if (k0 + kTfK < A.kb1) load_tile(k0 + kTfK);   // in flight during this tile
#pragma unroll 8
        for (int k = 0; k < kn; ++k) {
            const double2 x = *reinterpret_cast<const double2 *>(sa + k * 32 + 2 * sp);
            const double bv = sb[k * kTfCols + c];
            a0 = fma(x.x, bv, a0);
            a1 = fma(x.y, bv, a1);
        }
    }
    // residual sums of the block's columns, per slot
    const int col = c0 + c;
    const bool ok = col < ncol_e;
    const size_t gc = (size_t)e * ncol_e + (ok ? col : 0);    // column over all edges
    const size_t pt = gc >> 1;                                  // its R point
    const double rx = ok ? A.Rx[gc] : 0.0;
    const double w = ok ? A.wgt[pt] : 0.0;
    const double tc = ok ? (A.magnitude ? A.tgt[pt] : A.tgt[gc]) : 0.0;
    double S4[2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const double z = (h ? a1 : a0) + rx;
        const double zp = __shfl_xor(z, 1);     // the point's other part (lane c ^ 1)
        double d, ww, t;
        if (A.magnitude) {
            const bool use = ok && (c & 1) == 0;
            t = use ? tc : 0.0;
            d = use ? hypot(z, zp) - tc : 0.0;
            ww = use ? w : 0.0;
        } else {
            t = tc;
            d = ok ? z - tc : 0.0;
            ww = w;
        }
        S4[h][0] = d * t;
        S4[h][1] = d * d;
        S4[h][2] = ww * d * t;
        S4[h][3] = ww * d * d;
    }
#pragma unroll
    for (int m = kTfCols / 2; m >= 1; m >>= 1)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int v = 0; v < 4; ++v) S4[h][v] += __shfl_xor(S4[h][v], m);
    if (c == 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int s = s0 + 2 * sp + h;
            if (s >= nrep) continue;
            double *o = A.sums + (((size_t)s * A.nsp + e) * A.nbc + blockIdx.x) * 4;
#pragma unroll
            for (int v = 0; v < 4; ++v) o[v] = S4[h][v];
        }
    }
}

// chi^2 of slot s from its residual sums (nsp, nbc, 4), by one wavefront (every lane gets
// it): the blocks are added in block order per lane, then the lanes (rows of 16, then rows).
__device__ double c2_from_sums(const double *sums, int nsp, int nbc, const double *q0, int profile,
                               double nfit) {
    const int lane = threadIdx.x & 63;
    // every edge's first two block rounds requested together (bounds-checked, no branch)
    const Buf bs = make_buf(sums, (uint32_t)(sizeof(double) * 4 * (size_t)nsp * nbc));
    double P[chm::kMaxNsp][2][4];
#pragma unroll
    for (int e = 0; e < chm::kMaxNsp; ++e)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int b = lane + 64 * h;
            const bool ok = e < nsp && b < nbc;
            const uint32_t off = (uint32_t)(sizeof(double) * 4 * ((size_t)(ok ? e : 0) * nbc + b));
#pragma unroll
            for (int v = 0; v < 4; ++v) P[e][h][v] = buf_f64(bs, ok, off + 8u * v);
        }
    double U[chm::kMaxNsp][4];
#pragma unroll
    for (int e = 0; e < chm::kMaxNsp; ++e)
#pragma unroll
        for (int v = 0; v < 4; ++v) U[e][v] = P[e][0][v] + P[e][1][v];
#pragma unroll
    for (int e = 0; e < chm::kMaxNsp; ++e) {
        if (e >= nsp) break;
        for (int b = lane + 128; b < nbc; b += 64)
#pragma unroll
            for (int v = 0; v < 4; ++v) U[e][v] += sums[((size_t)e * nbc + b) * 4 + v];
    }
    // the sums over the wavefront: each row of 16 lanes by DPP, then the four rows
#pragma unroll
    for (int e = 0; e < chm::kMaxNsp; ++e) {
        if (e >= nsp) break;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double x = chm_smp::group_sum_d<16>(U[e][v]);
            x += __shfl_xor(x, 16);
            U[e][v] = x + __shfl_xor(x, 32);
        }
    }
    double tot = 0.0;
#pragma unroll
    for (int e = 0; e < chm::kMaxNsp; ++e) {
        if (e >= nsp) break;
        const double T0 = q0[2 * e], Q0 = q0[2 * e + 1];
        double sc = 1.0, sm1 = 0.0;
        if (profile) {
            const double P2 = fmax(U[e][1] + 2.0 * U[e][0] + T0, 1e-30);
            const double P1 = U[e][0] + T0;
            sc = P1 / P2;
            sm1 = (P1 - P2) / P2;
        }
        tot += sc * sc * U[e][3] + 2.0 * sc * sm1 * U[e][2] + sm1 * sm1 * Q0;
    }
    return tot / nfit;
}

// ------------------------------------------------------------------------------ P4
struct PsAcceptArgs {
    int nrep, natoms, natoms_pad, npairs, nsp, nk, nkw, nwin, ncol_e, maxaff, magnitude, profile;
    double nfit, ebeta;
    int has_prior;
    uint64_t seed;
    const int *rung_of;
    uint64_t *stepc;
    const double *beta_slot;
    const double *sums;      // (nrep, nsp, nbc, 4) residual sums of the trial chi(R)
    const double *q0;        // (nsp, 2) per edge T0 = sum t^2, Q0 = sum t^2 / sigma^2
    int nbc;                 // column blocks per edge
    const PsMove *mv;
    const AffRec *aff;
    const double *trial;
    uint8_t *sp;
    double *pos, *r, *chi, *c2, *best_c2, *best_pos;
    int *pt;
    uint8_t *best_sp;
    long long *acc_rung;
    int *dcount;             // (nrep, nrow_tot) zeroed for the next step (MS only)
    int nrow_tot;
    int4 *log;               // (log_cap, nrep): a, b, is-swap | accepted << 1, slot
    double *log_c2;          // (log_cap, nrep): trial chi^2
    int log_cap;
    long long rec_base;
};

constexpr int kAccThreads = 256;
constexpr int kAccParts = 8;   // workgroups per replica: each takes the same decision and a
                               // share of the commit and best-state copies

// Workgroup (part, slot) of the accept launch, grid kAccParts x nrep.  Every workgroup
// decides (the same arithmetic on the same inputs: the residual sums and the step number,
// chi^2 and best chi^2 that P1 copied into the move, none of which this launch writes, so a
// workgroup that starts after the lead has finished decides the same); part 0 makes the
// scalar updates (log,
// chi^2, counters, the moved species / position, step counter); the copies (trial chi into
// the running chi, the affected pairs' distances and types, and on a new best the species
// and positions into the best state, with the move applied) are split between the parts.
__global__ __launch_bounds__(kAccThreads) void k_ps_accept(PsAcceptArgs A) {
    __shared__ double red[1];
    __shared__ int accs, imps;
    const int s = blockIdx.y, part = blockIdx.x, nparts = gridDim.x;
    const bool lead = part == 0;
    const int rung = A.rung_of[s];
    const PsMove mv = A.mv[s];
    const uint64_t n = mv.n;
    const double bslot = A.beta_slot[s], c2c = mv.c2c, bestc = mv.bestc;
    if (threadIdx.x < 64) {
        const double v = c2_from_sums(A.sums + (size_t)s * A.nsp * A.nbc * 4, A.nsp, A.nbc, A.q0, A.profile, A.nfit);
        if (threadIdx.x == 0) red[0] = v;
    }
    __syncthreads();
    const double c2t = red[0];
    if (threadIdx.x == 0) {
        double x = bslot * (c2t - c2c) * A.nfit;
        if (A.has_prior) x += A.ebeta * mv.de;
        const uint64_t h = splitmix64(A.seed ^ (n * kGolden) ^ ((uint64_t)rung << 32) ^ kAccSalt);
        const bool acc = metropolis(x, u01(splitmix64(h)));
        accs = acc ? 1 : 0;
        const double cur = acc ? c2t : c2c;
        imps = (cur < bestc) ? 1 : 0;
        if (lead && A.log_cap > 0) {
            const long long ri = (long long)n - 1 - A.rec_base;
            if (ri >= 0) {
                const size_t at = (size_t)(ri % A.log_cap) * A.nrep + rung;
                A.log[at] = make_int4(mv.a, mv.b, mv.sw | (acc ? 2 : 0), s);
                A.log_c2[at] = c2t;
            }
        }
    }
    __syncthreads();
    const int acc = accs;
    const int tid = part * (int)blockDim.x + (int)threadIdx.x, nthr = nparts * (int)blockDim.x;
    if (acc) {
        for (int q = tid; q < mv.naff; q += nthr) {
            const AffRec ar = A.aff[(size_t)s * A.maxaff + q];
            A.r[(size_t)s * A.npairs + ar.p] = ar.rnew;
            A.pt[(size_t)s * A.npairs + ar.p] = ar.tnew;
        }
        const size_t ck = (size_t)A.nsp * A.nk;
        for (size_t q = tid; q < ck; q += nthr) A.chi[s * ck + q] = A.trial[s * ck + q];
        if (lead && threadIdx.x == 0) {
            uint8_t *sp = A.sp + (size_t)s * A.natoms_pad;
            sp[mv.a] = (uint8_t)mv.na;
            sp[mv.b] = (uint8_t)mv.nb;
            if (!mv.sw) {
                double *P = A.pos + ((size_t)s * A.natoms + mv.a) * 3;
                P[0] = mv.np[0];
                P[1] = mv.np[1];
                P[2] = mv.np[2];
            }
            A.c2[s] = c2t;
            A.acc_rung[rung] += 1;
        }
    }
    if (imps) {
        // the best state after this step: the species and positions as stored (the lead may
        // be writing the move into them now), with this step's move applied when accepted
        {
            const uint4 *src = reinterpret_cast<const uint4 *>(A.sp + (size_t)s * A.natoms_pad);
            uint4 *dst = reinterpret_cast<uint4 *>(A.best_sp + (size_t)s * A.natoms_pad);
            for (int q = tid; q < A.natoms_pad / 16; q += nthr) {
                uint4 w = src[q];
                if (acc) {
                    unsigned char *b = reinterpret_cast<unsigned char *>(&w);
                    if (mv.a >= 16 * q && mv.a < 16 * q + 16) b[mv.a - 16 * q] = (unsigned char)mv.na;
                    if (mv.b >= 16 * q && mv.b < 16 * q + 16) b[mv.b - 16 * q] = (unsigned char)mv.nb;
                }
                dst[q] = w;
            }
        }
        {
            const int np = 3 * A.natoms;
            const double *src = A.pos + (size_t)s * np;
            double *dst = A.best_pos + (size_t)s * np;
            const Buf bp = make_buf(src, (uint32_t)(sizeof(double) * np));
            const bool moved = acc && !mv.sw;
            constexpr int U = 8;
            for (int q0 = 0; q0 < np; q0 += U * nthr) {
                double v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int q = q0 + u * nthr + tid;
                    v[u] = buf_f64(bp, q < np, 8u * (uint32_t)q);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int q = q0 + u * nthr + tid;
                    const int k = q - 3 * mv.a;   // (a selection: no indexed register array)
                    const double pk = (k == 0) ? mv.np[0] : (k == 1) ? mv.np[1] : mv.np[2];
                    if (q < np) dst[q] = (moved && k >= 0 && k < 3) ? pk : v[u];
                }
            }
        }
        if (lead && threadIdx.x == 0) A.best_c2[s] = acc ? c2t : c2c;
    }
    if (A.dcount)
        for (int q = tid; q < A.nrow_tot; q += nthr) A.dcount[(size_t)s * A.nrow_tot + q] = 0;
    if (lead && threadIdx.x == 0) A.stepc[s] = n;
}

// chi^2 of every row of `sums` (for init and chm_psampler_chi2), one wavefront per row
__global__ __launch_bounds__(64) void k_ps_chi2(const double *__restrict__ sums, int nsp, int nbc,
                                               const double *__restrict__ q0, int profile, double nfit,
                                               double *__restrict__ out) {
    const double v = c2_from_sums(sums + (size_t)blockIdx.x * nsp * nbc * 4, nsp, nbc, q0, profile, nfit);
    if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// chi rows (nconf, nsp, nk) -> k-major (nsp, nk, nconf), the transform's input layout
__global__ void k_ps_to_kmajor(int nconf, int nsp, int nk, const double *__restrict__ chi, double *__restrict__ outT) {
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)nconf * nsp * nk) return;
    const int s = (int)(tid / ((long)nsp * nk));
    const int q = (int)(tid - (long)s * nsp * nk);
    outT[(size_t)q * nconf + s] = chi[tid];
}

__global__ void k_ps_exchange(int nrep, const double *__restrict__ beta, const double *__restrict__ c2,
                              int *__restrict__ slot_of, int *__restrict__ rung_of,
                              double *__restrict__ beta_slot, uint64_t *__restrict__ iex,
                              long long *__restrict__ att, long long *__restrict__ acc, uint64_t seed,
                              double nfit) {
    exchange_body(nrep, beta, c2, slot_of, rung_of, beta_slot, iex, att, acc, seed, nfit);
}

// per slot: species (u8) from int32, positions copied, best = current
__global__ void k_ps_init_slots(int nrep, int natoms, int natoms_pad, const int *__restrict__ sp0,
                                const double *__restrict__ pos0, uint8_t *__restrict__ sp,
                                uint8_t *__restrict__ best_sp, double *__restrict__ pos,
                                double *__restrict__ best_pos) {
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (tid < natoms_pad) {
        const uint8_t v = tid < natoms ? (uint8_t)sp0[tid] : 0;
        sp[(size_t)s * natoms_pad + tid] = v;
        best_sp[(size_t)s * natoms_pad + tid] = v;
    }
    if (tid < 3L * natoms) {
        pos[(size_t)s * natoms * 3 + tid] = pos0[tid];
        best_pos[(size_t)s * natoms * 3 + tid] = pos0[tid];
    }
}

__global__ void k_ps_types(int nrep, int npairs, int natoms_pad, int nsp, const uint8_t *__restrict__ sp,
                           const int *__restrict__ iidx, const int *__restrict__ jidx, int *__restrict__ pt) {
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)nrep * npairs) return;
    const int s = (int)(tid / npairs), p = (int)(tid % npairs);
    const uint8_t *S = sp + (size_t)s * natoms_pad;
    pt[tid] = (int)S[iidx[p]] * nsp + (int)S[jidx[p]];
}

__global__ void k_ps_copy_rows(int nrep, long n, const double *__restrict__ src, double *__restrict__ dst) {
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < (long)nrep * n) dst[tid] = src[tid % n];
}

__global__ void k_ps_mark(int nrep, const double *__restrict__ best_c2, const long long *__restrict__ acc_rung,
                          double *__restrict__ out) {
    double b = __builtin_huge_val();
    long long a = 0;
    for (int s = threadIdx.x; s < nrep; s += 64) {
        b = fmin(b, best_c2[s]);
        a += acc_rung[s];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        b = fmin(b, __shfl_xor(b, m));
        a += __shfl_xor(a, m);
    }
    if (threadIdx.x == 0) {
        out[0] = b;
        out[1] = (double)a;
        out[2] = (double)wall_clock64();
    }
}

}  // namespace

// ================================================================================ host
struct chm_psampler {
    chm_ctx *ctx = nullptr;
    int nrep = 0, natoms = 0, natoms_pad = 0, npairs = 0, ncoord = 0, nsp = 0, nk = 0, nshell = 0;
    int nkw = 0, nwin = 0, ncol_e = 0, magnitude = 0, profile = 1, maxaff = 0, msmax = 1;
    double nfit = 1.0, swap_frac = 0.75, disp_sigma = 0.015, max_disp = 0.30, ebeta = 0.0;
    int has_prior = 0, has_ms = 0;
    int k0 = 0, nchunk = 0;          // first k > 0; 64-point chunks of the k > 0 points
    int kb0 = 0, kb1 = 0;            // k rows of the window transform that are not all zero
    Geo g{};
    size_t lds_delta = 0, lds_prop = 0;
    uint64_t seed = 0;
    bool inited = false;
    long long rec_base = 0;
    long long hstep = 0;   // steps completed (host mirror of the device counters)
    int log_cap = 0, mark_cap = 0;
    double cnt[chm::kMaxNsp] = {0};
    chm::DevBuf rev, ideal, etab, Te, Te_raw, Rx, tgt, wgt, q0, ms_sten, tms;
    int nbc = 1;   // 16-column blocks per edge of the transform
    chm::DevBuf sp, pos, r, pt, chi, c2, best_c2, best_sp, best_pos, rung_of, slot_of, beta, beta_slot,
        stepc, iex, acc_rung, exch_att, exch_acc;
    chm::DevBuf mv, aff, dcount, dms, trial, trialT, sums, log, log_c2, marks, tmp, tmp2;
    int nsplit = 1;   // workgroups per slot of the proposal (MS gather share)
    int ms_groups = 1;   // entry groups of the MS product (partials added by the delta kernel)
    static int ms_groups_for(int nblk) { return std::max(4, (nblk + kMsRowsMax - 1) / kMsRowsMax); }
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
};

namespace {

void ps_drop_graph(chm_psampler *S) {
    if (S->gexec) hipGraphExecDestroy(S->gexec);
    if (S->graph) hipGraphDestroy(S->graph);
    S->gexec = nullptr;
    S->graph = nullptr;
    S->graph_steps = 0;
}

PsArgs prop_args(chm_psampler *S) {
    chm_ctx *c = S->ctx;
    PsArgs A{};
    A.nrep = S->nrep;
    A.natoms = S->natoms;
    A.natoms_pad = S->natoms_pad;
    A.ncoord = S->ncoord;
    A.npairs = S->npairs;
    A.nsp = S->nsp;
    A.nk = S->nk;
    A.nshell = S->nshell;
    A.maxaff = S->maxaff;
    A.msmax = S->msmax;
    A.seed = S->seed;
    A.swap_frac = S->swap_frac;
    A.disp_sigma = S->disp_sigma;
    A.max_disp = S->max_disp;
    A.rung_of = S->rung_of.as<int>();
    A.stepc = S->stepc.as<uint64_t>();
    A.c2 = S->c2.as<double>();
    A.best_c2 = S->best_c2.as<double>();
    A.sp = S->sp.as<uint8_t>();
    A.pos = S->pos.as<double>();
    A.ideal = S->ideal.as<double>();
    A.rev = S->rev.as<int>();
    A.iidx = c->i_idx.as<int>();
    A.jidx = c->j_idx.as<int>();
    A.shell = c->shell.as<int>();
    A.shift = c->shift.as<double>();
    A.sig2 = c->sig2.as<double>();
    A.r = S->r.as<double>();
    A.pt = S->pt.as<int>();
    A.etab = S->has_prior ? S->etab.as<double>() : nullptr;
    A.mv = S->mv.as<PsMove>();
    A.aff = S->aff.as<AffRec>();
    A.dcount = S->has_ms ? S->dcount.as<int>() : nullptr;
    A.has_ms = S->has_ms;
    A.g = S->g;
    A.ms_sten = S->ms_sten.as<int4>();
    return A;
}

PsAcceptArgs accept_args(chm_psampler *S) {
    PsAcceptArgs A{};
    A.nrep = S->nrep;
    A.natoms = S->natoms;
    A.natoms_pad = S->natoms_pad;
    A.npairs = S->npairs;
    A.nsp = S->nsp;
    A.nk = S->nk;
    A.nkw = S->nkw;
    A.nwin = S->nwin;
    A.ncol_e = S->ncol_e;
    A.maxaff = S->maxaff;
    A.magnitude = S->magnitude;
    A.profile = S->profile;
    A.nfit = S->nfit;
    A.ebeta = S->ebeta;
    A.has_prior = S->has_prior;
    A.seed = S->seed;
    A.rung_of = S->rung_of.as<int>();
    A.stepc = S->stepc.as<uint64_t>();
    A.beta_slot = S->beta_slot.as<double>();
    A.sums = S->sums.as<double>();
    A.q0 = S->q0.as<double>();
    A.nbc = S->nbc;
    A.mv = S->mv.as<PsMove>();
    A.aff = S->aff.as<AffRec>();
    A.trial = S->trial.as<double>();
    A.sp = S->sp.as<uint8_t>();
    A.pos = S->pos.as<double>();
    A.r = S->r.as<double>();
    A.chi = S->chi.as<double>();
    A.c2 = S->c2.as<double>();
    A.best_c2 = S->best_c2.as<double>();
    A.best_pos = S->best_pos.as<double>();
    A.pt = S->pt.as<int>();
    A.best_sp = S->best_sp.as<uint8_t>();
    A.acc_rung = S->acc_rung.as<long long>();
    A.dcount = S->has_ms ? S->dcount.as<int>() : nullptr;
    A.nrow_tot = S->g.nrow_tot;
    A.log = S->log_cap ? S->log.as<int4>() : nullptr;
    A.log_c2 = S->log_cap ? S->log_c2.as<double>() : nullptr;
    A.log_cap = S->log_cap;
    A.rec_base = S->rec_base;
    return A;
}

template <int NSP>
int launch_delta(chm_psampler *S, hipStream_t st) {
    chm_ctx *c = S->ctx;
    PsDeltaArgs A{};
    A.nrep = S->nrep;
    A.npairs = S->npairs;
    A.nk = S->nk;
    A.k0 = S->k0;
    A.maxaff = S->maxaff;
    A.msmax = S->msmax;
    A.nblk = std::max(1, S->g.nblk);
    A.s02 = c->s02;
    A.mv = S->mv.as<PsMove>();
    A.aff = S->aff.as<AffRec>();
    A.ftab = c->f.as<double>();
    A.dtab = c->delta.as<double>();
    A.lamtab = c->lam.as<double>();
    A.ktab = c->k.as<double>();
    A.drtab = c->drtab.as<double>();
    A.c3tab = c->c3tab.as<double>();
    A.s2tab = c->s2tab.as<double>();
    A.dms = S->has_ms ? S->dms.as<double>() : nullptr;
    A.ngrp = S->ms_groups;
    A.chi = S->chi.as<double>();
    A.trial = S->trial.as<double>();
    A.trialT = S->trialT.as<double>();
    k_ps_delta<NSP><<<dim3(S->nchunk, S->nrep), 64 * kDeltaWaves, S->lds_delta, st>>>(A, chm::kPolyCoefs);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

template <int NSP>
int delta_attr(chm_psampler *S) {
    S->lds_delta = sizeof(double) * (NSP + NSP * NSP + kDeltaWaves * NSP) * 64 + sizeof(TermRec) * 2 * S->maxaff;
    CHM_REQUIRE(S->lds_delta <= 160 * 1024, CHM_E_ARG, "coordination too large for the LDS of the delta kernel");
    CHM_HIP(hipFuncSetAttribute((const void *)k_ps_delta<NSP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)S->lds_delta));
    return CHM_OK;
}

// trial chi(R) residual sums of nconf k-major chi rows
int launch_transform(chm_psampler *S, hipStream_t st, int nconf, const double *trialT, double *sums) {
    TfArgs T{};
    T.nrep = nconf;
    T.nk = S->nk;
    T.nsp = S->nsp;
    T.kb0 = S->kb0;
    T.kb1 = S->kb1;
    T.ncol_e = S->ncol_e;
    T.nbc = S->nbc;
    T.magnitude = S->magnitude;
    T.trialT = trialT;
    T.Te = S->Te.as<double>();
    T.Rx = S->Rx.as<double>();
    T.tgt = S->tgt.as<double>();
    T.wgt = S->wgt.as<double>();
    T.sums = sums;
    k_ps_transform<<<dim3(S->nbc, S->nsp, (nconf + 31) / 32), 256, 0, st>>>(T);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int launch_step(chm_psampler *S, hipStream_t st) {
    {
        PsArgs A = prop_args(S);
        k_ps_propose<<<dim3(S->nsplit, S->nrep), kPropThreads, S->lds_prop, st>>>(A);
        CHM_LAUNCH_CHECK();
    }
    if (S->has_ms) {
        const Geo &g = S->g;
        k_ps_msgemm<<<dim3((S->nk + 31) / 32, S->nsp, ((S->nrep + 31) / 32) * S->ms_groups), 256, 0, st>>>(
            S->nrep, S->nk, S->nsp, g.nblk, g.rstride, g.nss_rows, g.nrow_tot, S->ms_groups, S->dcount.as<int>(),
            S->tms.as<double>(), S->dms.as<double>());
        CHM_LAUNCH_CHECK();
    }
    int rc = CHM_E_ARG;
    switch (S->nsp) {
        case 2: rc = launch_delta<2>(S, st); break;
        case 3: rc = launch_delta<3>(S, st); break;
        case 4: rc = launch_delta<4>(S, st); break;
        case 5: rc = launch_delta<5>(S, st); break;
        case 6: rc = launch_delta<6>(S, st); break;
    }
    if (rc) return rc;
    if ((rc = launch_transform(S, st, S->nrep, S->trialT.as<double>(), S->sums.as<double>()))) return rc;
    k_ps_accept<<<dim3(kAccParts, S->nrep), kAccThreads, 0, st>>>(accept_args(S));
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int launch_exchange(chm_psampler *S, hipStream_t st) {
    k_ps_exchange<<<1, 64, 0, st>>>(S->nrep, S->beta.as<double>(), S->c2.as<double>(), S->slot_of.as<int>(),
                                    S->rung_of.as<int>(), S->beta_slot.as<double>(), S->iex.as<uint64_t>(),
                                    S->exch_att.as<long long>(), S->exch_acc.as<long long>(), S->seed,
                                    S->nfit);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

}  // namespace

extern "C" int chm_psampler_create(chm_ctx *ctx, int nrep, const double *ideal, const int32_t *rev,
                                   int nkw, int nwin, const double *Te, const double *Rx,
                                   const double *target, const double *sigma, int magnitude,
                                   int profile_scale, chm_psampler **out) {
    CHM_REQUIRE(ctx && out && ideal && rev && Te && target && sigma, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(ctx->npairs > 0 && ctx->nsp >= 2, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(nrep >= 1 && nkw >= 1 && nwin >= 1, CHM_E_ARG, "bad nrep / window");
    const int natoms = ctx->natoms, npairs = ctx->npairs;
    CHM_REQUIRE(npairs % natoms == 0, CHM_E_ARG, "expected uniform coordination");
    const int ncoord = npairs / natoms;
    CHM_REQUIRE(4 * ncoord <= 4096, CHM_E_ARG, "coordination too large");
    for (int p = 0; p < npairs; ++p) CHM_REQUIRE(rev[p] >= 0 && rev[p] < npairs, CHM_E_ARG, "reverse index out of range");
    CHM_HIP(hipSetDevice(ctx->device));
    // the pair list must be i-sorted with ncoord pairs per site, and rev its reverse
    {
        std::vector<int> ii(npairs), jj(npairs);
        CHM_HIP(hipMemcpy(ii.data(), ctx->i_idx.p, sizeof(int) * npairs, hipMemcpyDeviceToHost));
        CHM_HIP(hipMemcpy(jj.data(), ctx->j_idx.p, sizeof(int) * npairs, hipMemcpyDeviceToHost));
        for (int p = 0; p < npairs; ++p) {
            CHM_REQUIRE(ii[p] == p / ncoord, CHM_E_ARG, "pairs must be i-sorted with uniform coordination");
            CHM_REQUIRE(ii[rev[p]] == jj[p] && jj[rev[p]] == ii[p], CHM_E_ARG, "rev is not the reverse index");
        }
    }
    auto *S = new chm_psampler();
    S->ctx = ctx;
    S->nrep = nrep;
    S->natoms = natoms;
    S->natoms_pad = (natoms + 15) & ~15;
    S->npairs = npairs;
    S->ncoord = ncoord;
    S->nsp = ctx->nsp;
    S->nk = ctx->nk;
    S->nshell = ctx->nshell;
    S->nkw = nkw;
    S->nwin = nwin;
    S->ncol_e = nkw * nwin * 2;
    S->magnitude = magnitude ? 1 : 0;
    S->profile = profile_scale ? 1 : 0;
    S->maxaff = 4 * ncoord;
    S->nfit = (double)S->nsp * nkw * nwin * (magnitude ? 1 : 2);
    auto fail = [&](int rc) {
        delete S;
        return rc;
    };
    int rc = CHM_OK;
    while (S->k0 < S->nk && !(ctx->h_k[S->k0] > 0.0)) ++S->k0;
    if (S->k0 >= S->nk) return fail((chm::set_error("no k point above 0"), CHM_E_ARG));
    S->nchunk = (S->nk - S->k0 + 63) / 64;
    S->lds_prop = (size_t)S->natoms_pad;
    if (S->lds_prop > 120 * 1024)
        return fail((chm::set_error("at most 122880 sites (the proposal stages the species in LDS)"), CHM_E_ARG));
    {
        hipError_t e = hipFuncSetAttribute((const void *)k_ps_propose, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)S->lds_prop);
        if (e != hipSuccess) return fail((chm::set_error(hipGetErrorString(e)), CHM_E_HIP));
    }
    switch (S->nsp) {
        case 2: rc = delta_attr<2>(S); break;
        case 3: rc = delta_attr<3>(S); break;
        case 4: rc = delta_attr<4>(S); break;
        case 5: rc = delta_attr<5>(S); break;
        case 6: rc = delta_attr<6>(S); break;
    }
    if (rc) return fail(rc);
    auto up = [&](chm::DevBuf &b, const void *h, size_t n) -> int {
        int r = b.ensure(n);
        if (r) return r;
        CHM_HIP(hipMemcpyAsync(b.p, h, n, hipMemcpyHostToDevice, ctx->stream));
        return CHM_OK;
    };
    const size_t npt = (size_t)S->nsp * nkw * nwin;
    std::vector<double> wgt(npt), q0(2 * (size_t)S->nsp, 0.0);
    for (size_t q = 0; q < npt; ++q) {
        if (!(sigma[q] > 0.0)) return fail((chm::set_error("sigma must be positive"), CHM_E_ARG));
        wgt[q] = 1.0 / (sigma[q] * sigma[q]);
        const size_t e = q / ((size_t)nkw * nwin);
        const double t2 = magnitude ? target[q] * target[q]
                                    : target[2 * q] * target[2 * q] + target[2 * q + 1] * target[2 * q + 1];
        q0[2 * e] += t2;
        q0[2 * e + 1] += t2 * wgt[q];
    }
    S->nbc = (S->ncol_e + kTfCols - 1) / kTfCols;
    std::vector<double> rx((size_t)S->nsp * S->ncol_e, 0.0);
    if (Rx) std::copy(Rx, Rx + rx.size(), rx.begin());
    const size_t te_n = (size_t)S->nsp * S->nk * S->ncol_e;
    // the k rows of the transform that are not all zero (the window vanishes outside)
    S->kb0 = S->nk;
    S->kb1 = 0;
    for (int e = 0; e < S->nsp; ++e)
        for (int k = 0; k < S->nk; ++k) {
            const double *row = Te + ((size_t)e * S->nk + k) * S->ncol_e;
            bool nz = false;
            for (int c = 0; c < S->ncol_e && !nz; ++c) nz = row[c] != 0.0;
            if (nz) {
                S->kb0 = std::min(S->kb0, k);
                S->kb1 = std::max(S->kb1, k + 1);
            }
        }
    if (S->kb1 <= S->kb0) {
        S->kb0 = 0;
        S->kb1 = 1;
    }

    if ((rc = up(S->rev, rev, sizeof(int) * npairs)) || (rc = up(S->ideal, ideal, sizeof(double) * 3 * natoms)) ||
        (rc = up(S->Te_raw, Te, sizeof(double) * te_n)) || (rc = S->Te.ensure(sizeof(double) * te_n)) ||
        (rc = up(S->Rx, rx.data(), sizeof(double) * rx.size())) ||
        (rc = up(S->tgt, target, sizeof(double) * npt * (magnitude ? 1 : 2))) ||
        (rc = up(S->wgt, wgt.data(), sizeof(double) * npt)) || (rc = up(S->q0, q0.data(), sizeof(double) * q0.size())))
        return fail(rc);
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    const size_t nr = nrep;
    if ((rc = S->sp.ensure(nr * S->natoms_pad)) || (rc = S->best_sp.ensure(nr * S->natoms_pad)) ||
        (rc = S->pos.ensure(sizeof(double) * 3 * nr * natoms)) ||
        (rc = S->best_pos.ensure(sizeof(double) * 3 * nr * natoms)) ||
        (rc = S->r.ensure(sizeof(double) * nr * npairs)) || (rc = S->pt.ensure(sizeof(int) * nr * npairs)) ||
        (rc = S->chi.ensure(sizeof(double) * nr * S->nsp * S->nk)) ||
        (rc = S->trial.ensure(sizeof(double) * nr * S->nsp * S->nk)) ||
        (rc = S->c2.ensure(sizeof(double) * nr)) || (rc = S->best_c2.ensure(sizeof(double) * nr)) ||
        (rc = S->rung_of.ensure(sizeof(int) * nr)) || (rc = S->slot_of.ensure(sizeof(int) * nr)) ||
        (rc = S->beta.ensure(sizeof(double) * nr)) || (rc = S->beta_slot.ensure(sizeof(double) * nr)) ||
        (rc = S->stepc.ensure(sizeof(uint64_t) * nr)) || (rc = S->iex.ensure(sizeof(uint64_t))) ||
        (rc = S->acc_rung.ensure(sizeof(long long) * nr)) ||
        (rc = S->exch_att.ensure(sizeof(long long) * nr)) || (rc = S->exch_acc.ensure(sizeof(long long) * nr)) ||
        (rc = S->mv.ensure(sizeof(PsMove) * nr)) || (rc = S->aff.ensure(sizeof(AffRec) * nr * S->maxaff)) ||

        (rc = S->trialT.ensure(sizeof(double) * nr * S->nsp * S->nk)) ||
        (rc = S->sums.ensure(sizeof(double) * nr * S->nsp * S->nbc * 4)) ||
        (rc = S->tmp.ensure(sizeof(double) * nr)))
        return fail(rc);
    CHM_HIP(hipMemsetAsync(S->acc_rung.p, 0, S->acc_rung.bytes, ctx->stream));
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    chm::ctx_retain(ctx);
    *out = S;
    return CHM_OK;
}

extern "C" int chm_psampler_destroy(chm_psampler *S) {
    if (!S) return CHM_OK;
    hipSetDevice(S->ctx->device);
    hipStreamSynchronize(S->ctx->stream);
    ps_drop_graph(S);
    chm_ctx *c = S->ctx;
    delete S;
    chm::ctx_release(c);
    return CHM_OK;
}

extern "C" int chm_psampler_set_moves(chm_psampler *S, double swap_frac, double disp_sigma, double max_disp) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    CHM_REQUIRE(swap_frac >= 0.0 && swap_frac <= 1.0 && disp_sigma >= 0.0 && max_disp > 0.0, CHM_E_ARG,
                "swap_frac in [0, 1], disp_sigma >= 0, max_disp > 0");
    if (swap_frac != S->swap_frac || disp_sigma != S->disp_sigma || max_disp != S->max_disp) {
        CHM_HIP(hipSetDevice(S->ctx->device));
        CHM_HIP(hipStreamSynchronize(S->ctx->stream));
        ps_drop_graph(S);
    }
    S->swap_frac = swap_frac;
    S->disp_sigma = disp_sigma;
    S->max_disp = max_disp;
    return CHM_OK;
}

extern "C" int chm_psampler_set_prior(chm_psampler *S, const double *etab, double ebeta) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    ps_drop_graph(S);
    if (!etab) {
        S->has_prior = 0;
        S->ebeta = 0.0;
        return CHM_OK;
    }
    const size_t n = (size_t)S->nshell * S->nsp * S->nsp;
    int rc = S->etab.ensure(sizeof(double) * n);
    if (rc) return rc;
    CHM_HIP(hipMemcpy(S->etab.p, etab, sizeof(double) * n, hipMemcpyHostToDevice));
    S->has_prior = 1;
    S->ebeta = ebeta;
    return CHM_OK;
}

extern "C" int chm_psampler_set_ms(chm_psampler *S, const chm_geom *gm, const int32_t *site_xyz,
                                   const int8_t *stencil, const double *tms) {
    CHM_REQUIRE(S && gm && site_xyz && stencil && tms, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(gm->natoms == S->natoms && gm->nsp == S->nsp && gm->nblk > 0 && gm->nstencil > 0,
                CHM_E_ARG, "geometry does not match the sampler");
    Geo g;
    int rc = geo_from(gm, site_xyz, g);
    if (rc) return rc;
    if ((rc = check_stencil(g, stencil))) return rc;
    CHM_REQUIRE(sizeof(int) * (size_t)g.nrow_tot <= 64 * 1024, CHM_E_ARG, "too many MS entries for LDS");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    ps_drop_graph(S);
    const int msmax = g.nsp * g.nblk;
    if ((rc = S->ms_sten.ensure(16 * (size_t)g.nstencil)) ||
        (rc = S->tms.ensure(sizeof(double) * (size_t)g.nsp * g.nblk * S->nk)) ||
        (rc = S->dcount.ensure(sizeof(int) * (size_t)S->nrep * g.nrow_tot)) ||
        (rc = S->dms.ensure(sizeof(double) * (size_t)S->ms_groups_for(g.nblk) * S->nrep * g.nsp * S->nk)))
        return rc;
    S->ms_groups = S->ms_groups_for(g.nblk);
    CHM_HIP(hipMemset(S->dcount.p, 0, S->dcount.bytes));
    // proposal workgroups per slot: about 3 gather items per lane, at most 16
    S->nsplit = 1 + std::max(1, std::min(16, (2 * g.nstencil + 3 * kPropThreads - 1) / (3 * kPropThreads)));
    CHM_HIP(hipMemcpy(S->ms_sten.p, stencil, 16 * (size_t)g.nstencil, hipMemcpyHostToDevice));
    CHM_HIP(hipMemcpy(S->tms.p, tms, sizeof(double) * (size_t)g.nsp * g.nblk * S->nk, hipMemcpyHostToDevice));
    S->g = g;
    S->msmax = msmax;
    S->has_ms = 1;
    switch (S->nsp) {
        case 2: rc = delta_attr<2>(S); break;
        case 3: rc = delta_attr<3>(S); break;
        case 4: rc = delta_attr<4>(S); break;
        case 5: rc = delta_attr<5>(S); break;
        case 6: rc = delta_attr<6>(S); break;
    }
    if (rc) return rc;
    S->lds_prop = sizeof(int) * (size_t)((g.nrow_tot + 3) & ~3) + (size_t)S->natoms_pad;
    CHM_REQUIRE(S->lds_prop <= 160 * 1024, CHM_E_ARG, "sites and MS entries too many for the LDS of the proposal");
    CHM_HIP(hipFuncSetAttribute((const void *)k_ps_propose, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)S->lds_prop));
    return CHM_OK;
}

// Start every replica at (species0, pos0): distances and types of every pair, chi from the
// full path sum (+ MS counts @ table, computed by the caller and passed as chi_ms, or null),
// chi^2, best = start; ladder beta (rung order); counters and the exchange stream reset.
extern "C" int chm_psampler_init(chm_psampler *S, const int32_t *species0, const double *pos0,
                                 const double *chi_ms, const double *beta, uint64_t seed) {
    CHM_REQUIRE(S && species0 && pos0 && beta, CHM_E_ARG, "bad arguments");
    chm_ctx *c = S->ctx;
    const int nsp = S->nsp, natoms = S->natoms, nr = S->nrep;
    std::vector<int> cnt(nsp, 0);
    for (int q = 0; q < natoms; ++q) {
        CHM_REQUIRE(species0[q] >= 0 && species0[q] < nsp, CHM_E_ARG, "species index out of range");
        ++cnt[species0[q]];
    }
    for (int q = 0; q < 3 * natoms; ++q) CHM_REQUIRE(std::isfinite(pos0[q]), CHM_E_ARG, "positions must be finite");
    CHM_HIP(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    CHM_HIP(hipStreamSynchronize(st));
    ps_drop_graph(S);
    S->seed = seed;
    int rc;
    // transform matrix divided by each edge's absorber count (composition is conserved)
    {
        std::vector<double> te((size_t)nsp * S->nk * S->ncol_e);
        CHM_HIP(hipMemcpy(te.data(), S->Te_raw.p, sizeof(double) * te.size(), hipMemcpyDeviceToHost));
        const size_t per = (size_t)S->nk * S->ncol_e;
        for (int e = 0; e < nsp; ++e) {
            S->cnt[e] = std::max(cnt[e], 1);
            const double inv = 1.0 / S->cnt[e];
            for (size_t q = 0; q < per; ++q) te[e * per + q] *= inv;
        }
        CHM_HIP(hipMemcpy(S->Te.p, te.data(), sizeof(double) * te.size(), hipMemcpyHostToDevice));
    }
    chm::DevBuf sp0, p0;
    if ((rc = sp0.ensure(sizeof(int) * natoms)) || (rc = p0.ensure(sizeof(double) * 3 * natoms))) return rc;
    CHM_HIP(hipMemcpy(sp0.p, species0, sizeof(int) * natoms, hipMemcpyHostToDevice));
    CHM_HIP(hipMemcpy(p0.p, pos0, sizeof(double) * 3 * natoms, hipMemcpyHostToDevice));
    {
        const long m = std::max((long)S->natoms_pad, 3L * natoms);
        k_ps_init_slots<<<dim3((unsigned)((m + 255) / 256), nr), 256, 0, st>>>(
            nr, natoms, S->natoms_pad, sp0.as<int>(), p0.as<double>(), S->sp.as<uint8_t>(),
            S->best_sp.as<uint8_t>(), S->pos.as<double>(), S->best_pos.as<double>());
        CHM_LAUNCH_CHECK();
    }
    // one configuration's distances and chi, copied to every slot
    if ((rc = chm::distances_into(c, st, 1, p0.as<double>(), S->r.as<double>()))) return rc;
    {
        const long tot = (long)nr * S->npairs;
        k_ps_types<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, st>>>(
            nr, S->npairs, S->natoms_pad, nsp, S->sp.as<uint8_t>(), c->i_idx.as<int>(), c->j_idx.as<int>(),
            S->pt.as<int>());
        CHM_LAUNCH_CHECK();
    }
    chm::ChiShape sh1;
    if ((rc = chm::chi_shape(c, 1, sh1))) return rc;
    chm::DevBuf part, chi1;
    if ((rc = part.ensure(sizeof(double) * (size_t)sh1.nwg * nsp * S->nk)) ||
        (rc = chi1.ensure(sizeof(double) * nsp * S->nk)))
        return rc;
    if ((rc = chm::chi_into(c, st, 1, CHM_PREC_F64, sh1, S->r.as<double>(), S->pt.as<int>(), part.as<double>(),
                            chi1.as<double>())))
        return rc;
    CHM_HIP(hipStreamSynchronize(st));
    {
        std::vector<double> h((size_t)nsp * S->nk);
        CHM_HIP(hipMemcpy(h.data(), chi1.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
        if (chi_ms)
            for (size_t q = 0; q < h.size(); ++q) h[q] += chi_ms[q];
        CHM_HIP(hipMemcpy(chi1.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
    }
    const long nkk = (long)nsp * S->nk;
    k_ps_copy_rows<<<dim3((unsigned)((nr * nkk + 255) / 256)), 256, 0, st>>>(nr, nkk, chi1.as<double>(),
                                                                              S->chi.as<double>());
    CHM_LAUNCH_CHECK();
    k_ps_copy_rows<<<dim3((unsigned)(((long)nr * S->npairs + 255) / 256)), 256, 0, st>>>(
        nr, S->npairs, S->r.as<double>(), S->r.as<double>());
    CHM_LAUNCH_CHECK();
    // chi^2 of the start: the step's transform and residual sums on chi
    {
        const long tot = (long)nr * nsp * S->nk;
        k_ps_to_kmajor<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, st>>>(nr, nsp, S->nk, S->chi.as<double>(),
                                                                            S->trialT.as<double>());
        CHM_LAUNCH_CHECK();
    }
    if ((rc = launch_transform(S, st, nr, S->trialT.as<double>(), S->sums.as<double>()))) return rc;
    k_ps_chi2<<<nr, 64, 0, st>>>(S->sums.as<double>(), nsp, S->nbc, S->q0.as<double>(), S->profile, S->nfit,
                                 S->c2.as<double>());
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(S->best_c2.p, S->c2.p, sizeof(double) * nr, hipMemcpyDeviceToDevice, st));
    std::vector<int> ident(nr);
    for (int q = 0; q < nr; ++q) ident[q] = q;
    CHM_HIP(hipMemcpyAsync(S->rung_of.p, ident.data(), sizeof(int) * nr, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->slot_of.p, ident.data(), sizeof(int) * nr, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->beta.p, beta, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->beta_slot.p, beta, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemsetAsync(S->stepc.p, 0, S->stepc.bytes, st));
    CHM_HIP(hipMemsetAsync(S->iex.p, 0, S->iex.bytes, st));
    CHM_HIP(hipMemsetAsync(S->acc_rung.p, 0, S->acc_rung.bytes, st));
    CHM_HIP(hipMemsetAsync(S->exch_att.p, 0, S->exch_att.bytes, st));
    CHM_HIP(hipMemsetAsync(S->exch_acc.p, 0, S->exch_acc.bytes, st));
    if (S->has_ms) CHM_HIP(hipMemsetAsync(S->dcount.p, 0, S->dcount.bytes, st));
    CHM_HIP(hipStreamSynchronize(st));   // sp0, p0, part, chi1 go out of scope
    S->rec_base = 0;
    S->hstep = 0;
    S->inited = true;
    return CHM_OK;
}

// nsteps steps; replica exchange after every exchange_every-th completed step (0: never).
// Exchange periods run from one captured graph.  Asynchronous: returns once launched.
extern "C" int chm_psampler_run(chm_psampler *S, int nsteps, int exchange_every) {
    CHM_REQUIRE(S && S->inited, CHM_E_STATE, "init first");
    CHM_REQUIRE(nsteps >= 0 && exchange_every >= 0, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    int rc = CHM_OK;
    long long &cur = S->hstep;   // steps done so far (every slot advances together)
    const int E = exchange_every;
    int left = nsteps;
    while (left > 0) {
        int to_ex = E ? (int)(E - cur % E) : left;   // steps until the next exchange
        if (E && to_ex == E && left >= E) {
            if (!S->gexec || S->graph_steps != E) {
                // one exchange period per graph, captured on the sampler's own stream
                ps_drop_graph(S);
                CHM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                for (int t = 0; t < E && rc == CHM_OK; ++t) rc = launch_step(S, st);
                if (rc == CHM_OK) rc = launch_exchange(S, st);
                hipGraph_t gph = nullptr;
                hipError_t e = hipStreamEndCapture(st, &gph);
                if (rc) {
                    if (gph) hipGraphDestroy(gph);
                    return rc;
                }
                if (e != hipSuccess) {
                    chm::set_error(std::string("graph capture: ") + hipGetErrorString(e));
                    return CHM_E_HIP;
                }
                S->graph = gph;
                CHM_HIP(hipGraphInstantiate(&S->gexec, gph, nullptr, nullptr, 0));
                S->graph_steps = E;
            }
            while (left >= E) {
                CHM_HIP(hipGraphLaunch(S->gexec, st));
                left -= E;
                cur += E;
            }
            continue;
        }
        const int n = std::min(left, to_ex);
        for (int t = 0; t < n; ++t)
            if ((rc = launch_step(S, st))) return rc;
        left -= n;
        cur += n;
        if (E && cur % E == 0)
            if ((rc = launch_exchange(S, st))) return rc;
    }
    return CHM_OK;
}

extern "C" int chm_psampler_exchange(chm_psampler *S) {
    CHM_REQUIRE(S && S->inited, CHM_E_STATE, "init first");
    CHM_HIP(hipSetDevice(S->ctx->device));
    int rc = launch_exchange(S, S->ctx->stream);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    return CHM_OK;
}

extern "C" int chm_psampler_set_beta(chm_psampler *S, const double *beta) {
    CHM_REQUIRE(S && S->inited && beta, CHM_E_ARG, "bad arguments");
    const int nr = S->nrep;
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    std::vector<int> slot_of(nr);
    CHM_HIP(hipMemcpyAsync(slot_of.data(), S->slot_of.p, sizeof(int) * nr, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    std::vector<double> bs(nr);
    for (int q = 0; q < nr; ++q) bs[slot_of[q]] = beta[q];
    CHM_HIP(hipMemcpyAsync(S->beta.p, beta, sizeof(double) * nr, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->beta_slot.p, bs.data(), sizeof(double) * nr, hipMemcpyHostToDevice, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_psampler_set_recording(chm_psampler *S, int log_steps) {
    CHM_REQUIRE(S && log_steps >= 0, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    ps_drop_graph(S);
    int rc;
    if (log_steps > 0 && ((rc = S->log.ensure(sizeof(int4) * (size_t)log_steps * S->nrep)) ||
                          (rc = S->log_c2.ensure(sizeof(double) * (size_t)log_steps * S->nrep))))
        return rc;
    if (log_steps > 0) {
        CHM_HIP(hipMemset(S->log.p, 0xFF, sizeof(int4) * (size_t)log_steps * S->nrep));
        CHM_HIP(hipMemset(S->log_c2.p, 0, sizeof(double) * (size_t)log_steps * S->nrep));
    }
    S->rec_base = S->inited ? S->hstep : 0;   // init restarts the count
    S->log_cap = log_steps;
    return CHM_OK;
}

namespace {
__global__ void k_ps_gather_rows(int nrep, long row, const int *__restrict__ slot_of, const uint8_t *__restrict__ src8,
                                 const double *__restrict__ srcd, const int *__restrict__ srci, long pitch,
                                 void *__restrict__ out, int kind) {
    // kind 0: u8 -> int32, 1: double, 2: int32; rung order
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int rg = blockIdx.y;
    if (tid >= row) return;
    const int s = slot_of[rg];
    if (kind == 0)
        reinterpret_cast<int *>(out)[(size_t)rg * row + tid] = src8[(size_t)s * pitch + tid];
    else if (kind == 1)
        reinterpret_cast<double *>(out)[(size_t)rg * row + tid] = srcd[(size_t)s * pitch + tid];
    else
        reinterpret_cast<int *>(out)[(size_t)rg * row + tid] = srci[(size_t)s * pitch + tid];
}
}  // namespace

// what: 0 species (nrep, natoms) int32, 1 chi^2 (nrep), 2 best chi^2 (nrep), 3 best species,
// 4 positions (nrep, natoms, 3), 5 best positions, 6 chi (nrep, nsp, nk), 7 r (nrep, npairs),
// 8 pair types (nrep, npairs) int32, 9 accepted per rung (nrep) int64, 10 exchange attempts
// and accepts (2, nrep - 1) int64, 11 completed steps (1) int64, 12 move log (log_cap, nrep, 4)
// int32, 13 logged trial chi^2 (log_cap, nrep), 14 rung_of (nrep) int32, 15 energy-prior
// change of each replica's last proposal (nrep).  Rung order except
// where noted (the log is indexed by rung already; rung_of by slot).
extern "C" int chm_psampler_read(chm_psampler *S, int what, void *out, int64_t nbytes) {
    const size_t bytes = nbytes < 0 ? 0 : (size_t)nbytes;
    CHM_REQUIRE(S && S->inited && out, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    const int nr = S->nrep;
    auto gather = [&](long row, const uint8_t *s8, const double *sd, const int *si, long pitch, int kind,
                      size_t esz) -> int {
        CHM_REQUIRE(bytes == esz * (size_t)row * nr, CHM_E_ARG, "output size does not match");
        int rc = S->tmp2.ensure(bytes);
        if (rc) return rc;
        k_ps_gather_rows<<<dim3((unsigned)((row + 255) / 256), nr), 256, 0, st>>>(
            nr, row, S->slot_of.as<int>(), s8, sd, si, pitch, S->tmp2.p, kind);
        CHM_LAUNCH_CHECK();
        CHM_HIP(hipMemcpyAsync(out, S->tmp2.p, bytes, hipMemcpyDeviceToHost, st));
        CHM_HIP(hipStreamSynchronize(st));
        return CHM_OK;
    };
    auto plain = [&](const chm::DevBuf &b, size_t n) -> int {
        CHM_REQUIRE(bytes == n, CHM_E_ARG, "output size does not match");
        CHM_HIP(hipMemcpyAsync(out, b.p, n, hipMemcpyDeviceToHost, st));
        CHM_HIP(hipStreamSynchronize(st));
        return CHM_OK;
    };
    switch (what) {
        case 0: return gather(S->natoms, S->sp.as<uint8_t>(), nullptr, nullptr, S->natoms_pad, 0, 4);
        case 1: return gather(1, nullptr, S->c2.as<double>(), nullptr, 1, 1, 8);
        case 2: return gather(1, nullptr, S->best_c2.as<double>(), nullptr, 1, 1, 8);
        case 3: return gather(S->natoms, S->best_sp.as<uint8_t>(), nullptr, nullptr, S->natoms_pad, 0, 4);
        case 4: return gather(3L * S->natoms, nullptr, S->pos.as<double>(), nullptr, 3L * S->natoms, 1, 8);
        case 5: return gather(3L * S->natoms, nullptr, S->best_pos.as<double>(), nullptr, 3L * S->natoms, 1, 8);
        case 6: return gather((long)S->nsp * S->nk, nullptr, S->chi.as<double>(), nullptr, (long)S->nsp * S->nk, 1, 8);
        case 7: return gather(S->npairs, nullptr, S->r.as<double>(), nullptr, S->npairs, 1, 8);
        case 8: return gather(S->npairs, nullptr, nullptr, S->pt.as<int>(), S->npairs, 2, 4);
        case 9: return plain(S->acc_rung, sizeof(long long) * nr);
        case 10: {
            const size_t m = (size_t)std::max(nr - 1, 1);
            CHM_REQUIRE(bytes == 2 * sizeof(long long) * m, CHM_E_ARG, "output size does not match");
            CHM_HIP(hipMemcpyAsync(out, S->exch_att.p, sizeof(long long) * m, hipMemcpyDeviceToHost, st));
            CHM_HIP(hipMemcpyAsync(reinterpret_cast<long long *>(out) + m, S->exch_acc.p, sizeof(long long) * m,
                                   hipMemcpyDeviceToHost, st));
            CHM_HIP(hipStreamSynchronize(st));
            return CHM_OK;
        }
        case 11: {
            CHM_REQUIRE(bytes == sizeof(long long), CHM_E_ARG, "output size does not match");
            CHM_HIP(hipMemcpyAsync(out, S->stepc.p, sizeof(long long), hipMemcpyDeviceToHost, st));
            CHM_HIP(hipStreamSynchronize(st));
            return CHM_OK;
        }
        case 12: return plain(S->log, sizeof(int4) * (size_t)S->log_cap * nr);
        case 13: return plain(S->log_c2, sizeof(double) * (size_t)S->log_cap * nr);
        case 14: return plain(S->rung_of, sizeof(int) * nr);
        case 15: {
            // energy-prior change of each slot's last proposal, rung order
            CHM_REQUIRE(bytes == sizeof(double) * nr, CHM_E_ARG, "output size does not match");
            std::vector<PsMove> mv(nr);
            std::vector<int> slot_of(nr);
            CHM_HIP(hipMemcpyAsync(mv.data(), S->mv.p, sizeof(PsMove) * nr, hipMemcpyDeviceToHost, st));
            CHM_HIP(hipMemcpyAsync(slot_of.data(), S->slot_of.p, sizeof(int) * nr, hipMemcpyDeviceToHost, st));
            CHM_HIP(hipStreamSynchronize(st));
            double *o = reinterpret_cast<double *>(out);
            for (int q = 0; q < nr; ++q) o[q] = mv[slot_of[q]].de;
            return CHM_OK;
        }
    }
    chm::set_error("unknown read code");
    return CHM_E_ARG;
}

extern "C" int chm_psampler_reserve_marks(chm_psampler *S, int n) {
    CHM_REQUIRE(S && n >= 0, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    if (n > S->mark_cap) {
        CHM_HIP(hipStreamSynchronize(S->ctx->stream));
        int rc = S->marks.ensure(sizeof(double) * 3 * (size_t)n);
        if (rc) return rc;
        S->mark_cap = n;
    }
    return CHM_OK;
}

extern "C" int chm_psampler_mark(chm_psampler *S, int idx) {
    CHM_REQUIRE(S && S->inited && idx >= 0 && idx < S->mark_cap, CHM_E_ARG, "mark index out of range");
    CHM_HIP(hipSetDevice(S->ctx->device));
    k_ps_mark<<<1, 64, 0, S->ctx->stream>>>(S->nrep, S->best_c2.as<double>(), S->acc_rung.as<long long>(),
                                            S->marks.as<double>() + 3 * (size_t)idx);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

extern "C" int chm_psampler_read_marks(chm_psampler *S, int n, double *out) {
    CHM_REQUIRE(S && out && n >= 0 && n <= S->mark_cap, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    int clk = 0;
    CHM_HIP(hipDeviceGetAttribute(&clk, hipDeviceAttributeWallClockRate, S->ctx->device));
    if (n) CHM_HIP(hipMemcpy(out, S->marks.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
    const double t0 = n ? out[2] : 0.0;
    for (int q = 0; q < n; ++q) out[3 * q + 2] = (out[3 * q + 2] - t0) / (clk * 1000.0);
    return CHM_OK;
}

// chi^2 of nconf unnormalised chi (nconf, nsp, nk) against the sampler's target, window
// and scale profiling (the value the chain holds for a state with that chi).
extern "C" int chm_psampler_chi2(chm_psampler *S, int nconf, const double *chi, double *out) {
    CHM_REQUIRE(S && S->inited && chi && out && nconf >= 1, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    const size_t nkk = (size_t)S->nsp * S->nk;
    chm::DevBuf in, inT, sums, c2;
    int rc;
    if ((rc = in.ensure(sizeof(double) * nconf * nkk)) || (rc = inT.ensure(sizeof(double) * nconf * nkk)) ||
        (rc = sums.ensure(sizeof(double) * nconf * S->nsp * S->nbc * 4)) || (rc = c2.ensure(sizeof(double) * nconf)))
        return rc;
    CHM_HIP(hipMemcpyAsync(in.p, chi, sizeof(double) * nconf * nkk, hipMemcpyHostToDevice, st));
    {
        const long tot = (long)nconf * nkk;
        k_ps_to_kmajor<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, st>>>(nconf, S->nsp, S->nk, in.as<double>(),
                                                                            inT.as<double>());
        CHM_LAUNCH_CHECK();
    }
    if ((rc = launch_transform(S, st, nconf, inT.as<double>(), sums.as<double>()))) return rc;
    k_ps_chi2<<<nconf, 64, 0, st>>>(sums.as<double>(), S->nsp, S->nbc, S->q0.as<double>(), S->profile, S->nfit,
                                    c2.as<double>());
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(out, c2.p, sizeof(double) * nconf, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}
