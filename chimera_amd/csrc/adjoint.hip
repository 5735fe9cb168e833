// Adjoints of the pair-space path sum and the relaxed-occupancy path sum (gfx950), and
// HIP-event timing of the forward-model kernels.
//
// dL/dr (hard and soft occupancy): a workgroup owns 1024 pairs of one configuration, one
// pair per lane (16 wavefronts share one copy of the staged tables).  The per-k tables
// (f, delta or f cos(delta) / f sin(delta), k, -2/lambda) and the configuration's dL/dchi
// are staged in LDS once; every lane walks
// the k grid with its pair's accumulators in registers, so all 64 lanes of a wavefront
// read the same k column of a table row (a handful of distinct LDS words per step).
// dL/dpos and dL/dw: one lane per (configuration, atom) gathers its pairs through CSR
// lists built with the topology (pairs where the atom is the absorber, then where it
// is the scatterer), always in the same order: the gradients are deterministic (no
// floating-point atomics).
//
// Relaxed-occupancy forward: lanes own the k > 0 points, one each (as the hard path
// sum), a pair record carries both sites' weight vectors; T_ab = base f_b sin(2kr +
// delta_ab) is formed from one sincos(2kr) per (pair, k) and the tables f_b cos(delta_ab),
// f_b sin(delta_ab).
//
// Phases are carried in revolutions (2 k r / 2 pi = (k / pi) r) and sin / cos / exp are
// the range-specific polynomials of fp64math.h.  The
// per-k constants sit in LDS as one KRow per k point, read by all lanes at once.
#include <cmath>

#include "common.h"
#include "fp64math.h"

namespace chm {
int distances_device(chm_ctx *ctx, int nconf, bool want_evec);
int chi_device(chm_ctx *ctx, int nconf, int precision);
int ptypes_device(chm_ctx *ctx, int nconf);
int reduce_partials(chm_ctx *ctx, const double *partial, double *out, int nwg, int n_per_conf,
                    int nconf);
}  // namespace chm

namespace {

constexpr int kLanes = 1024;   // adjoint workgroups: pairs per workgroup (tables staged once)
constexpr int kMaxK = 512;   // k points above 0 in the soft forward (one lane each)
constexpr int kSoftRegNsp = 3;   // soft forward: table columns in registers up to this
constexpr double kInvPi = 0.31830988618379067154;
constexpr double kC3Rev = (4.0 / 3.0) * 0.15915494309189533577;   // (4/3) / (2 pi)

// Per-k constants of a term: e^(c1 r + c2 sigma^2) is the damping, kp r the phase
// 2 k r in revolutions, c3k C3 the cumulant phase in revolutions, pref = S0^2 / k
// (0 for k <= 0, where the term is defined as 0), tk = 2 k = d(2 k r)/dr.
struct KRow {
    double kp, c1, c2, pref, c3k, tk;
};

__device__ inline void stage_krows(KRow *rows, const double *__restrict__ ktab,
                                   const double *__restrict__ lamtab, int nk, double s02) {
    for (int q = threadIdx.x; q < nk; q += blockDim.x) {
        const double kv = ktab[q];
        KRow kr;
        kr.kp = kv * kInvPi;
        kr.c1 = -2.0 / lamtab[q];
        kr.c2 = -2.0 * kv * kv;
        kr.pref = kv > 0.0 ? s02 / kv : 0.0;
        kr.c3k = kC3Rev * kv * kv * kv;
        kr.tk = 2.0 * kv;
        rows[q] = kr;
    }
}

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

// ---------------------------------------------------------------- hard occupancy

// gr[ic][p] = sum_k gchi[ic][a_p][k] dT_p/dr (r_eff = r + dr, sigma^2 + s2tab, C3 phase)
template <int NSP>
__global__ __launch_bounds__(1024) void k_adjoint_r(
    const double *__restrict__ r_all, const int *__restrict__ pt_all,
    const int *__restrict__ shell, const double *__restrict__ sig2,
    const double *__restrict__ ftab, const double *__restrict__ dtab,
    const double *__restrict__ lamtab, const double *__restrict__ ktab,
    const double *__restrict__ drtab, const double *__restrict__ c3tab,
    const double *__restrict__ s2tab, const double *__restrict__ gchi,
    double *__restrict__ gr, int npairs, int nk, int kw, double s02, const chm::PolyCoefs pc) {
    extern __shared__ double smem[];
    double *sf = smem;                  // f (NSP, kw): the columns of the current k chunk
    double *sd = sf + NSP * kw;         // delta / 2 pi (NSP*NSP, kw)
    double *sg = sd + NSP * NSP * kw;   // this configuration's dL/dchi (NSP, kw)
    KRow *sk = reinterpret_cast<KRow *>(sg + NSP * kw);
    const int ic = blockIdx.y;
    const int p = blockIdx.x * kLanes + threadIdx.x;
    const bool valid = p < npairs;
    const long o = (long)ic * npairs + (valid ? p : 0);
    const int t = pt_all[o];
    const int sh = shell[valid ? p : 0];
    const int sa = t / NSP, sb = t - sa * NSP;
    const double re = r_all[o] + drtab[sh * NSP * NSP + t];
    const double ir = 1.0 / re;
    const double ir2 = ir * ir;
    const double s2 = sig2[valid ? p : 0] + s2tab[sh * NSP * NSP + t];
    const double c3 = c3tab[sh];
    double acc = 0.0;
    // the k grid in chunks of kw columns (one chunk when the tables fit the LDS), each
    // staged by the workgroup and summed by every lane in k order
    for (int kb = 0; kb < nk; kb += kw) {
        const int kc = min(kw, nk - kb);
        __syncthreads();                 // the previous chunk's columns are no longer read
        for (int q = threadIdx.x; q < NSP * kc; q += kLanes) {
            const int row = q / kc, c = q - row * kc;
            sf[row * kw + c] = ftab[row * nk + kb + c];
            sg[row * kw + c] = gchi[((long)ic * NSP + row) * nk + kb + c];
        }
        for (int q = threadIdx.x; q < NSP * NSP * kc; q += kLanes) {
            const int row = q / kc, c = q - row * kc;
            sd[row * kw + c] = dtab[row * nk + kb + c] * (0.5 * kInvPi);
        }
        stage_krows(sk, ktab + kb, lamtab + kb, kc, s02);
        __syncthreads();
        if (!valid) continue;
        const double *gs = sg + sa * kw;
        const double *fs = sf + sb * kw;
        const double *ds = sd + t * kw;
        for (int j = 0; j < kc; ++j) {
            const double g = gs[j];
            const KRow kr = sk[j];
            if (g == 0.0 || kr.pref == 0.0) continue;   // the k = 0 term is defined as 0
            const double ph = fma(kr.kp, re, fma(-kr.c3k, c3, ds[j]));
            double sn, cs;
            chm::sincos2pi(ph, pc, &sn, &cs);
            const double amp = kr.pref * fs[j] * ir2 * chm::exp_fin(fma(kr.c1, re, kr.c2 * s2), pc);
            acc += g * amp * ((kr.c1 - 2.0 * ir) * sn + kr.tk * cs);
        }
    }
    if (valid) gr[o] = acc;
}

// The same dL/dr on a uniform k grid with C3 = 0: e^{2ikr} advances along k by one
// complex rotation e^{2i dk r} per point (4 FMAs) and is recomputed exactly every
// kAnchorEvery points (and at the first), so no per-term sine or cosine is evaluated.
// The tables are f_b cos(delta_t), f_b sin(delta_t) (one row per pair type):
// f sin(2kr + delta) = s fc + c fs and f cos(2kr + delta) = c fc - s fs.
constexpr int kAnchorEvery = 32;

template <int NSP>
__global__ __launch_bounds__(1024) void k_adjoint_r_rot(
    const double *__restrict__ r_all, const int *__restrict__ pt_all,
    const int *__restrict__ shell, const double *__restrict__ sig2,
    const double *__restrict__ fctab, const double *__restrict__ fstab,
    const double *__restrict__ lamtab, const double *__restrict__ ktab,
    const double *__restrict__ drtab, const double *__restrict__ s2tab,
    const double *__restrict__ gchi, double *__restrict__ gr, int npairs, int nk, double dkp,
    double s02, const chm::PolyCoefs pc) {
    extern __shared__ double smem[];
    double *sfc = smem;                  // f_b cos(delta_t) (NSP*NSP, nk)
    double *sfs = sfc + NSP * NSP * nk;  // f_b sin(delta_t)
    double *sg = sfs + NSP * NSP * nk;   // this configuration's dL/dchi (NSP, nk)
    KRow *sk = reinterpret_cast<KRow *>(sg + NSP * nk);
    const int ic = blockIdx.y;
    for (int q = threadIdx.x; q < NSP * NSP * nk; q += kLanes) {
        sfc[q] = fctab[q];
        sfs[q] = fstab[q];
    }
    for (int q = threadIdx.x; q < NSP * nk; q += kLanes) sg[q] = gchi[(long)ic * NSP * nk + q];
    stage_krows(sk, ktab, lamtab, nk, s02);
    __syncthreads();
    const int p = blockIdx.x * kLanes + threadIdx.x;
    if (p >= npairs) return;
    const long o = (long)ic * npairs + p;
    const int t = pt_all[o];
    const int sh = shell[p];
    const int sa = t / NSP;
    const double re = r_all[o] + drtab[sh * NSP * NSP + t];
    const double ir = 1.0 / re;
    const double ir2 = ir * ir;
    const double s2 = sig2[p] + s2tab[sh * NSP * NSP + t];
    const double *gs = sg + sa * nk;
    const double *fc = sfc + t * nk;
    const double *fs = sfs + t * nk;
    double rs, rc;                       // e^{2i dk r}
    chm::sincos2pi(dkp * re, pc, &rs, &rc);
    double s = 0.0, c = 1.0;             // e^{2ikr}
    double acc = 0.0;
    for (int ik = 0; ik < nk; ++ik) {
        const KRow kr = sk[ik];
        if (ik % kAnchorEvery == 0) {
            chm::sincos2pi(kr.kp * re, pc, &s, &c);
        } else {
            const double c2 = fma(c, rc, -s * rs);
            s = fma(s, rc, c * rs);
            c = c2;
        }
        const double g = gs[ik];
        if (g == 0.0 || kr.pref == 0.0) continue;   // the k = 0 term is defined as 0
        const double fsn = fma(s, fc[ik], c * fs[ik]);    // f sin(2kr + delta)
        const double fcs = fma(c, fc[ik], -s * fs[ik]);   // f cos(2kr + delta)
        const double amp = kr.pref * ir2 * chm::exp_fin(fma(kr.c1, re, kr.c2 * s2), pc);
        acc += g * amp * fma(kr.c1 - 2.0 * ir, fsn, kr.tk * fcs);
    }
    gr[o] = acc;
}

// gpos[ic][a] = sum_{p: i_p = a} -gr_p e_p + sum_{p: j_p = a} gr_p e_p; with gwi/gwj also
// gw[ic][a][s] = sum_{p: i_p = a} gwi_p[s] + sum_{p: j_p = a} gwj_p[s].
__global__ void k_gather_atoms(const double *__restrict__ gr, const double *__restrict__ evec,
                               const double *__restrict__ gwi, const double *__restrict__ gwj,
                               const int *__restrict__ ioff, const int *__restrict__ ilist,
                               const int *__restrict__ joff, const int *__restrict__ jlist,
                               double *__restrict__ gpos, double *__restrict__ gw, int npairs,
                               int natoms, int nconf, int nsp) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)natoms * nconf) return;
    const int ic = (int)(tid / natoms), a = (int)(tid % natoms);
    const double *g = gr + (long)ic * npairs;
    const double *e = evec + (long)ic * npairs * 3;
    double x = 0.0, y = 0.0, z = 0.0;
    double ws[chm::kMaxNsp];
    for (int s = 0; s < chm::kMaxNsp; ++s) ws[s] = 0.0;
    for (int q = ioff[a]; q < ioff[a + 1]; ++q) {
        const int p = ilist[q];
        const double v = g[p];
        x -= v * e[3 * p];
        y -= v * e[3 * p + 1];
        z -= v * e[3 * p + 2];
        if (gw)
            for (int s = 0; s < nsp; ++s) ws[s] += gwi[((long)ic * npairs + p) * nsp + s];
    }
    for (int q = joff[a]; q < joff[a + 1]; ++q) {
        const int p = jlist[q];
        const double v = g[p];
        x += v * e[3 * p];
        y += v * e[3 * p + 1];
        z += v * e[3 * p + 2];
        if (gw)
            for (int s = 0; s < nsp; ++s) ws[s] += gwj[((long)ic * npairs + p) * nsp + s];
    }
    gpos[3 * tid] = x;
    gpos[3 * tid + 1] = y;
    gpos[3 * tid + 2] = z;
    if (gw)
        for (int s = 0; s < nsp; ++s) gw[tid * nsp + s] = ws[s];
}

// ---------------------------------------------------------------- soft occupancy

// fc[t][k] = f_b(k) cos(delta_t(k)), fs[t][k] = f_b(k) sin(delta_t(k)), t = a*nsp + b
__global__ void k_soft_tables(const double *__restrict__ ftab, const double *__restrict__ dtab,
                              double *__restrict__ fc, double *__restrict__ fs, int nsp, int nk) {
    int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nsp * nsp * nk) return;
    const int t = tid / nk, ik = tid - t * nk;
    const double f = ftab[(t % nsp) * nk + ik];
    double sn, cs;
    sincos(dtab[tid], &sn, &cs);
    fc[tid] = f * cs;
    fs[tid] = f * sn;
}

template <int NSP>
struct SoftRec {
    double r, lir2, s2;   // r, ln(1 / r^2) (folded into the exponent), sigma^2
    double wi[NSP], wj[NSP];
};

// One lane per k point above 0 (lane l owns k index k0 + l); each round stages blockDim
// pair records with both sites' weights.
template <int NSP>
__global__ __launch_bounds__(512) void k_chi_soft(
    const double *__restrict__ r_all, const double *__restrict__ w_all,
    const int *__restrict__ iidx, const int *__restrict__ jidx,
    const double *__restrict__ sig2, const double *__restrict__ fctab,
    const double *__restrict__ fstab, const double *__restrict__ lamtab,
    const double *__restrict__ ktab, double *__restrict__ partial, int npairs, int natoms,
    int nk, int k0, int pairs_per_wg, int kspan, double s02, const chm::PolyCoefs pc) {
    extern __shared__ double smem[];
    // the lane's own column of f cos(delta), f sin(delta): in registers up to 3 species
    // (18 values), else the workgroup's columns in LDS
    constexpr bool kReg = NSP <= kSoftRegNsp;
    double *sfc = smem;
    double *sfs = sfc + (kReg ? 0 : NSP * NSP * kspan);
    SoftRec<NSP> *tile = reinterpret_cast<SoftRec<NSP> *>(sfs + (kReg ? 0 : NSP * NSP * kspan));
    const int ic = blockIdx.y, wg = blockIdx.x, nwg = gridDim.x;
    const int nthr = blockDim.x;
    // this workgroup's k columns kb .. kb + kc - 1: chunk blockIdx.z of kspan columns
    const int kb = k0 + (int)blockIdx.z * kspan;
    const int kc = min(kspan, nk - kb);
    const int ik = kb + (int)threadIdx.x;
    const bool live = (int)threadIdx.x < kc;
    const int ikl = live ? (int)threadIdx.x : kc - 1;   // column within the chunk
    const int ikc = kb + ikl;                             // column of the k grid
    double fcr[kReg ? NSP * NSP : 1], fsr[kReg ? NSP * NSP : 1];
    if constexpr (kReg) {
#pragma unroll
        for (int q = 0; q < NSP * NSP; ++q) {
            fcr[q] = fctab[q * nk + ikc];
            fsr[q] = fstab[q * nk + ikc];
        }
    } else {
        for (int q = threadIdx.x; q < NSP * NSP * kc; q += nthr) {
            const int row = q / kc, c = q - row * kc;
            sfc[row * kspan + c] = fctab[row * nk + kb + c];
            sfs[row * kspan + c] = fstab[row * nk + kb + c];
        }
    }
    const double kv = live ? ktab[ik] : 0.0;
    const double kp = kv * kInvPi;
    const double c1 = -2.0 / (live ? lamtab[ik] : 1.0);
    const double c2 = -2.0 * kv * kv;
    const double pref = (live && kv > 0.0) ? s02 / kv : 0.0;
    double acc[NSP];
#pragma unroll
    for (int s = 0; s < NSP; ++s) acc[s] = 0.0;
    const double *rc = r_all + (long)ic * npairs;
    const double *wc = w_all + (long)ic * natoms * NSP;
    const int pbeg = wg * pairs_per_wg, pend = min(npairs, pbeg + pairs_per_wg);
    for (int p0 = pbeg; p0 < pend; p0 += nthr) {
        __syncthreads();
        {
            const int p = p0 + (int)threadIdx.x;
            SoftRec<NSP> rec;
            if (p < pend) {
                const double re = rc[p];
                rec.r = re;
                rec.lir2 = -2.0 * log(re);
                rec.s2 = sig2[p];
                const double *wi = wc + (long)iidx[p] * NSP;
                const double *wj = wc + (long)jidx[p] * NSP;
#pragma unroll
                for (int s = 0; s < NSP; ++s) {
                    rec.wi[s] = wi[s];
                    rec.wj[s] = wj[s];
                }
            } else {
                rec.r = 1.0;
                rec.lir2 = -2000.0;   // e^-2000 = 0: contributes exactly 0
                rec.s2 = 0.0;
#pragma unroll
                for (int s = 0; s < NSP; ++s) rec.wi[s] = rec.wj[s] = 0.0;
            }
            tile[threadIdx.x] = rec;
        }
        __syncthreads();
        const int nt = min(nthr, pend - p0);
        for (int q = 0; q < nt; ++q) {
            const SoftRec<NSP> rec = tile[q];
            const double base = chm::exp_fin(fma(c1, rec.r, fma(c2, rec.s2, rec.lir2)), pc);
            double sn, cs;
            chm::sincos2pi(kp * rec.r, pc, &sn, &cs);
            // sum_b w_j[b] f sin(2kr + delta_ab) = sn A_a + cs B_a with
            // A_a = sum_b w_j[b] fc_ab, B_a = sum_b w_j[b] fs_ab
#pragma unroll
            for (int a = 0; a < NSP; ++a) {
                double A = 0.0, B = 0.0;
#pragma unroll
                for (int b = 0; b < NSP; ++b) {
                    const int t = (a * NSP + b) * kspan + ikl;
                    A = fma(rec.wj[b], kReg ? fcr[a * NSP + b] : sfc[t], A);
                    B = fma(rec.wj[b], kReg ? fsr[a * NSP + b] : sfs[t], B);
                }
                acc[a] = fma(rec.wi[a] * base, fma(sn, A, cs * B), acc[a]);
            }
        }
    }
    double *out = partial + ((long)ic * nwg + wg) * NSP * nk;
    if (blockIdx.z == 0)
        for (int q = threadIdx.x; q < k0; q += nthr)
#pragma unroll
            for (int s = 0; s < NSP; ++s) out[s * nk + q] = 0.0;
    if (live) {
#pragma unroll
        for (int s = 0; s < NSP; ++s) out[s * nk + ik] = pref * acc[s];
    }
}

// Per pair: gr (dL/dr), gwi[s] (dL/dw_i[s]), gwj[s] (dL/dw_j[s]) of the soft path sum.
template <int NSP>
__global__ __launch_bounds__(1024) void k_adjoint_soft(
    const double *__restrict__ r_all, const double *__restrict__ w_all,
    const int *__restrict__ iidx, const int *__restrict__ jidx,
    const double *__restrict__ sig2, const double *__restrict__ fctab,
    const double *__restrict__ fstab, const double *__restrict__ lamtab,
    const double *__restrict__ ktab, const double *__restrict__ gchi,
    double *__restrict__ gr, double *__restrict__ gwi_out, double *__restrict__ gwj_out,
    int npairs, int natoms, int nk, int kw, double dkp, double s02, const chm::PolyCoefs pc) {
    extern __shared__ double smem[];
    double *sfc = smem;                   // the columns of the current k chunk (kw wide)
    double *sfs = sfc + NSP * NSP * kw;
    double *sg = sfs + NSP * NSP * kw;
    KRow *sk = reinterpret_cast<KRow *>(sg + NSP * kw);
    const int ic = blockIdx.y;
    const int p = blockIdx.x * kLanes + threadIdx.x;
    const bool valid = p < npairs;
    const int pp = valid ? p : 0;
    const long o = (long)ic * npairs + pp;
    const double re = r_all[o];
    const double ir = 1.0 / re;
    const double ir2 = ir * ir;
    const double s2 = sig2[pp];
    double wi[NSP], wj[NSP], gwi[NSP], gwj[NSP];
    const double *wc = w_all + (long)ic * natoms * NSP;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        wi[s] = wc[(long)iidx[pp] * NSP + s];
        wj[s] = wc[(long)jidx[pp] * NSP + s];
        gwi[s] = gwj[s] = 0.0;
    }
    double acc = 0.0;
    // uniform k grid (dkp > 0): e^{2ikr} by the rotation recurrence of k_adjoint_r_rot
    double rs = 0.0, rc = 1.0, sn = 0.0, cs = 1.0;
    if (dkp > 0.0) chm::sincos2pi(dkp * re, pc, &rs, &rc);
    // the k grid in chunks of kw columns (one chunk when the tables fit the LDS), each
    // staged by the workgroup and summed by every lane in k order
    for (int kb = 0; kb < nk; kb += kw) {
      const int kc = min(kw, nk - kb);
      __syncthreads();                   // the previous chunk's columns are no longer read
      for (int q = threadIdx.x; q < NSP * NSP * kc; q += kLanes) {
          const int row = q / kc, c = q - row * kc;
          sfc[row * kw + c] = fctab[row * nk + kb + c];
          sfs[row * kw + c] = fstab[row * nk + kb + c];
      }
      for (int q = threadIdx.x; q < NSP * kc; q += kLanes) {
          const int row = q / kc, c = q - row * kc;
          sg[row * kw + c] = gchi[((long)ic * NSP + row) * nk + kb + c];
      }
      stage_krows(sk, ktab + kb, lamtab + kb, kc, s02);
      __syncthreads();
      if (!valid) continue;
      for (int ik = kb; ik < kb + kc; ++ik) {
        const int j = ik - kb;
        const KRow kr = sk[j];
        if (dkp == 0.0 || ik % kAnchorEvery == 0) {
            chm::sincos2pi(kr.kp * re, pc, &sn, &cs);
        } else {
            const double c2 = fma(cs, rc, -sn * rs);
            sn = fma(sn, rc, cs * rs);
            cs = c2;
        }
        if (kr.pref == 0.0) continue;
        const double base = kr.pref * ir2 * chm::exp_fin(fma(kr.c1, re, kr.c2 * s2), pc);
        const double dfac = kr.c1 - 2.0 * ir;
        // with A_a = sum_b w_j[b] fc_ab, B_a = sum_b w_j[b] fs_ab the pair's weighted
        // f sin / f cos over b are sn A + cs B and cs A - sn B; with u_a = g_a w_i[a],
        // C_b = sum_a u_a fc_ab, D_b = sum_a u_a fs_ab the w_j gradient is
        // base (sn C_b + cs D_b)
        double C[NSP], D[NSP];
#pragma unroll
        for (int b = 0; b < NSP; ++b) C[b] = D[b] = 0.0;
        double accv = 0.0;
#pragma unroll
        for (int a = 0; a < NSP; ++a) {
            const double ga = sg[a * kw + j];
            const double u = ga * wi[a];
            double A = 0.0, B = 0.0;
#pragma unroll
            for (int b = 0; b < NSP; ++b) {
                const int t = (a * NSP + b) * kw + j;
                const double fc = sfc[t], fs = sfs[t];
                A = fma(wj[b], fc, A);
                B = fma(wj[b], fs, B);
                C[b] = fma(u, fc, C[b]);
                D[b] = fma(u, fs, D[b]);
            }
            const double st = fma(sn, A, cs * B);    // sum_b w_j[b] f sin(2kr + delta_ab)
            const double dt = fma(cs, A, -sn * B);   // sum_b w_j[b] f cos(2kr + delta_ab)
            gwi[a] = fma(ga * base, st, gwi[a]);
            accv = fma(u, fma(dfac, st, kr.tk * dt), accv);
        }
        acc = fma(base, accv, acc);
#pragma unroll
        for (int b = 0; b < NSP; ++b) gwj[b] = fma(base, fma(sn, C[b], cs * D[b]), gwj[b]);
      }
    }
    if (!valid) return;
    gr[o] = acc;
#pragma unroll
    for (int s = 0; s < NSP; ++s) {
        gwi_out[o * NSP + s] = gwi[s];
        gwj_out[o * NSP + s] = gwj[s];
    }
}

// ---------------------------------------------------------------- launchers

size_t adjoint_lds(int nsp, int nk, bool soft) {
    return soft ? sizeof(double) * (2 * nsp * nsp + nsp) * nk + sizeof(KRow) * nk
                : sizeof(double) * (nsp + nsp * nsp + nsp) * nk + sizeof(KRow) * nk;
}

// k columns per chunk of the adjoint kernels: all of them when the tables fit 160 KB of
// LDS, else the most that do (the kernels then restage chunk by chunk)
int adjoint_kw(int nsp, int nk, bool soft) {
    const size_t per = adjoint_lds(nsp, 1, soft);
    return (int)std::max<size_t>(1, std::min<size_t>(nk, (160 * 1024) / per));
}

int soft_tables(chm_ctx *ctx);

// k grid uniform to rounding (|k_i - k_0 - i dk| <= 8 eps max|k|): the rotation
// recurrence then tracks 2kr to ~1e-13 rad; returns dk / pi, or 0 when it is not.
double uniform_dkp(const chm_ctx *ctx) {
    const int nk = ctx->nk;
    if (nk < 2) return 0.0;
    const std::vector<double> &k = ctx->h_k;
    const double dk = (k[nk - 1] - k[0]) / (nk - 1);
    double kmax = 0.0;
    for (double v : k) kmax = std::max(kmax, std::fabs(v));
    const double tol = 8.0 * 2.220446049250313e-16 * kmax;
    if (!(dk > 0.0)) return 0.0;
    for (int i = 0; i < nk; ++i)
        if (std::fabs(k[i] - (k[0] + i * dk)) > tol) return 0.0;
    return dk * kInvPi;
}

template <int NSP>
int launch_adjoint_r(chm_ctx *ctx, int nconf) {
    const int nk = ctx->nk, npairs = ctx->npairs;
    const dim3 grid((npairs + kLanes - 1) / kLanes, nconf);
    const double dkp = uniform_dkp(ctx);
    const size_t lds_rot = sizeof(double) * (2 * NSP * NSP + NSP) * nk + sizeof(KRow) * nk;
    if (ctx->c3_zero && dkp > 0.0 && lds_rot <= 80 * 1024) {
        int rc = soft_tables(ctx);       // f cos(delta), f sin(delta)
        if (rc) return rc;
        hipFuncSetAttribute((const void *)k_adjoint_r_rot<NSP>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_rot);
        k_adjoint_r_rot<NSP><<<grid, kLanes, lds_rot, ctx->stream>>>(
            ctx->r.as<double>(), ctx->ptype.as<int>(), ctx->shell.as<int>(),
            ctx->sig2.as<double>(), ctx->soft_fc.as<double>(), ctx->soft_fs.as<double>(),
            ctx->lam.as<double>(), ctx->k.as<double>(), ctx->drtab.as<double>(),
            ctx->s2tab.as<double>(), ctx->gchi.as<double>(), ctx->gr.as<double>(), npairs, nk,
            dkp, ctx->s02, chm::kPolyCoefs);
        CHM_LAUNCH_CHECK();
        return CHM_OK;
    }
    const int kw = adjoint_kw(NSP, nk, false);
    const size_t lds = adjoint_lds(NSP, kw, false);
    hipFuncSetAttribute((const void *)k_adjoint_r<NSP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds);
    k_adjoint_r<NSP><<<grid, kLanes, lds, ctx->stream>>>(
        ctx->r.as<double>(), ctx->ptype.as<int>(), ctx->shell.as<int>(), ctx->sig2.as<double>(),
        ctx->f.as<double>(), ctx->delta.as<double>(), ctx->lam.as<double>(), ctx->k.as<double>(),
        ctx->drtab.as<double>(), ctx->c3tab.as<double>(), ctx->s2tab.as<double>(),
        ctx->gchi.as<double>(), ctx->gr.as<double>(), npairs, nk, kw, ctx->s02, chm::kPolyCoefs);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

template <int NSP>
int launch_chi_soft(chm_ctx *ctx, int nconf) {
    const int nk = ctx->nk, npairs = ctx->npairs;
    int k0 = 0;
    while (k0 < nk && !(ctx->h_k[k0] > 0.0)) ++k0;
    CHM_REQUIRE(k0 < nk, CHM_E_ARG, "no k point above 0");
    const int nkz = nk - k0;
    // k chunks (grid z): at most kMaxK lanes each, and (beyond three species) the chunk's
    // f cos(delta) / f sin(delta) columns plus one pair tile within 160 KB of LDS
    auto lds_of = [&](int span) {
        const int thr = (span + 63) / 64 * 64;
        return sizeof(double) * (NSP <= kSoftRegNsp ? 0 : 2 * NSP * NSP * (size_t)span) +
               sizeof(SoftRec<NSP>) * thr;
    };
    int nkc = (nkz + kMaxK - 1) / kMaxK;
    while (lds_of((nkz + nkc - 1) / nkc) > 160 * 1024) ++nkc;
    const int kspan = (nkz + nkc - 1) / nkc;
    const int nthr = (kspan + 63) / 64 * 64;
    const size_t lds = lds_of(kspan);
    const int rounds = (npairs + nthr - 1) / nthr;
    int nwg = std::max(1, std::min(rounds, (2048 + nconf - 1) / nconf));
    const int ppw = ((rounds + nwg - 1) / nwg) * nthr;
    nwg = (npairs + ppw - 1) / ppw;
    int rc = ctx->partial.ensure(sizeof(double) * (size_t)nconf * nwg * NSP * nk);
    if (rc) return rc;
    hipFuncSetAttribute((const void *)k_chi_soft<NSP>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds);
    k_chi_soft<NSP><<<dim3(nwg, nconf, nkc), nthr, lds, ctx->stream>>>(
        ctx->r.as<double>(), ctx->w.as<double>(), ctx->i_idx.as<int>(), ctx->j_idx.as<int>(),
        ctx->sig2.as<double>(), ctx->soft_fc.as<double>(), ctx->soft_fs.as<double>(),
        ctx->lam.as<double>(), ctx->k.as<double>(), ctx->partial.as<double>(), npairs, ctx->natoms,
        nk, k0, ppw, kspan, ctx->s02, chm::kPolyCoefs);
    CHM_LAUNCH_CHECK();
    if ((rc = ctx->chi.ensure(sizeof(double) * (size_t)nconf * NSP * nk))) return rc;
    return chm::reduce_partials(ctx, ctx->partial.as<double>(), ctx->chi.as<double>(), nwg,
                                NSP * nk, nconf);
}

template <int NSP>
int launch_adjoint_soft(chm_ctx *ctx, int nconf) {
    const int nk = ctx->nk, npairs = ctx->npairs;
    const int kw = adjoint_kw(NSP, nk, true);
    const size_t lds = adjoint_lds(NSP, kw, true);
    hipFuncSetAttribute((const void *)k_adjoint_soft<NSP>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    dim3 grid((npairs + kLanes - 1) / kLanes, nconf);
    k_adjoint_soft<NSP><<<grid, kLanes, lds, ctx->stream>>>(
        ctx->r.as<double>(), ctx->w.as<double>(), ctx->i_idx.as<int>(), ctx->j_idx.as<int>(),
        ctx->sig2.as<double>(), ctx->soft_fc.as<double>(), ctx->soft_fs.as<double>(),
        ctx->lam.as<double>(), ctx->k.as<double>(), ctx->gchi.as<double>(), ctx->gr.as<double>(),
        ctx->gwi.as<double>(), ctx->gwj.as<double>(), npairs, ctx->natoms, nk, kw, uniform_dkp(ctx),
        ctx->s02, chm::kPolyCoefs);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

#define CHM_NSP_DISPATCH(fn, ...)                                                     \
    switch (ctx->nsp) {                                                               \
        case 2: return fn<2>(__VA_ARGS__);                                            \
        case 3: return fn<3>(__VA_ARGS__);                                            \
        case 4: return fn<4>(__VA_ARGS__);                                            \
        case 5: return fn<5>(__VA_ARGS__);                                            \
        case 6: return fn<6>(__VA_ARGS__);                                            \
    }                                                                                 \
    chm::set_error("unsupported species count");                                      \
    return CHM_E_ARG;

int adjoint_r(chm_ctx *ctx, int nconf) { CHM_NSP_DISPATCH(launch_adjoint_r, ctx, nconf) }
int chi_soft(chm_ctx *ctx, int nconf) { CHM_NSP_DISPATCH(launch_chi_soft, ctx, nconf) }
int adjoint_soft(chm_ctx *ctx, int nconf) { CHM_NSP_DISPATCH(launch_adjoint_soft, ctx, nconf) }

int soft_tables(chm_ctx *ctx) {
    const int n = ctx->nsp * ctx->nsp * ctx->nk;
    int rc;
    if ((rc = ctx->soft_fc.ensure(sizeof(double) * n)) || (rc = ctx->soft_fs.ensure(sizeof(double) * n)))
        return rc;
    k_soft_tables<<<dim3((n + 255) / 256), 256, 0, ctx->stream>>>(
        ctx->f.as<double>(), ctx->delta.as<double>(), ctx->soft_fc.as<double>(),
        ctx->soft_fs.as<double>(), ctx->nsp, ctx->nk);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int gather_atoms(chm_ctx *ctx, int nconf, bool with_w) {
    const long tot = (long)nconf * ctx->natoms;
    int rc = ctx->gpos.ensure(sizeof(double) * 3 * tot);
    if (rc) return rc;
    if (with_w && (rc = ctx->gw.ensure(sizeof(double) * ctx->nsp * tot))) return rc;
    k_gather_atoms<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, ctx->stream>>>(
        ctx->gr.as<double>(), ctx->evec.as<double>(), with_w ? ctx->gwi.as<double>() : nullptr,
        with_w ? ctx->gwj.as<double>() : nullptr, ctx->csr_ioff.as<int>(), ctx->csr_i.as<int>(),
        ctx->csr_joff.as<int>(), ctx->csr_j.as<int>(), ctx->gpos.as<double>(),
        with_w ? ctx->gw.as<double>() : nullptr, ctx->npairs, ctx->natoms, nconf, ctx->nsp);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int check_species(const chm_ctx *ctx, int nconf, const int32_t *species) {
    const long n = (long)nconf * ctx->natoms;
    for (long q = 0; q < n; ++q)
        CHM_REQUIRE(species[q] >= 0 && species[q] < ctx->nsp, CHM_E_ARG,
                    "species index out of range");
    return CHM_OK;
}

// Everything the hard adjoint needs on the device: pos -> r, evec; species -> types.
int prepare_hard(chm_ctx *ctx, int nconf, const double *pos, const int32_t *species,
                 const double *gchi) {
    int rc = check_species(ctx, nconf, species);
    if (rc) return rc;
    const size_t na = (size_t)nconf * ctx->natoms;
    if ((rc = upload(ctx, ctx->pos, pos, sizeof(double) * 3 * na))) return rc;
    if ((rc = upload(ctx, ctx->species, species, sizeof(int32_t) * na))) return rc;
    if ((rc = upload(ctx, ctx->gchi, gchi, sizeof(double) * (size_t)nconf * ctx->nsp * ctx->nk)))
        return rc;
    if ((rc = chm::distances_device(ctx, nconf, true))) return rc;
    const long tot = (long)nconf * ctx->npairs;
    if ((rc = ctx->ptype.ensure(sizeof(int32_t) * tot))) return rc;
    if ((rc = ctx->gr.ensure(sizeof(double) * tot))) return rc;
    return CHM_OK;
}

int prepare_soft(chm_ctx *ctx, int nconf, const double *pos, const double *w, const double *gchi) {
    const size_t na = (size_t)nconf * ctx->natoms;
    int rc;
    if ((rc = upload(ctx, ctx->pos, pos, sizeof(double) * 3 * na))) return rc;
    if ((rc = upload(ctx, ctx->w, w, sizeof(double) * ctx->nsp * na))) return rc;
    if (gchi && (rc = upload(ctx, ctx->gchi, gchi,
                             sizeof(double) * (size_t)nconf * ctx->nsp * ctx->nk)))
        return rc;
    if ((rc = chm::distances_device(ctx, nconf, gchi != nullptr))) return rc;
    if ((rc = soft_tables(ctx))) return rc;
    if (gchi) {
        const long tot = (long)nconf * ctx->npairs;
        if ((rc = ctx->gr.ensure(sizeof(double) * tot)) ||
            (rc = ctx->gwi.ensure(sizeof(double) * ctx->nsp * tot)) ||
            (rc = ctx->gwj.ensure(sizeof(double) * ctx->nsp * tot)))
            return rc;
    }
    return CHM_OK;
}

}  // namespace

extern "C" int chm_fm_grad_pos(chm_ctx *ctx, int nconf, const double *pos,
                               const int32_t *species, const double *gchi, double *gpos) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(nconf > 0 && pos && species && gchi && gpos, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(ctx->device));
    int rc = prepare_hard(ctx, nconf, pos, species, gchi);
    if (rc) return rc;
    if ((rc = chm::ptypes_device(ctx, nconf))) return rc;
    if ((rc = adjoint_r(ctx, nconf))) return rc;
    if ((rc = gather_atoms(ctx, nconf, false))) return rc;
    ctx->prof_nconf[2] = nconf;
    return download(ctx, gpos, ctx->gpos, sizeof(double) * 3 * (size_t)nconf * ctx->natoms);
}

extern "C" int chm_fm_chi_soft(chm_ctx *ctx, int nconf, const double *pos, const double *w,
                               double *chi) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(nconf > 0 && pos && w && chi, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(ctx->device));
    int rc = prepare_soft(ctx, nconf, pos, w, nullptr);
    if (rc) return rc;
    if ((rc = chi_soft(ctx, nconf))) return rc;
    ctx->prof_nconf[3] = nconf;
    return download(ctx, chi, ctx->chi, sizeof(double) * (size_t)nconf * ctx->nsp * ctx->nk);
}

extern "C" int chm_fm_grad_soft(chm_ctx *ctx, int nconf, const double *pos, const double *w,
                                const double *gchi, double *gpos, double *gw) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(nconf > 0 && pos && w && gchi && gpos && gw, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(ctx->device));
    int rc = prepare_soft(ctx, nconf, pos, w, gchi);
    if (rc) return rc;
    if ((rc = adjoint_soft(ctx, nconf))) return rc;
    if ((rc = gather_atoms(ctx, nconf, true))) return rc;
    ctx->prof_nconf[4] = nconf;
    const size_t na = (size_t)nconf * ctx->natoms;
    CHM_HIP(hipMemcpyAsync(gw, ctx->gw.p, sizeof(double) * ctx->nsp * na, hipMemcpyDeviceToHost,
                           ctx->stream));
    return download(ctx, gpos, ctx->gpos, sizeof(double) * 3 * na);
}

extern "C" int chm_fm_profile(chm_ctx *ctx, int what, int nconf, int iters, double *ms) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(what >= 0 && what <= 4 && nconf > 0 && iters > 0 && ms, CHM_E_ARG,
                "bad arguments");
    CHM_REQUIRE(ctx->prof_nconf[what] == nconf, CHM_E_STATE,
                "run the profiled call once with this nconf before profiling it");
    CHM_HIP(hipSetDevice(ctx->device));
    hipEvent_t e0, e1;
    CHM_HIP(hipEventCreate(&e0));
    CHM_HIP(hipEventCreate(&e1));
    int rc = CHM_OK;
    auto once = [&]() -> int {
        switch (what) {
            case 0: return chm::chi_device(ctx, nconf, CHM_PREC_F64);
            case 1: return chm::chi_device(ctx, nconf, CHM_PREC_F32);
            case 2: return adjoint_r(ctx, nconf);
            case 3: return chi_soft(ctx, nconf);
            default: return adjoint_soft(ctx, nconf);
        }
    };
    if ((rc = once())) return rc;   // warm
    hipEventRecord(e0, ctx->stream);
    for (int it = 0; it < iters && !rc; ++it) rc = once();
    hipEventRecord(e1, ctx->stream);
    hipEventSynchronize(e1);
    float t = 0.f;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    *ms = (double)t / iters;
    return CHM_OK;
}
