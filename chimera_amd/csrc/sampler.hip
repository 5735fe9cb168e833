// Device-resident parallel-tempering RMC swap sampler in count space (gfx950).
//
// State per replica slot: species (u8 per site), integer counts of every basis row
// (ordered pairs N[a][b][shell] and multiple-scattering entries C[e]), chi^2.  On the
// ideal lattice a pair's (or path's) chi(k) term depends only on the species at its
// sites, so chi(R) of a configuration is counts @ basis_R, where basis_R holds the
// windowed transform of every row (built once).  One step, for all replicas:
//
//   K1 propose + gather:  splitmix64 proposal (methods/method_a_rmc.py:253-259,
//      exafs_gpu/kernels.py:362-398); the directed pairs and the multiple-scattering
//      paths through the two swapped sites (kernels.py:425-468, chimera/mslibrary.py:
//      466-486) become integer count changes, histogrammed in LDS and added to a
//      per-replica delta with integer atomics (order-independent, so deterministic).
//   K2 trial transform:   chi(R)_trial = (counts + delta) @ basis_R, one FP64 GEMM
//      over replicas x rows for every edge (the reference's resample + window + FFT of
//      the trial chi, chimera/multift.py:58-63, done as a contraction).
//   K3 accept:            per-edge profiled scale and reduced chi^2
//      (method_a_rmc.py:190-227), Metropolis with a counter-based uniform
//      (method_a_rmc.py:280-285), commit of counts and species, best retention
//      (method_a_rmc.py:301-307), move log.
//   K4 exchange:          every exchange_every steps, alternating neighbour rungs swap
//      temperatures (method_a_rmc.py:332-373): the rung <-> slot permutation changes,
//      no configuration moves.
//
// Computing chi(R)_trial from the exact integer counts every step means the running
// state never drifts: there is no floating-point accumulator to re-anchor.
#include <cmath>
#include <cstring>

#include "common.h"

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kAccSalt = 0xD1B54A32D192ED03ull;
constexpr uint64_t kExcSalt = 0x94D049BB133111EBull;

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += kGolden;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double u01(uint64_t h) {
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

struct Geo {
    int natoms, natoms_pad, nsp, ncoord, nshell, nblk, nstencil, nabsrec;
    int nrow_e, nrow_tot, nss_rows, ncol_e, ncol_tot, nkw, nwin;
    int box[3], ncell[3], nbasis;
    int bpar[8];
    int has_const;
};

__device__ __forceinline__ int site_index(const Geo &g, int x, int y, int z) {
    int b = g.bpar[(x & 1) | ((y & 1) << 1) | ((z & 1) << 2)];
    int cell = ((x >> 1) * g.ncell[1] + (y >> 1)) * g.ncell[2] + (z >> 1);
    return cell * g.nbasis + b;
}

__device__ __forceinline__ int wrap(int v, int n) {
    v = (v < 0) ? v + n : v;
    return (v >= n) ? v - n : v;
}

struct Prop {
    int a, b, na, nb, sw;
};

// kernels.py:378-398: h = splitmix64(seed ^ step*golden ^ key<<32), then swap flag,
// a, and up to 8 candidates.  swap_frac = 1: the flag is always set.  All 8 candidates
// repeating a's species makes a null move.
__device__ Prop propose(const uint8_t *sp, int natoms, uint64_t seed, uint64_t stepc,
                        uint64_t key) {
    uint64_t h = splitmix64(seed ^ (stepc * kGolden) ^ (key << 32));
    h = splitmix64(h);  // swap flag draw, u < 1 always
    h = splitmix64(h);
    int a = (int)(u01(h) * natoms);
    if (a >= natoms) a = natoms - 1;
    int sa = sp[a];
    Prop p{a, a, sa, sa, 0};
    for (int t = 0; t < 8; ++t) {
        h = splitmix64(h);
        int c = (int)(u01(h) * natoms);
        if (c >= natoms) c = natoms - 1;
        if (sp[c] != sa) {
            p.b = c;
            p.na = sp[c];
            p.nb = sa;
            p.sw = 1;
            break;
        }
    }
    return p;
}

__device__ __forceinline__ void load_species(uint8_t *lsp, const uint8_t *gsp, int natoms_pad) {
    const uint4 *src = reinterpret_cast<const uint4 *>(gsp);
    uint4 *dst = reinterpret_cast<uint4 *>(lsp);
    for (int q = threadIdx.x; q < natoms_pad / 16; q += blockDim.x) dst[q] = src[q];
}

// ---------------------------------------------------------------- K1
// grid (nsplit, nrep); block 256.  LDS: species (natoms_pad B) + hist (nrow_tot int).
__global__ __launch_bounds__(256) void k_propose_gather(
    Geo g, const uint8_t *__restrict__ species, const int *__restrict__ rung_of,
    const uint64_t *__restrict__ stepc_arr, uint64_t seed, const int *__restrict__ site_xyz,
    const int *__restrict__ nbr_j, const int *__restrict__ nbr_sh,
    const int4 *__restrict__ stencil, int *__restrict__ dcounts, int *__restrict__ prop_out) {
    extern __shared__ int lds[];
    int *hist = lds;
    uint8_t *lsp = reinterpret_cast<uint8_t *>(lds + ((g.nrow_tot + 3) & ~3));
    __shared__ Prop P;
    const int s = blockIdx.y;
    load_species(lsp, species + (long)s * g.natoms_pad, g.natoms_pad);
    for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x) hist[q] = 0;
    __syncthreads();
    if (threadIdx.x == 0) {
        P = propose(lsp, g.natoms, seed, stepc_arr[s] + 1, (uint64_t)rung_of[s]);
        if (blockIdx.x == 0) {
            int *o = prop_out + 5 * s;
            o[0] = P.a;
            o[1] = P.b;
            o[2] = P.na;
            o[3] = P.nb;
            o[4] = P.sw;
        }
    }
    __syncthreads();
    const Prop p = P;
    if (!p.sw) return;
    const int nss = 2 * g.ncoord;
    const int nitems = nss + 2 * g.nstencil;
    const int stride = gridDim.x * blockDim.x;
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < nitems; it += stride) {
        if (it < nss) {
            // ordered pairs x->y and y->x for every neighbour y of x in {a, b}
            const int x = (it < g.ncoord) ? p.a : p.b;
            const int o = (it < g.ncoord) ? it : it - g.ncoord;
            const int y = nbr_j[x * g.ncoord + o];
            if (x == p.b && y == p.a) continue;  // covered from a's side
            const int sh = nbr_sh[x * g.ncoord + o];
            const int sx = lsp[x], sy = lsp[y];
            const int nx = (x == p.a) ? p.na : p.nb;
            const int ny = (y == p.a) ? p.na : ((y == p.b) ? p.nb : sy);
            atomicAdd(&hist[sx * g.nrow_e + sy * g.nshell + sh], -1);
            atomicAdd(&hist[nx * g.nrow_e + ny * g.nshell + sh], 1);
            atomicAdd(&hist[sy * g.nrow_e + sx * g.nshell + sh], -1);
            atomicAdd(&hist[ny * g.nrow_e + nx * g.nshell + sh], 1);
        } else {
            const int m = it - nss;
            const int x = (m < g.nstencil) ? p.a : p.b;
            const int4 rec = stencil[(m < g.nstencil) ? m : m - g.nstencil];
            const int8_t *rb = reinterpret_cast<const int8_t *>(&rec);
            const int nslot = rb[12], wgt = rb[13];
            const int cls_off = (int)(uint8_t)rb[14] | ((int)rb[15] << 8);
            const int xyz = site_xyz[x];
            const int x0 = xyz & 1023, y0 = (xyz >> 10) & 1023, z0 = (xyz >> 20) & 1023;
            int sites[4];
            bool has_a = false;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q > nslot) {
                    sites[q] = -1;
                    continue;
                }
                int X = wrap(x0 + rb[3 * q], g.box[0]);
                int Y = wrap(y0 + rb[3 * q + 1], g.box[1]);
                int Z = wrap(z0 + rb[3 * q + 2], g.box[2]);
                sites[q] = site_index(g, X, Y, Z);
                has_a |= (sites[q] == p.a);
            }
            if (x == p.b && has_a) continue;  // a path through both sites counts once
            int eo = 0, en = 0, pw = 1;
#pragma unroll
            for (int q = 1; q < 4; ++q) {
                if (q > nslot) break;
                const int st = sites[q];
                const int so = lsp[st];
                const int sn = (st == p.a) ? p.na : ((st == p.b) ? p.nb : so);
                eo += so * pw;
                en += sn * pw;
                pw *= g.nsp;
            }
            const int sa0 = lsp[sites[0]];
            const int sa1 = (sites[0] == p.a) ? p.na : ((sites[0] == p.b) ? p.nb : sa0);
            eo += sa0 * g.nrow_e + g.nss_rows + cls_off;
            en += sa1 * g.nrow_e + g.nss_rows + cls_off;
            if (eo != en) {
                atomicAdd(&hist[eo], -wgt);
                atomicAdd(&hist[en], wgt);
            }
        }
    }
    __syncthreads();
    int *dc = dcounts + (long)s * g.nrow_tot;
    for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x)
        if (hist[q] != 0) atomicAdd(&dc[q], hist[q]);
}

// ---------------------------------------------------------------- full counts
// grid (nsplit, nrep).  counts must be zero on entry (the const rows are set here).
__global__ __launch_bounds__(256) void k_full_counts(
    Geo g, const uint8_t *__restrict__ species, const int *__restrict__ site_xyz,
    const int *__restrict__ nbr_j, const int *__restrict__ nbr_sh,
    const int4 *__restrict__ stencil, int *__restrict__ counts) {
    extern __shared__ int lds[];
    int *hist = lds;
    uint8_t *lsp = reinterpret_cast<uint8_t *>(lds + ((g.nrow_tot + 3) & ~3));
    const int s = blockIdx.y;
    load_species(lsp, species + (long)s * g.natoms_pad, g.natoms_pad);
    for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x) hist[q] = 0;
    __syncthreads();
    const long npairs = (long)g.natoms * g.ncoord;
    const long nms = (long)g.natoms * g.nabsrec;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long it = (long)blockIdx.x * blockDim.x + threadIdx.x; it < npairs + nms; it += stride) {
        if (it < npairs) {
            const int x = (int)(it / g.ncoord);
            const int y = nbr_j[it];
            atomicAdd(&hist[lsp[x] * g.nrow_e + lsp[y] * g.nshell + nbr_sh[it]], 1);
        } else {
            const long m = it - npairs;
            const int x = (int)(m / g.nabsrec);
            const int4 rec = stencil[m % g.nabsrec];
            const int8_t *rb = reinterpret_cast<const int8_t *>(&rec);
            const int nslot = rb[12], wgt = rb[13];
            const int cls_off = (int)(uint8_t)rb[14] | ((int)rb[15] << 8);
            const int xyz = site_xyz[x];
            const int x0 = xyz & 1023, y0 = (xyz >> 10) & 1023, z0 = (xyz >> 20) & 1023;
            int e = lsp[x] * g.nrow_e + g.nss_rows + cls_off;
            int pw = 1;
            for (int q = 1; q <= nslot; ++q) {
                int X = wrap(x0 + rb[3 * q], g.box[0]);
                int Y = wrap(y0 + rb[3 * q + 1], g.box[1]);
                int Z = wrap(z0 + rb[3 * q + 2], g.box[2]);
                e += lsp[site_index(g, X, Y, Z)] * pw;
                pw *= g.nsp;
            }
            atomicAdd(&hist[e], wgt);
        }
    }
    __syncthreads();
    int *dc = counts + (long)s * g.nrow_tot;
    for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x) {
        int v = hist[q];
        if (g.has_const && blockIdx.x == 0 && (q % g.nrow_e) == g.nrow_e - 1) v += 1;
        if (v != 0) atomicAdd(&dc[q], v);
    }
}

// ---------------------------------------------------------------- K2
// mtrial[s][e*ncol_e + col] = sum_k (counts + dcounts)[s][e*nrow_e + k] * basisR[e][k][col]
// Workgroup: 16 columns x 16 K-slices, RB replica rows held per lane.
constexpr int RB = 32;
constexpr int NCW = 16;
constexpr int KS = 16;

__global__ __launch_bounds__(256) void k_trial_gemm(Geo g, int nrep, const int *__restrict__ counts,
                                                     const int *__restrict__ dcounts,
                                                     const double *__restrict__ basisR,
                                                     double *__restrict__ mtrial) {
    extern __shared__ double dl[];
    const int e = blockIdx.z;
    const int r0 = blockIdx.y * RB;
    const int c0 = blockIdx.x * NCW;
    const int K = g.nrow_e;
    const int ld = RB + 1;
    // stage the coefficients of this edge for RB replicas: As[k][m]
    for (int q = threadIdx.x; q < RB * K; q += blockDim.x) {
        int m = q / K, k = q % K;
        int slot = r0 + m;
        double v = 0.0;
        if (slot < nrep) {
            long o = (long)slot * g.nrow_tot + (long)e * K + k;
            v = (double)(counts[o] + dcounts[o]);
        }
        dl[k * ld + m] = v;
    }
    __syncthreads();
    const int tn = threadIdx.x % NCW, ks = threadIdx.x / NCW;
    const int col = c0 + tn;
    const int kc = (K + KS - 1) / KS;
    const int kb = ks * kc, ke = min(K, kb + kc);
    double acc[RB];
#pragma unroll
    for (int m = 0; m < RB; ++m) acc[m] = 0.0;
    if (col < g.ncol_e) {
        const double *Bc = basisR + (long)e * K * g.ncol_e + col;
        for (int k = kb; k < ke; ++k) {
            const double bv = Bc[(long)k * g.ncol_e];
            const double *Ak = dl + k * ld;
#pragma unroll
            for (int m = 0; m < RB; ++m) acc[m] = fma(Ak[m], bv, acc[m]);
        }
    }
    __syncthreads();
    // partials P[ks][m][tn]
    double *P = dl;
#pragma unroll
    for (int m = 0; m < RB; ++m) P[(ks * RB + m) * NCW + tn] = acc[m];
    __syncthreads();
    for (int q = threadIdx.x; q < RB * NCW; q += blockDim.x) {
        int m = q / NCW, t = q % NCW;
        int slot = r0 + m, cc = c0 + t;
        if (slot >= nrep || cc >= g.ncol_e) continue;
        double sum = 0.0;
        for (int z = 0; z < KS; ++z) sum += P[(z * RB + m) * NCW + t];
        mtrial[(long)slot * g.ncol_tot + (long)e * g.ncol_e + cc] = sum;
    }
}

// ---------------------------------------------------------------- K3
struct AccArgs {
    const double *target;   // (ncol_tot) fit window [re, im] per (chan, j)
    const double *wsig;     // (ncol_tot/2) 1/sigma^2
    const double *mtrial;   // (nrep, ncol_tot)
    const double *beta;     // per rung
    const int *prop;        // (nrep, 5)
    int *counts, *dcounts;
    uint8_t *species, *best_sp;
    double *c2, *best_c2;
    const int *rung_of;
    uint64_t *stepc;
    long long *acc_rung;
    int *log;               // (log_cap, nrep, 4) or null
    int log_cap;
    int *series;            // (series_cap, nsp*nsp*nshell) or null
    int series_cap, series_rung;
    uint64_t seed;
    int profile;
    int mode;               // 0 step, 1 init, 2 evaluate rows of mtrial into c2
    double nfit;
};

__device__ double block_sum(double v, double *red) {
    // fixed-order tree over the 256 lanes
    __syncthreads();
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_accept(Geo g, AccArgs A) {
    __shared__ double red[256];
    __shared__ double scale[chm::kMaxNsp];
    __shared__ int accd;
    const int s = blockIdx.x;
    const double *m = A.mtrial + (long)s * g.ncol_tot;
    const int ncplx_e = g.ncol_e / 2;
    // per-edge profiled amplitude (method_a_rmc.py:190-209)
    for (int e = 0; e < g.nsp; ++e) {
        double num = 0.0, den = 0.0;
        if (A.profile) {
            for (int q = threadIdx.x; q < ncplx_e; q += blockDim.x) {
                long o = (long)e * g.ncol_e + 2 * q;
                double mr = m[o], mi = m[o + 1];
                num += mr * A.target[o] + mi * A.target[o + 1];
                den += mr * mr + mi * mi;
            }
            num = block_sum(num, red);
            den = block_sum(den, red);
        }
        if (threadIdx.x == 0) scale[e] = A.profile ? num / fmax(den, 1e-30) : 1.0;
    }
    __syncthreads();
    double res = 0.0;
    for (int q = threadIdx.x; q < g.ncol_tot / 2; q += blockDim.x) {
        int e = q / ncplx_e;
        double se = scale[e];
        double dr = se * m[2 * q] - A.target[2 * q];
        double di = se * m[2 * q + 1] - A.target[2 * q + 1];
        res += (dr * dr + di * di) * A.wsig[q];
    }
    res = block_sum(res, red);
    const double c2t = res / A.nfit;
    const int *pp = A.prop + 5 * s;
    const int rung = A.rung_of[s];
    uint64_t step = A.stepc[s] + 1;
    if (A.mode == 2) {
        if (threadIdx.x == 0) A.c2[s] = c2t;
        return;
    }
    if (A.mode == 1) {
        if (threadIdx.x == 0) {
            A.c2[s] = c2t;
            A.best_c2[s] = c2t;
        }
        for (int q = threadIdx.x; q < g.natoms_pad; q += blockDim.x)
            A.best_sp[(long)s * g.natoms_pad + q] = A.species[(long)s * g.natoms_pad + q];
        return;
    }
    if (threadIdx.x == 0) {
        int acc = 0;
        if (pp[4]) {
            double x = A.beta[rung] * (c2t - A.c2[s]) * A.nfit;
            uint64_t h = splitmix64(A.seed ^ (step * kGolden) ^ ((uint64_t)rung << 32) ^ kAccSalt);
            double u = u01(splitmix64(h));
            acc = (x <= 0.0) || (u < exp(-fmin(x, 60.0)));
        }
        accd = acc;
        if (acc) {
            A.species[(long)s * g.natoms_pad + pp[0]] = (uint8_t)pp[2];
            A.species[(long)s * g.natoms_pad + pp[1]] = (uint8_t)pp[3];
            A.c2[s] = c2t;
            A.acc_rung[rung] += 1;
        }
        if (A.log && step - 1 < (uint64_t)A.log_cap) {
            int *L = A.log + ((long)(step - 1) * gridDim.x + s) * 4;
            L[0] = pp[0];
            L[1] = pp[1];
            L[2] = acc;
            L[3] = rung;
        }
        A.stepc[s] = step;
    }
    __syncthreads();
    const int acc = accd;
    int *cn = A.counts + (long)s * g.nrow_tot;
    int *dc = A.dcounts + (long)s * g.nrow_tot;
    for (int q = threadIdx.x; q < g.nrow_tot; q += blockDim.x) {
        if (acc) cn[q] += dc[q];
        dc[q] = 0;
    }
    __syncthreads();
    if (A.series && rung == A.series_rung && step - 1 < (uint64_t)A.series_cap) {
        const int nss = g.nsp * g.nsp * g.nshell;
        for (int q = threadIdx.x; q < nss; q += blockDim.x) {
            int a = q / (g.nsp * g.nshell), r = q % (g.nsp * g.nshell);
            A.series[(long)(step - 1) * nss + q] = cn[a * g.nrow_e + r];
        }
    }
    const double cur = A.c2[s];
    if (cur < A.best_c2[s]) {
        __syncthreads();
        for (int q = threadIdx.x; q < g.natoms_pad; q += blockDim.x)
            A.best_sp[(long)s * g.natoms_pad + q] = A.species[(long)s * g.natoms_pad + q];
        __syncthreads();
        if (threadIdx.x == 0) A.best_c2[s] = cur;
    }
}

// ---------------------------------------------------------------- K4
__global__ void k_exchange(int nrep, const double *__restrict__ beta, const double *__restrict__ c2,
                           int *__restrict__ slot_of, int *__restrict__ rung_of,
                           uint64_t *__restrict__ iex, long long *__restrict__ att,
                           long long *__restrict__ acc, uint64_t seed, double nfit) {
    __shared__ uint64_t hx;
    if (threadIdx.x == 0) {
        uint64_t n = iex[0] + 1;
        iex[0] = n;
        hx = splitmix64(seed ^ (n * kGolden) ^ kExcSalt);
    }
    __syncthreads();
    const int start = (int)(hx >> 63);
    for (int i = start + 2 * threadIdx.x; i < nrep - 1; i += 2 * blockDim.x) {
        int si = slot_of[i], sj = slot_of[i + 1];
        double d = (beta[i] - beta[i + 1]) * (c2[si] - c2[sj]) * nfit;
        double u = u01(splitmix64(hx ^ ((uint64_t)(i + 1) << 32)));
        bool ok = (d >= 0.0) || (u < exp(fmax(d, -60.0)));
        att[i] += 1;
        if (ok) {
            acc[i] += 1;
            slot_of[i] = sj;
            slot_of[i + 1] = si;
            rung_of[sj] = i;
            rung_of[si] = i + 1;
        }
    }
}

// chi(k) of every slot: counts @ basis_k per edge (for reading back)
__global__ void k_counts_to_f64(const int *__restrict__ counts, double *__restrict__ out, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (double)counts[t];
}

}  // namespace

struct chm_sampler {
    chm_ctx *ctx = nullptr;
    Geo g{};
    int nrep = 0, nk = 0, nsplit = 1, profile = 1;
    uint64_t seed = 0;
    bool inited = false;
    long long hstep = 0;
    double nfit = 1.0;
    size_t lds_k1 = 0;
    chm::DevBuf site_xyz, nbr_j, nbr_sh, stencil, basisR, basisK, target, wsig;
    chm::DevBuf species, counts, dcounts, c2, best_c2, best_sp, rung_of, slot_of, beta, stepc, iex,
        prop, mtrial, exch_att, exch_acc, acc_rung, log, series, tmp;
    int log_cap = 0, series_cap = 0, series_rung = -1;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
};

namespace {

AccArgs acc_args(chm_sampler *S, int mode) {
    AccArgs A;
    A.target = S->target.as<double>();
    A.wsig = S->wsig.as<double>();
    A.mtrial = S->mtrial.as<double>();
    A.beta = S->beta.as<double>();
    A.prop = S->prop.as<int>();
    A.counts = S->counts.as<int>();
    A.dcounts = S->dcounts.as<int>();
    A.species = S->species.as<uint8_t>();
    A.best_sp = S->best_sp.as<uint8_t>();
    A.c2 = S->c2.as<double>();
    A.best_c2 = S->best_c2.as<double>();
    A.rung_of = S->rung_of.as<int>();
    A.stepc = S->stepc.as<uint64_t>();
    A.acc_rung = S->acc_rung.as<long long>();
    A.log = S->log_cap ? S->log.as<int>() : nullptr;
    A.log_cap = S->log_cap;
    A.series = S->series_cap ? S->series.as<int>() : nullptr;
    A.series_cap = S->series_cap;
    A.series_rung = S->series_rung;
    A.seed = S->seed;
    A.profile = S->profile;
    A.mode = mode;
    A.nfit = S->nfit;
    return A;
}

size_t lds_gemm(const Geo &g) {
    size_t a = sizeof(double) * (size_t)g.nrow_e * (RB + 1);
    size_t p = sizeof(double) * (size_t)KS * RB * NCW;
    return std::max(a, p);
}

int launch_gemm(chm_sampler *S) {
    const Geo &g = S->g;
    dim3 grid((g.ncol_e + NCW - 1) / NCW, (S->nrep + RB - 1) / RB, g.nsp);
    k_trial_gemm<<<grid, 256, lds_gemm(g), S->ctx->stream>>>(
        g, S->nrep, S->counts.as<int>(), S->dcounts.as<int>(), S->basisR.as<double>(),
        S->mtrial.as<double>());
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int launch_step(chm_sampler *S) {
    const Geo &g = S->g;
    hipStream_t st = S->ctx->stream;
    k_propose_gather<<<dim3(S->nsplit, S->nrep), 256, S->lds_k1, st>>>(
        g, S->species.as<uint8_t>(), S->rung_of.as<int>(), S->stepc.as<uint64_t>(), S->seed,
        S->site_xyz.as<int>(), S->nbr_j.as<int>(), S->nbr_sh.as<int>(), S->stencil.as<int4>(),
        S->dcounts.as<int>(), S->prop.as<int>());
    CHM_LAUNCH_CHECK();
    int rc = launch_gemm(S);
    if (rc) return rc;
    k_accept<<<dim3(S->nrep), 256, 0, st>>>(g, acc_args(S, 0));
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int launch_exchange(chm_sampler *S) {
    k_exchange<<<1, 256, 0, S->ctx->stream>>>(
        S->nrep, S->beta.as<double>(), S->c2.as<double>(), S->slot_of.as<int>(),
        S->rung_of.as<int>(), S->iex.as<uint64_t>(), S->exch_att.as<long long>(),
        S->exch_acc.as<long long>(), S->seed, S->nfit);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

void drop_graph(chm_sampler *S) {
    if (S->gexec) hipGraphExecDestroy(S->gexec);
    if (S->graph) hipGraphDestroy(S->graph);
    S->gexec = nullptr;
    S->graph = nullptr;
    S->graph_steps = 0;
}

}  // namespace

extern "C" int chm_sampler_create(chm_ctx *ctx, const chm_geom *gm, int nrep, int nk,
                                  const int32_t *site_xyz, const int32_t *nbr_j,
                                  const int32_t *nbr_shell, const int8_t *stencil,
                                  const double *basis_R, const double *basis_k,
                                  const double *target, const double *sigma, int profile_scale,
                                  chm_sampler **out) {
    CHM_REQUIRE(ctx && gm && out, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(nrep >= 1 && nk >= 2, CHM_E_ARG, "bad nrep / nk");
    CHM_REQUIRE(gm->nsp >= 2 && gm->nsp <= chm::kMaxNsp, CHM_E_ARG, "2..6 species");
    CHM_REQUIRE(gm->natoms > 0 && gm->ncoord > 0 && gm->nshell > 0, CHM_E_ARG, "bad geometry");
    CHM_REQUIRE(gm->nkw > 0 && gm->nwin > 0, CHM_E_ARG, "bad window");
    CHM_REQUIRE(gm->nblk >= 0 && gm->nstencil >= 0 && gm->nabsrec >= 0 &&
                    gm->nabsrec <= gm->nstencil,
                CHM_E_ARG, "bad stencil sizes");
    CHM_REQUIRE(gm->nblk == 0 || gm->nstencil > 0, CHM_E_ARG, "MS entries need a stencil");
    CHM_REQUIRE(site_xyz && nbr_j && nbr_shell && basis_R && basis_k && target && sigma,
                CHM_E_ARG, "null input");
    CHM_REQUIRE(gm->nstencil == 0 || stencil, CHM_E_ARG, "null stencil");
    for (int d = 0; d < 3; ++d)
        CHM_REQUIRE(gm->box[d] == 2 * gm->ncell[d] && gm->box[d] < 1024, CHM_E_ARG,
                    "box must be 2*ncell (< 1024)");
    CHM_REQUIRE(gm->nbasis >= 1 && gm->natoms == gm->nbasis * gm->ncell[0] * gm->ncell[1] *
                                                     gm->ncell[2],
                CHM_E_ARG, "natoms != nbasis * cells");
    Geo g{};
    g.natoms = gm->natoms;
    g.natoms_pad = (gm->natoms + 15) & ~15;
    g.nsp = gm->nsp;
    g.ncoord = gm->ncoord;
    g.nshell = gm->nshell;
    g.nblk = gm->nblk;
    g.nstencil = gm->nstencil;
    g.nabsrec = gm->nabsrec;
    g.nss_rows = gm->nsp * gm->nshell;
    g.has_const = gm->has_const ? 1 : 0;
    g.nrow_e = g.nss_rows + g.nblk + g.has_const;
    g.nrow_tot = g.nsp * g.nrow_e;
    g.nkw = gm->nkw;
    g.nwin = gm->nwin;
    g.ncol_e = gm->nkw * gm->nwin * 2;
    g.ncol_tot = g.nsp * g.ncol_e;
    g.nbasis = gm->nbasis;
    for (int d = 0; d < 3; ++d) {
        g.box[d] = gm->box[d];
        g.ncell[d] = gm->ncell[d];
    }
    for (int q = 0; q < 8; ++q) g.bpar[q] = gm->basis_of_parity[q];
    // host-side validation of every index a kernel will follow
    for (long q = 0; q < (long)g.natoms * g.ncoord; ++q) {
        CHM_REQUIRE(nbr_j[q] >= 0 && nbr_j[q] < g.natoms, CHM_E_ARG, "neighbour out of range");
        CHM_REQUIRE(nbr_shell[q] >= 0 && nbr_shell[q] < g.nshell, CHM_E_ARG, "shell out of range");
    }
    for (int q = 0; q < g.natoms; ++q) {
        int v = site_xyz[q];
        int x = v & 1023, y = (v >> 10) & 1023, z = (v >> 20) & 1023;
        CHM_REQUIRE(x < g.box[0] && y < g.box[1] && z < g.box[2], CHM_E_ARG, "site outside box");
        int b = g.bpar[(x & 1) | ((y & 1) << 1) | ((z & 1) << 2)];
        int cell = ((x >> 1) * g.ncell[1] + (y >> 1)) * g.ncell[2] + (z >> 1);
        CHM_REQUIRE(b >= 0 && cell * g.nbasis + b == q, CHM_E_ARG,
                    "site coordinates do not map back to their index");
    }
    for (int q = 0; q < g.nstencil; ++q) {
        const int8_t *rb = stencil + 16 * q;
        int nslot = rb[12], w = rb[13];
        int cls_off = (int)(uint8_t)rb[14] | ((int)rb[15] << 8);
        CHM_REQUIRE(nslot >= 1 && nslot <= 3 && w >= 1 && w <= 2, CHM_E_ARG, "bad stencil record");
        int pw = 1;
        for (int z = 0; z < nslot; ++z) pw *= g.nsp;
        CHM_REQUIRE(cls_off >= 0 && cls_off + pw <= g.nblk, CHM_E_ARG, "stencil entry out of range");
        for (int t = 0; t < 3 * (nslot + 1); ++t)
            CHM_REQUIRE(rb[t] > -g.box[t % 3] && rb[t] < g.box[t % 3], CHM_E_ARG,
                        "stencil offset exceeds the box");
        CHM_REQUIRE((q < g.nabsrec) == (rb[0] == 0 && rb[1] == 0 && rb[2] == 0), CHM_E_ARG,
                    "absorber-role records must come first");
    }
    CHM_HIP(hipSetDevice(ctx->device));
    auto *S = new chm_sampler();
    S->ctx = ctx;
    S->g = g;
    S->nrep = nrep;
    S->nk = nk;
    S->profile = profile_scale ? 1 : 0;
    S->nfit = (double)g.ncol_tot;  // nsp*nkw*nwin complex points, Re and Im counted
    int items = 2 * g.ncoord + 2 * g.nstencil;
    S->nsplit = std::max(1, std::min(16, (items + 511) / 512));
    S->lds_k1 = sizeof(int) * ((g.nrow_tot + 3) & ~3) + (size_t)g.natoms_pad;
    auto fail = [&](int rc) {
        delete S;
        return rc;
    };
    auto up = [&](chm::DevBuf &b, const void *h, size_t n) -> int {
        int rc = b.ensure(n);
        if (rc) return rc;
        CHM_HIP(hipMemcpyAsync(b.p, h, n, hipMemcpyHostToDevice, ctx->stream));
        return CHM_OK;
    };
    int rc;
    // wsig = 1/sigma^2, target already restricted to the window
    std::vector<double> ws(g.ncol_tot / 2);
    for (int q = 0; q < g.ncol_tot / 2; ++q) {
        if (!(sigma[q] > 0.0)) return fail((chm::set_error("sigma must be positive"), CHM_E_ARG));
        ws[q] = 1.0 / (sigma[q] * sigma[q]);
    }
    if ((rc = up(S->site_xyz, site_xyz, sizeof(int) * g.natoms)) ||
        (rc = up(S->nbr_j, nbr_j, sizeof(int) * g.natoms * g.ncoord)) ||
        (rc = up(S->nbr_sh, nbr_shell, sizeof(int) * g.natoms * g.ncoord)) ||
        (rc = up(S->stencil, g.nstencil ? (const void *)stencil : (const void *)"", 16 * (size_t)g.nstencil)) ||
        (rc = up(S->basisR, basis_R, sizeof(double) * (size_t)g.nsp * g.nrow_e * g.ncol_e)) ||
        (rc = up(S->basisK, basis_k, sizeof(double) * (size_t)g.nsp * g.nrow_e * nk)) ||
        (rc = up(S->target, target, sizeof(double) * g.ncol_tot)) ||
        (rc = up(S->wsig, ws.data(), sizeof(double) * ws.size())))
        return fail(rc);
    size_t nr = nrep;
    if ((rc = S->species.ensure(nr * g.natoms_pad)) || (rc = S->best_sp.ensure(nr * g.natoms_pad)) ||
        (rc = S->counts.ensure(sizeof(int) * nr * g.nrow_tot)) ||
        (rc = S->dcounts.ensure(sizeof(int) * nr * g.nrow_tot)) ||
        (rc = S->c2.ensure(sizeof(double) * nr)) || (rc = S->best_c2.ensure(sizeof(double) * nr)) ||
        (rc = S->rung_of.ensure(sizeof(int) * nr)) || (rc = S->slot_of.ensure(sizeof(int) * nr)) ||
        (rc = S->beta.ensure(sizeof(double) * nr)) || (rc = S->stepc.ensure(sizeof(uint64_t) * nr)) ||
        (rc = S->iex.ensure(sizeof(uint64_t))) || (rc = S->prop.ensure(sizeof(int) * 5 * nr)) ||
        (rc = S->mtrial.ensure(sizeof(double) * nr * g.ncol_tot)) ||
        (rc = S->exch_att.ensure(sizeof(long long) * nr)) ||
        (rc = S->exch_acc.ensure(sizeof(long long) * nr)) ||
        (rc = S->acc_rung.ensure(sizeof(long long) * nr)))
        return fail(rc);
    hipFuncSetAttribute((const void *)k_trial_gemm, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds_gemm(g));
    hipFuncSetAttribute((const void *)k_propose_gather,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)S->lds_k1);
    hipFuncSetAttribute((const void *)k_full_counts, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)S->lds_k1);
    if (lds_gemm(g) > 160 * 1024 || S->lds_k1 > 160 * 1024)
        return fail((chm::set_error("problem too large for LDS staging"), CHM_E_ARG));
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        chm::set_error(std::string("sampler upload: ") + hipGetErrorString(e));
        return fail(CHM_E_HIP);
    }
    *out = S;
    return CHM_OK;
}

extern "C" int chm_sampler_destroy(chm_sampler *S) {
    if (!S) return CHM_OK;
    hipSetDevice(S->ctx->device);
    hipStreamSynchronize(S->ctx->stream);
    drop_graph(S);
    delete S;
    return CHM_OK;
}

extern "C" int chm_sampler_set_recording(chm_sampler *S, int log_capacity, int series_capacity,
                                         int series_rung) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    CHM_REQUIRE(log_capacity >= 0 && series_capacity >= 0, CHM_E_ARG, "negative capacity");
    CHM_REQUIRE(series_capacity == 0 || (series_rung >= 0 && series_rung < S->nrep), CHM_E_ARG,
                "series rung out of range");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    const Geo &g = S->g;
    int rc;
    if ((rc = S->log.ensure(sizeof(int) * 4 * (size_t)log_capacity * S->nrep))) return rc;
    if ((rc = S->series.ensure(sizeof(int) * (size_t)series_capacity * g.nsp * g.nsp * g.nshell)))
        return rc;
    S->log_cap = log_capacity;
    S->series_cap = series_capacity;
    S->series_rung = series_rung;
    if (log_capacity) CHM_HIP(hipMemset(S->log.p, 0xff, S->log.bytes));
    return CHM_OK;
}

extern "C" int chm_sampler_init(chm_sampler *S, const int32_t *species0, const double *beta,
                                uint64_t seed) {
    CHM_REQUIRE(S && species0 && beta, CHM_E_ARG, "bad arguments");
    const Geo &g = S->g;
    for (int q = 0; q < g.natoms; ++q)
        CHM_REQUIRE(species0[q] >= 0 && species0[q] < g.nsp, CHM_E_ARG, "species out of range");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    drop_graph(S);
    const int nrep = S->nrep;
    std::vector<uint8_t> sp((size_t)nrep * g.natoms_pad, 0);
    for (int r = 0; r < nrep; ++r)
        for (int q = 0; q < g.natoms; ++q) sp[(size_t)r * g.natoms_pad + q] = (uint8_t)species0[q];
    std::vector<int> ident(nrep);
    for (int r = 0; r < nrep; ++r) ident[r] = r;
    CHM_HIP(hipMemcpyAsync(S->species.p, sp.data(), sp.size(), hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->beta.p, beta, sizeof(double) * nrep, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->rung_of.p, ident.data(), sizeof(int) * nrep, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->slot_of.p, ident.data(), sizeof(int) * nrep, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemsetAsync(S->counts.p, 0, S->counts.bytes, st));
    CHM_HIP(hipMemsetAsync(S->dcounts.p, 0, S->dcounts.bytes, st));
    CHM_HIP(hipMemsetAsync(S->stepc.p, 0, S->stepc.bytes, st));
    CHM_HIP(hipMemsetAsync(S->iex.p, 0, S->iex.bytes, st));
    CHM_HIP(hipMemsetAsync(S->prop.p, 0, S->prop.bytes, st));
    CHM_HIP(hipMemsetAsync(S->exch_att.p, 0, S->exch_att.bytes, st));
    CHM_HIP(hipMemsetAsync(S->exch_acc.p, 0, S->exch_acc.bytes, st));
    CHM_HIP(hipMemsetAsync(S->acc_rung.p, 0, S->acc_rung.bytes, st));
    if (S->log_cap) CHM_HIP(hipMemsetAsync(S->log.p, 0xff, S->log.bytes, st));
    S->seed = seed;
    S->hstep = 0;
    long long npairs = (long long)g.natoms * g.ncoord;
    long long nitems = npairs + (long long)g.natoms * g.nabsrec;
    int nsplit = (int)std::max(1LL, std::min(64LL, nitems / 8192 + 1));
    k_full_counts<<<dim3(nsplit, nrep), 256, S->lds_k1, st>>>(
        g, S->species.as<uint8_t>(), S->site_xyz.as<int>(), S->nbr_j.as<int>(), S->nbr_sh.as<int>(),
        S->stencil.as<int4>(), S->counts.as<int>());
    CHM_LAUNCH_CHECK();
    int rc = launch_gemm(S);
    if (rc) return rc;
    k_accept<<<dim3(nrep), 256, 0, st>>>(g, acc_args(S, 1));
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipStreamSynchronize(st));
    S->inited = true;
    return CHM_OK;
}

extern "C" int chm_sampler_run(chm_sampler *S, int nsteps, int exchange_every) {
    CHM_REQUIRE(S && S->inited, CHM_E_STATE, "init_state first");
    CHM_REQUIRE(nsteps >= 0 && exchange_every >= 0, CHM_E_ARG, "negative step count");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    const int E = exchange_every;
    int rc = 0;
    if (E > 0 && (S->graph_steps != E)) {
        drop_graph(S);
        CHM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        for (int t = 0; t < E && !rc; ++t) rc = launch_step(S);
        if (!rc) rc = launch_exchange(S);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(st, &gr);
        if (rc) {
            if (gr) hipGraphDestroy(gr);
            return rc;
        }
        if (e != hipSuccess) {
            chm::set_error(std::string("graph capture: ") + hipGetErrorString(e));
            return CHM_E_HIP;
        }
        S->graph = gr;
        CHM_HIP(hipGraphInstantiate(&S->gexec, gr, nullptr, nullptr, 0));
        S->graph_steps = E;
    }
    long long left = nsteps;
    while (left > 0) {
        if (E > 0 && S->hstep % E == 0 && left >= E) {
            CHM_HIP(hipGraphLaunch(S->gexec, st));
            S->hstep += E;
            left -= E;
        } else {
            if ((rc = launch_step(S))) return rc;
            S->hstep += 1;
            left -= 1;
            if (E > 0 && S->hstep % E == 0 && (rc = launch_exchange(S))) return rc;
        }
    }
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_sampler_read(chm_sampler *S, int what, void *host, int64_t nbytes) {
    CHM_REQUIRE(S && host, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    CHM_HIP(hipStreamSynchronize(st));
    const Geo &g = S->g;
    const int nrep = S->nrep;
    std::vector<int> slot_of(nrep);
    CHM_HIP(hipMemcpy(slot_of.data(), S->slot_of.p, sizeof(int) * nrep, hipMemcpyDeviceToHost));
    auto need = [&](int64_t n) -> int {
        CHM_REQUIRE(nbytes >= n, CHM_E_ARG, "host buffer too small");
        return CHM_OK;
    };
    int rc;
    switch (what) {
        case 0:
        case 4: {
            if ((rc = need((int64_t)nrep * g.natoms * 4))) return rc;
            std::vector<uint8_t> sp((size_t)nrep * g.natoms_pad);
            CHM_HIP(hipMemcpy(sp.data(), what == 0 ? S->species.p : S->best_sp.p, sp.size(),
                              hipMemcpyDeviceToHost));
            int32_t *o = static_cast<int32_t *>(host);
            for (int r = 0; r < nrep; ++r) {
                int s = (what == 0 || what == 4) ? slot_of[r] : r;
                for (int q = 0; q < g.natoms; ++q)
                    o[(size_t)r * g.natoms + q] = sp[(size_t)s * g.natoms_pad + q];
            }
            return CHM_OK;
        }
        case 1:
        case 3: {
            if ((rc = need((int64_t)nrep * 8))) return rc;
            std::vector<double> v(nrep);
            CHM_HIP(hipMemcpy(v.data(), what == 1 ? S->c2.p : S->best_c2.p, 8 * nrep,
                              hipMemcpyDeviceToHost));
            double *o = static_cast<double *>(host);
            for (int r = 0; r < nrep; ++r) o[r] = v[slot_of[r]];
            return CHM_OK;
        }
        case 2: {
            if ((rc = need((int64_t)nrep * g.nrow_tot * 4))) return rc;
            std::vector<int> v((size_t)nrep * g.nrow_tot);
            CHM_HIP(hipMemcpy(v.data(), S->counts.p, 4 * v.size(), hipMemcpyDeviceToHost));
            int *o = static_cast<int *>(host);
            for (int r = 0; r < nrep; ++r)
                memcpy(o + (size_t)r * g.nrow_tot, v.data() + (size_t)slot_of[r] * g.nrow_tot,
                       4 * g.nrow_tot);
            return CHM_OK;
        }
        case 5: {
            if ((rc = need((int64_t)nrep * 4))) return rc;
            CHM_HIP(hipMemcpy(host, S->rung_of.p, 4 * nrep, hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 6: {
            int64_t n = (int64_t)4 * 4 * S->log_cap * nrep;
            if ((rc = need(std::min<int64_t>(n, nbytes)))) return rc;
            CHM_HIP(hipMemcpy(host, S->log.p, (size_t)std::min<int64_t>(n, nbytes),
                              hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 7: {
            int64_t n = (int64_t)4 * S->series_cap * g.nsp * g.nsp * g.nshell;
            CHM_HIP(hipMemcpy(host, S->series.p, (size_t)std::min<int64_t>(n, nbytes),
                              hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 8: {
            int64_t n = (int64_t)8 * 2 * std::max(nrep - 1, 1);
            if ((rc = need(n))) return rc;
            int64_t *o = static_cast<int64_t *>(host);
            int m = std::max(nrep - 1, 1);
            CHM_HIP(hipMemcpy(o, S->exch_att.p, 8 * m, hipMemcpyDeviceToHost));
            CHM_HIP(hipMemcpy(o + m, S->exch_acc.p, 8 * m, hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 9: {
            if ((rc = need((int64_t)8 * nrep))) return rc;
            CHM_HIP(hipMemcpy(host, S->acc_rung.p, 8 * nrep, hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 10: {
            if ((rc = need(8))) return rc;
            int64_t v = S->hstep;
            memcpy(host, &v, 8);
            return CHM_OK;
        }
        case 11: {
            // chi(k) = counts @ basis_k per edge, rung order
            int64_t n = (int64_t)8 * nrep * g.nsp * S->nk;
            if ((rc = need(n))) return rc;
            long nc = (long)nrep * g.nrow_tot;
            if ((rc = S->tmp.ensure(sizeof(double) * (nc + (size_t)nrep * g.nsp * S->nk)))) return rc;
            double *cf = S->tmp.as<double>();
            double *ck = cf + nc;
            k_counts_to_f64<<<dim3((unsigned)((nc + 255) / 256)), 256, 0, st>>>(S->counts.as<int>(),
                                                                              cf, nc);
            CHM_LAUNCH_CHECK();
            rc = chm::gemm_f64(st, g.nsp, nrep, S->nk, g.nrow_e, cf, g.nrow_tot, g.nrow_e,
                               S->basisK.as<double>(), S->nk, (long)g.nrow_e * S->nk, ck,
                               g.nsp * S->nk, S->nk);
            if (rc) return rc;
            std::vector<double> v((size_t)nrep * g.nsp * S->nk);
            CHM_HIP(hipMemcpyAsync(v.data(), ck, 8 * v.size(), hipMemcpyDeviceToHost, st));
            CHM_HIP(hipStreamSynchronize(st));
            double *o = static_cast<double *>(host);
            size_t row = (size_t)g.nsp * S->nk;
            for (int r = 0; r < nrep; ++r)
                memcpy(o + r * row, v.data() + (size_t)slot_of[r] * row, 8 * row);
            return CHM_OK;
        }
        case 12: {
            int64_t n = (int64_t)8 * nrep * g.ncol_tot;
            if ((rc = need(n))) return rc;
            CHM_HIP(hipMemcpy(host, S->mtrial.p, n, hipMemcpyDeviceToHost));
            return CHM_OK;
        }
    }
    chm::set_error("unknown read selector");
    return CHM_E_ARG;
}

extern "C" int chm_sampler_counts(chm_sampler *S, int nconf, const int32_t *species,
                                  int32_t *counts) {
    CHM_REQUIRE(S && nconf > 0 && species && counts, CHM_E_ARG, "bad arguments");
    const Geo &g = S->g;
    for (long q = 0; q < (long)nconf * g.natoms; ++q)
        CHM_REQUIRE(species[q] >= 0 && species[q] < g.nsp, CHM_E_ARG, "species out of range");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    chm::DevBuf sp, cn;
    int rc;
    if ((rc = sp.ensure((size_t)nconf * g.natoms_pad)) ||
        (rc = cn.ensure(sizeof(int) * (size_t)nconf * g.nrow_tot)))
        return rc;
    std::vector<uint8_t> h((size_t)nconf * g.natoms_pad, 0);
    for (int c = 0; c < nconf; ++c)
        for (int q = 0; q < g.natoms; ++q)
            h[(size_t)c * g.natoms_pad + q] = (uint8_t)species[(size_t)c * g.natoms + q];
    CHM_HIP(hipMemcpyAsync(sp.p, h.data(), h.size(), hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemsetAsync(cn.p, 0, cn.bytes, st));
    long long nitems = (long long)g.natoms * g.ncoord + (long long)g.natoms * g.nabsrec;
    int nsplit = (int)std::max(1LL, std::min(64LL, nitems / 8192 + 1));
    k_full_counts<<<dim3(nsplit, nconf), 256, S->lds_k1, st>>>(
        g, sp.as<uint8_t>(), S->site_xyz.as<int>(), S->nbr_j.as<int>(), S->nbr_sh.as<int>(),
        S->stencil.as<int4>(), cn.as<int>());
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(counts, cn.p, sizeof(int) * (size_t)nconf * g.nrow_tot,
                           hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_sampler_chi2(chm_sampler *S, int nrows, const double *chiR, double *c2) {
    CHM_REQUIRE(S && nrows > 0 && chiR && c2, CHM_E_ARG, "bad arguments");
    const Geo &g = S->g;
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    chm::DevBuf m, o;
    int rc;
    if ((rc = m.ensure(sizeof(double) * (size_t)nrows * g.ncol_tot)) ||
        (rc = o.ensure(sizeof(double) * nrows)))
        return rc;
    CHM_HIP(hipMemcpyAsync(m.p, chiR, sizeof(double) * (size_t)nrows * g.ncol_tot,
                           hipMemcpyHostToDevice, st));
    AccArgs A = acc_args(S, 2);
    A.mtrial = m.as<double>();
    A.c2 = o.as<double>();
    k_accept<<<dim3(nrows), 256, 0, st>>>(g, A);
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(c2, o.p, sizeof(double) * nrows, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_sampler_exchange(chm_sampler *S) {
    CHM_REQUIRE(S && S->inited, CHM_E_STATE, "init_state first");
    CHM_HIP(hipSetDevice(S->ctx->device));
    int rc = launch_exchange(S);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    return CHM_OK;
}

extern "C" int chm_sampler_chik(chm_sampler *S, int nconf, const int32_t *counts, double *chi) {
    CHM_REQUIRE(S && nconf > 0 && counts && chi, CHM_E_ARG, "bad arguments");
    const Geo &g = S->g;
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    chm::DevBuf cf, ck;
    std::vector<double> h((size_t)nconf * g.nrow_tot);
    for (size_t q = 0; q < h.size(); ++q) h[q] = (double)counts[q];
    int rc;
    if ((rc = cf.ensure(sizeof(double) * h.size())) ||
        (rc = ck.ensure(sizeof(double) * (size_t)nconf * g.nsp * S->nk)))
        return rc;
    CHM_HIP(hipMemcpyAsync(cf.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
    rc = chm::gemm_f64(st, g.nsp, nconf, S->nk, g.nrow_e, cf.as<double>(), g.nrow_tot, g.nrow_e,
                       S->basisK.as<double>(), S->nk, (long)g.nrow_e * S->nk, ck.as<double>(),
                       g.nsp * S->nk, S->nk);
    if (rc) return rc;
    CHM_HIP(hipMemcpyAsync(chi, ck.p, sizeof(double) * (size_t)nconf * g.nsp * S->nk,
                           hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_sampler_profile(chm_sampler *S, int iters, double *ms) {
    CHM_REQUIRE(S && S->inited && iters > 0 && ms, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    const Geo &g = S->g;
    hipEvent_t e0, e1;
    CHM_HIP(hipEventCreate(&e0));
    CHM_HIP(hipEventCreate(&e1));
    float t = 0.f;
    int rc = 0;
    // K1: the same proposal every launch (the step counter does not move); the
    // accumulated deltas are cleared afterwards
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i) {
        k_propose_gather<<<dim3(S->nsplit, S->nrep), 256, S->lds_k1, st>>>(
            g, S->species.as<uint8_t>(), S->rung_of.as<int>(), S->stepc.as<uint64_t>(), S->seed,
            S->site_xyz.as<int>(), S->nbr_j.as<int>(), S->nbr_sh.as<int>(), S->stencil.as<int4>(),
            S->dcounts.as<int>(), S->prop.as<int>());
    }
    CHM_HIP(hipEventRecord(e1, st));
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[0] = t / iters;
    CHM_HIP(hipMemsetAsync(S->dcounts.p, 0, S->dcounts.bytes, st));
    // K2 on the current state
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters && !rc; ++i) rc = launch_gemm(S);
    CHM_HIP(hipEventRecord(e1, st));
    if (rc) return rc;
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[1] = t / iters;
    // K3 with every proposal marked null (nothing is committed)
    CHM_HIP(hipMemsetAsync(S->prop.p, 0, S->prop.bytes, st));
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i) k_accept<<<dim3(S->nrep), 256, 0, st>>>(g, acc_args(S, 0));
    CHM_HIP(hipEventRecord(e1, st));
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[2] = t / iters;
    // one full step (K1..K3) launched directly
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters && !rc; ++i) rc = launch_step(S);
    CHM_HIP(hipEventRecord(e1, st));
    if (rc) return rc;
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[3] = t / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}
