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
#include <algorithm>
#include <cmath>
#include <cstring>

#include "common.h"
#include "sampler_dev.h"
#include "sampler_state.h"
#include "learned_dev.h"

namespace {

using namespace chm_smp;

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
            atomicAdd(&hist[lsp[x] * g.rstride + lsp[y] * g.nshell + nbr_sh[it]], 1);
        } else {
            const long m = it - npairs;
            const int x = (int)(m / g.nabsrec);
            const int4 rec = stencil[m % g.nabsrec];
            const int8_t *rb = reinterpret_cast<const int8_t *>(&rec);
            const int nslot = rb[12], wgt = rb[13];
            const int cls_off = (int)(uint8_t)rb[14] | ((int)rb[15] << 8);
            const int xyz = site_xyz[x];
            const int x0 = xyz & 1023, y0 = (xyz >> 10) & 1023, z0 = (xyz >> 20) & 1023;
            int e = lsp[x] * g.rstride + g.nss_rows + cls_off;
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
        if (g.has_const && blockIdx.x == 0 && (q % g.rstride) == g.nrow_e - 1) v += 1;
        if (v != 0) atomicAdd(&dc[q], v);
    }
}

// workgroups per configuration of k_full_counts: ~8192 items each, and enough workgroups
// over the device when few configurations are counted (each workgroup ends in one integer
// atomic per non-zero row, so the split stays modest)
int full_counts_split(const Geo &g, int nconf) {
    const long long nitems = (long long)g.natoms * g.ncoord + (long long)g.natoms * g.nabsrec;
    const long long cap = std::max(64, 256 / std::max(1, nconf));
    return (int)std::max(1LL, std::min(cap, nitems / 8192 + 1));
}

// ---------------------------------------------------------------- K2
// Both trial-transform kernels end in the same epilogue: for a 32-slot x 16-column tile
// of one edge, m' (the trial chi(R) window values) is written out and, per slot, four
// sums of the residual d = m' - t are reduced over the tile's 16 columns:
//   U1 = sum d t,  U2 = sum d^2,  W1 = sum w d t,  W2 = sum w d^2      (w = 1/sigma^2)
// The profiled amplitude s = sum m't / sum m'^2 (method_a_rmc.py:190-209) then has
// s - 1 = -(U1 + U2) / (U2 + 2 U1 + T0) and the edge's chi^2 nfit share is
// s^2 W2 + 2 s (s-1) W1 + (s-1)^2 Q0, with T0 = sum t^2, Q0 = sum w t^2: every term
// is small near a fit, so nothing cancels (the plain expansion in m' loses ~6 digits).
constexpr int RB = 32;    // replica slots per tile
constexpr int NCW = 16;   // columns per tile
constexpr int KS = 16;    // K slices of the FP64 kernel

struct TileOut {
    uint64_t *stepc;        // per slot: advanced by `bump` steps per transform
    int bump;
    const double *best_c2;  // per slot: the best chi^2, copied into best_snap for the next
    double *best_snap;      //   per-replica launch (null: no copy)
    int *draws;             // per slot (kDrawInts): the next per-replica launch's acceptance
    int draw_mode;          //   uniforms and proposal draws, made by an extra workgroup that
    const int *rung_of;     //   also advances the step counter (null: none; 1 one step,
    uint64_t seed;          //   2 a paired round)
    int natoms;
    const int *anchor;      // re-anchor mode: slots with anchor[s] >= anchor_min get
    int anchor_min;         //   anchor_out[s] = counts[s] @ basis_R (FP64); nothing else
    double *anchor_out;     //   is written.  Null in the normal mode.
    const double *target;   // (ncol_tot)
    const double *wcol;     // (ncol_tot) 1/sigma^2 of each column
    double *mtrial;         // (nrep, ncol_tot)
    double *partials;       // (nrep, nsp, nct, 4)
    int acc32;              // the running chi(R) is an FP32 accumulator (precision study)
};

// Every workgroup of a per-replica launch decides "new best?" against the best chi^2 before
// that launch, while the launch's lead workgroup overwrites best_c2; a workgroup that
// starts after the lead has finished would read the new value.  So the trial transform
// just before copies the value for them: workgroup (0, rows, edge 0) for its 32 slots.
__device__ __forceinline__ void snap_best(const TileOut &O, int nrep, int r0, int ct, int e) {
    if (O.best_snap && ct == 0 && e == 0 && threadIdx.x < RB && r0 + (int)threadIdx.x < nrep)
        O.best_snap[r0 + threadIdx.x] = O.best_c2[r0 + threadIdx.x];
}

// lanes t..t+15 (t % 16 == 0) hold the 16 columns of one slot
__device__ __forceinline__ void tile_epilogue(const Geo &g, const TileOut &O, int nrep, int slot,
                                              int e, int ct, int col, double mp) {
    const bool valid = slot < nrep && col < g.ncol_e;
    double p1 = 0.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
    if (valid) {
        const long gc = (long)e * g.ncol_e + col;
        O.mtrial[(long)slot * g.ncol_tot + gc] = mp;
        const double t = O.target[gc], w = O.wcol[gc];
        const double d = mp - t;
        p1 = d * t;
        p2 = d * d;
        q1 = p1 * w;
        q2 = p2 * w;
    }
    p1 = group_sum_d<16>(p1);
    p2 = group_sum_d<16>(p2);
    q1 = group_sum_d<16>(q1);
    q2 = group_sum_d<16>(q2);
    if (slot < nrep && (threadIdx.x & 15) == 0) {
        double *P = O.partials + (((long)slot * g.nsp + e) * g.nct + ct) * 4;
        P[0] = p1;
        P[1] = p2;
        P[2] = q1;
        P[3] = q2;
    }
}

// Absolute FP64: m' = (counts + delta) @ basis_R.  Workgroup: 16 columns x 16 K-slices,
// RB replica rows held per lane; the K-slice partials are summed in a fixed order.
__device__ __forceinline__ void gemm_abs_body(const Geo &g, int nrep, const int *__restrict__ counts,
                                              const int *__restrict__ dcounts,
                                              const double *__restrict__ basisR, const TileOut &O, int ct,
                                              int rbi, int e) {
    extern __shared__ double dl[];
    __shared__ int due[RB];
    __shared__ int any_due;
    const int r0 = rbi * RB;
    const int c0 = ct * NCW;
    const int K = g.nrow_e;
    const int ld = RB + 1;
    const bool reanchor = O.anchor_out != nullptr;
    if (!reanchor) snap_best(O, nrep, r0, ct, e);
    if (reanchor) {
        if (threadIdx.x == 0) any_due = 0;
        __syncthreads();
        if (threadIdx.x < RB) {
            const int slot = r0 + threadIdx.x;
            due[threadIdx.x] = slot < nrep && O.anchor[slot] >= O.anchor_min;
            if (due[threadIdx.x]) atomicOr(&any_due, 1);
        }
        __syncthreads();
        if (!any_due) return;
    }
    for (int q = threadIdx.x; q < RB * K; q += blockDim.x) {
        int m = q / K, k = q % K;
        int slot = r0 + m;
        double v = 0.0;
        if (slot < nrep) {
            long o = (long)slot * g.nrow_tot + (long)e * g.rstride + k;
            v = (double)(counts[o] + (dcounts ? dw_row(dcounts + (long)slot * g.nrow_w, e * g.rstride + k) : 0));
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
    double *P = dl;
#pragma unroll
    for (int m = 0; m < RB; ++m) P[(ks * RB + m) * NCW + tn] = acc[m];
    __syncthreads();
    // thread t: column t % 16, slots t/16 and t/16 + 16
    const int t = threadIdx.x;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = t / NCW + 16 * h, cl = t % NCW;
        double sum = 0.0;
        for (int z = 0; z < KS; ++z) sum += P[(z * RB + m) * NCW + cl];
        if (reanchor) {
            if (due[m] && c0 + cl < g.ncol_e)
                O.anchor_out[(long)(r0 + m) * g.ncol_tot + (long)e * g.ncol_e + c0 + cl] =
                    O.acc32 ? (double)(float)sum : sum;
        } else {
            tile_epilogue(g, O, nrep, r0 + m, e, ct, c0 + cl, O.acc32 ? (double)(float)sum : sum);
        }
    }
    if (!reanchor && O.bump && ct == 0 && e == 0 && threadIdx.x < RB &&
        r0 + (int)threadIdx.x < nrep)
        atomicAdd(reinterpret_cast<unsigned long long *>(O.stepc + r0 + threadIdx.x), 1ull);
}

// The learned proposal's scoring workgroups ride along in a transform launch: the last
// LP.ntile_blocks workgroups of grid row (y, z) = (0, 0) run them (independent of the transform: they read the committed configuration and write
// buffers the transform does not touch); those of the other rows exit.
__device__ __forceinline__ bool lp_riders(const chm_lp_dev::LPArgs &LP) {
    const int nlp = LP.ntile_blocks;
    const int x0 = (int)gridDim.x - nlp;
    if (nlp == 0 || (int)blockIdx.x < x0) return false;
    if (blockIdx.y == 0 && blockIdx.z == 0) {
        extern __shared__ __align__(16) char lp_lds[];
        chm_lp_dev::lp_block(LP, (int)blockIdx.x - x0, lp_lds);
    }
    return true;
}

__global__ __launch_bounds__(256) void k_gemm_abs_f64(Geo g, int nrep, const int *__restrict__ counts,
                                                       const int *__restrict__ dcounts,
                                                       const double *__restrict__ basisR, TileOut O,
                                                       chm_lp_dev::LPArgs LP) {
    if (lp_riders(LP)) return;
    gemm_abs_body(g, nrep, counts, dcounts, basisR, O, blockIdx.x, blockIdx.y, blockIdx.z);
}

// FP64 re-anchor of the running chi(R): anchor_out[s] = counts[s] @ basis_R for every slot
// with anchor[s] >= anchor_min.  Grid (column tiles, slots, edges), 16 columns x 16 K
// slices per workgroup; a workgroup of a slot that is not due exits after one scalar
// load.  Each lane requests its K slice's counts and basis entries together (bounds-
// checked buffer loads, no branch), then multiplies in row order; the slices are added in
// slice order.  The same slices, order and operations as the FP64 transform, so the
// values are those of gemm_abs_body's re-anchor mode bit for bit.
__global__ __launch_bounds__(256) void k_reanchor_f64(Geo g, int nrep, const int *__restrict__ counts,
                                                      const double *__restrict__ basisR,
                                                      const int *__restrict__ anchor, int anchor_min,
                                                      double *__restrict__ anchor_out, int acc32) {
    __shared__ double P[KS * NCW];
    const int slot = blockIdx.y, ct = blockIdx.x, e = blockIdx.z;
    if (anchor[slot] < anchor_min) return;
    const int K = g.nrow_e;
    const int tn = threadIdx.x % NCW, ks = threadIdx.x / NCW;
    const int c0 = ct * NCW, col = c0 + tn;
    const int kc = (K + KS - 1) / KS;
    const int kb = ks * kc, ke = min(K, kb + kc);
    const bool live = col < g.ncol_e;
    const Buf bc = make_buf(counts + (size_t)slot * g.nrow_tot + (size_t)e * g.rstride, (uint32_t)(sizeof(int) * K));
    const Buf bb = make_buf(basisR + (size_t)e * K * g.ncol_e, (uint32_t)(sizeof(double) * (size_t)K * g.ncol_e));
    constexpr int kBatch = 32;
    double acc = 0.0;
    for (int k0 = kb; k0 < ke; k0 += kBatch) {
        int cv[kBatch];
        double bv[kBatch];
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const int k = k0 + j;
            const bool ok = live && k < ke;
            cv[j] = buf_i32(bc, ok, (uint32_t)(sizeof(int) * k));
            bv[j] = buf_f64(bb, ok, (uint32_t)(sizeof(double) * ((size_t)k * g.ncol_e + col)));
        }
#pragma unroll
        for (int j = 0; j < kBatch; ++j) {
            const double f = fma((double)cv[j], bv[j], acc);
            acc = (k0 + j < ke) ? f : acc;
        }
    }
    P[ks * NCW + tn] = acc;
    __syncthreads();
    if (threadIdx.x < NCW && c0 + (int)threadIdx.x < g.ncol_e) {
        double sum = 0.0;
        for (int z = 0; z < KS; ++z) sum += P[z * NCW + threadIdx.x];
        anchor_out[(size_t)slot * g.ncol_tot + (size_t)e * g.ncol_e + c0 + threadIdx.x] =
            acc32 ? (double)(float)sum : sum;
    }
}

__device__ __forceinline__ __bf16 bf16_of_small_int(int v) {
    // |v| <= 128: exactly representable; the float's top 16 bits are the bf16
    float f = (float)v;
    return __builtin_bit_cast(__bf16, (unsigned short)(__builtin_bit_cast(unsigned int, f) >> 16));
}

// Incremental split-BF16: delta(R) = delta_counts @ basis_R with basis_R = hi + mid + lo
// (three bf16 planes) and the small integer deltas exact in bf16 (|d| > 128 split into
// 256*hi + lo), on v_mfma_f32_16x16x32_bf16 with FP32 accumulation; then
// m' = m_cur (FP64 running accumulator) + delta.  Workgroup = 4 waves splitting the K
// blocks of one 32-slot x 16-column tile; their partial tiles are summed in LDS.
// Operand layouts of the 16x16x32 MFMA: lane l holds A[l%16][8(l/16)+i] and
// B[8(l/16)+i][l%16] (i < 8); D lane l holds D[4(l/16)+i][l%16] (i < 4).
constexpr int KBW = 4;   // K blocks per wave held in registers (nkb <= 4 * KBW)

// The random numbers the next per-replica launch needs, one slot per lane, made while
// the transform runs: for one step n (pre = n - 1 before the advance) the acceptance
// uniform of step n and the proposal draws of step n + 1; for a paired round (pre = n)
// the uniforms of steps n + 1, n + 2 and the draws of steps n + 3, n + 4.  The same
// counter-based streams as in the per-replica kernel (kernels.py:378-398,
// method_a_rmc.py:278-285).  This workgroup also advances the step counter.
constexpr int kDrawInts = 24;   // u (2 x double), draws (2 x (a, c[8])), pad
__device__ __forceinline__ void compute_draws(const TileOut &O, int nrep, int r0) {
    const int t = threadIdx.x, slot = r0 + t;
    if (t >= RB || slot >= nrep) return;
    const uint64_t pre = O.stepc[slot];
    const uint64_t rung = (uint64_t)O.rung_of[slot];
    int *D = O.draws + (long)slot * kDrawInts;
    const int nst = O.draw_mode == 2 ? 2 : 1;
    for (int k = 0; k < nst; ++k) {
        const uint64_t h = splitmix64(O.seed ^ ((pre + 1 + k) * kGolden) ^ (rung << 32) ^ kAccSalt);
        const double u = u01(splitmix64(h));
        const unsigned long long b = __builtin_bit_cast(unsigned long long, u);
        D[2 * k] = (int)(unsigned)b;
        D[2 * k + 1] = (int)(unsigned)(b >> 32);
    }
    for (int k = 0; k < nst; ++k) {
        const PropDraws d = propose_draws(O.natoms, O.seed, pre + 1 + nst + k, rung);
        D[4 + 9 * k] = d.a;
#pragma unroll
        for (int i = 0; i < 8; ++i) D[5 + 9 * k + i] = d.c[i];
    }
    O.stepc[slot] = pre + (uint64_t)O.bump;
}
__device__ __forceinline__ double draw_u(const int *D, int k) {
    return __builtin_bit_cast(double, (unsigned long long)(unsigned)D[2 * k] |
                                          ((unsigned long long)(unsigned)D[2 * k + 1] << 32));
}
__device__ __forceinline__ PropDraws draw_p(const int *D, int k) {
    PropDraws d;
    d.a = D[4 + 9 * k];
#pragma unroll
    for (int i = 0; i < 8; ++i) d.c[i] = D[5 + 9 * k + i];
    return d;
}

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// two integers in [-256, 256] as a packed bf16 pair (exact: the float's upper half)
__device__ __forceinline__ unsigned pk_bf16_small(int v0, int v1) {
    return __builtin_amdgcn_perm(__builtin_bit_cast(unsigned, (float)v1), __builtin_bit_cast(unsigned, (float)v0),
                                 0x07060302u);
}

// The A operand of one K block: 8 small integer counts per lane as bf16 lo (and hi when
// |v| > 128), into acc (and acch).  When every value of the wave lies in [-128, 127]
// (nearly always: a swap changes few counts by small amounts) lo is the value itself,
// packed with one conversion and one byte permute per pair, and hi is zero.
__device__ __forceinline__ void mfma_kblock(const int4 a, const bf16x8 (&B)[3], f32x4 &acc, f32x4 &acch) {
    // 8 consecutive rows, packed two per word
    const int vals[8] = {dw_lo(a.x), dw_hi(a.x), dw_lo(a.y), dw_hi(a.y),
                         dw_lo(a.z), dw_hi(a.z), dw_lo(a.w), dw_hi(a.w)};
    int wide = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) wide |= (vals[i] + 128) & ~255;
    if (!__any(wide)) {
        const u32x4 w = {pk_bf16_small(vals[0], vals[1]), pk_bf16_small(vals[2], vals[3]),
                         pk_bf16_small(vals[4], vals[5]), pk_bf16_small(vals[6], vals[7])};
        const bf16x8 alo = __builtin_bit_cast(bf16x8, w);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, B[0], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, B[1], acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, B[2], acc, 0, 0, 0);
        return;
    }
    bf16x8 alo, ahi;
    bool hi_any = false;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int v = vals[i];
        const int lo = ((v + 128) & 255) - 128;
        const int hi = (v - lo) >> 8;
        hi_any |= (hi != 0);
        alo[i] = bf16_of_small_int(lo);
        ahi[i] = bf16_of_small_int(hi);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, B[0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, B[1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, B[2], acc, 0, 0, 0);
    if (__any(hi_any)) {
        acch = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, B[0], acch, 0, 0, 0);
        acch = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, B[1], acch, 0, 0, 0);
        acch = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, B[2], acch, 0, 0, 0);
    }
}

// NA = 1: m' = m_cur + dc[0] @ B into O[0].  NA = 3 (a paired round): m1 = m_cur +
// dc[0] @ B, m2r = m_cur + dc[1] @ B and m2a = m1 + dc[2] @ B into O[0..2], each product
// accumulated in FP32 on its own and added to the FP64 value, exactly as in single steps.
template <int NA, int KB>
__device__ __forceinline__ void gemm_delta_bf16_tile(const Geo &g, int nrep, const int *const (&dc)[NA],
                                                     const bf16x8 *__restrict__ Tb,
                                                     const double *__restrict__ mcur, const TileOut *O, int ct,
                                                     int rbi, int e) {
    __shared__ float red[NA][4][RB][NCW + 1];
    const int r0 = rbi * RB;
    const int t = threadIdx.x;
    const int wave = t >> 6, lane = t & 63;
    snap_best(O[0], nrep, r0, ct, e);
    // epilogue operands first, so their latency hides behind the MFMA phase
    double mc_pre[2], tg_pre[2], w_pre[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = t / NCW + 16 * h, col = ct * NCW + t % NCW, slot = r0 + m;
        const bool v = slot < nrep && col < g.ncol_e;
        const long gc = (long)e * g.ncol_e + col;
        mc_pre[h] = v ? mcur[(long)slot * g.ncol_tot + gc] : 0.0;
        tg_pre[h] = v ? O[0].target[gc] : 0.0;
        w_pre[h] = v ? O[0].wcol[gc] : 0.0;
    }
    const int per = (g.nkb + 3) / 4;
    const int kb0 = wave * per, kb1 = min(g.nkb, kb0 + per);
    // all operands of this wave's K blocks, loaded before any MFMA
    bf16x8 B[KB][3];
    int4 Av[NA][KB][2];   // 8 consecutive rows of one slot, two per word
    Buf bA[NA];
#pragma unroll
    for (int a = 0; a < NA; ++a) bA[a] = make_buf(dc[a], 4u * nrep * g.nrow_w);
#pragma unroll
    for (int q = 0; q < KB; ++q) {
        const int kb = kb0 + q;
        if (kb < kb1) {
            const bf16x8 *Bp = Tb + ((((long)e * g.nct + ct) * g.nkb + kb) * 3) * 64 + lane;
            B[q][0] = Bp[0];
            B[q][1] = Bp[64];
            B[q][2] = Bp[128];
            const int k0 = kb * 32 + 8 * (lane >> 4);
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int rb = 0; rb < 2; ++rb) {
                    const int slot = r0 + rb * 16 + (lane & 15);
                    // bounds-checked (slots beyond nrep read zero): no branch, no wait
                    Av[a][q][rb] = buf_i128(bA[a], slot < nrep,
                                            4u * ((uint32_t)slot * g.nrow_w + ((e * g.rstride + k0) >> 1)));
                }
        }
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        f32x4 acc[2], acch[2];
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i) acc[rb][i] = acch[rb][i] = 0.f;
#pragma unroll
        for (int q = 0; q < KB; ++q) {
            if (kb0 + q >= kb1) break;
#pragma unroll
            for (int rb = 0; rb < 2; ++rb) mfma_kblock(Av[a][q][rb], B[q], acc[rb], acch[rb]);
        }
#pragma unroll
        for (int rb = 0; rb < 2; ++rb)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                red[a][wave][rb * 16 + 4 * (lane >> 4) + i][lane & 15] = acc[rb][i] + 256.f * acch[rb][i];
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = t / NCW + 16 * h, cl = t % NCW;
        const int slot = r0 + m, col = ct * NCW + cl;
        const bool valid = slot < nrep && col < g.ncol_e;
        double m1 = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            const float d = red[a][0][m][cl] + red[a][1][m][cl] + red[a][2][m][cl] + red[a][3][m][cl];
            const TileOut &Oa = O[a];
            double mp = valid ? ((a == 2) ? m1 : mc_pre[h]) + (double)d : 0.0;
            if (Oa.acc32) mp = (double)(float)mp;
            if (a == 0) m1 = mp;
            double p1 = 0.0, p2 = 0.0, q1 = 0.0, q2 = 0.0;
            if (valid) {
                Oa.mtrial[(long)slot * g.ncol_tot + (long)e * g.ncol_e + col] = mp;
                const double dd = mp - tg_pre[h];
                p1 = dd * tg_pre[h];
                p2 = dd * dd;
                q1 = p1 * w_pre[h];
                q2 = p2 * w_pre[h];
            }
            p1 = group_sum_d<16>(p1);
            p2 = group_sum_d<16>(p2);
            q1 = group_sum_d<16>(q1);
            q2 = group_sum_d<16>(q2);
            if (slot < nrep && cl == 0) {
                double *P = Oa.partials + (((long)slot * g.nsp + e) * g.nct + ct) * 4;
                P[0] = p1;
                P[1] = p2;
                P[2] = q1;
                P[3] = q2;
            }
        }
    }
    if (O[0].bump && !O[0].draws && ct == 0 && e == 0 && threadIdx.x < RB && r0 + (int)threadIdx.x < nrep)
        atomicAdd(reinterpret_cast<unsigned long long *>(O[0].stepc + r0 + threadIdx.x),
                  (unsigned long long)O[0].bump);
}

template <int KB = KBW>
__device__ __forceinline__ void gemm_delta_bf16_body(const Geo &g, int nrep, const int *__restrict__ dcounts,
                                                     const bf16x8 *__restrict__ Tb,
                                                     const double *__restrict__ mcur, const TileOut &O, int ct,
                                                     int rbi, int e) {
    const int *const dc[1] = {dcounts};
    gemm_delta_bf16_tile<1, KB>(g, nrep, dc, Tb, mcur, &O, ct, rbi, e);
}

__global__ __launch_bounds__(256) void k_gemm_delta_bf16(Geo g, int nrep, const int *__restrict__ dcounts,
                                                          const bf16x8 *__restrict__ Tb,
                                                          const double *__restrict__ mcur, TileOut O,
                                                          chm_lp_dev::LPArgs LP) {
    if (lp_riders(LP)) return;
    if ((int)blockIdx.x == g.nct) {   // the extra column of workgroups: draws + step counter
        if (blockIdx.z == 0) compute_draws(O, nrep, blockIdx.y * RB);
        return;
    }
    gemm_delta_bf16_body(g, nrep, dcounts, Tb, mcur, O, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Incremental FP32: delta(R) = delta_counts @ basis_R with basis_R rounded to FP32 and FP32
// FMA (no MFMA: the reference transform in FP32, for the precision study), added to the
// running chi(R).  Tiling as k_gemm_abs_f64: 16 columns x 16 K slices, 32 slots per lane.
__device__ __forceinline__ void gemm_delta_f32_body(const Geo &g, int nrep, const int *__restrict__ dcounts,
                                                    const float *__restrict__ basis32,
                                                    const double *__restrict__ mcur, const TileOut &O, int ct,
                                                    int rbi, int e) {
    extern __shared__ float fl[];
    const int r0 = rbi * RB;
    const int c0 = ct * NCW;
    const int K = g.nrow_e;
    const int ld = RB + 1;
    snap_best(O, nrep, r0, ct, e);
    for (int q = threadIdx.x; q < RB * K; q += blockDim.x) {
        const int m = q / K, k = q % K;
        const int slot = r0 + m;
        fl[k * ld + m] = (slot < nrep) ? (float)dw_row(dcounts + (long)slot * g.nrow_w, e * g.rstride + k) : 0.f;
    }
    __syncthreads();
    const int tn = threadIdx.x % NCW, ks = threadIdx.x / NCW;
    const int col = c0 + tn;
    const int kc = (K + KS - 1) / KS;
    const int kb = ks * kc, ke = min(K, kb + kc);
    float acc[RB];
#pragma unroll
    for (int m = 0; m < RB; ++m) acc[m] = 0.f;
    if (col < g.ncol_e) {
        const float *Bc = basis32 + (long)e * K * g.ncol_e + col;
        for (int k = kb; k < ke; ++k) {
            const float bv = Bc[(long)k * g.ncol_e];
            const float *Ak = fl + k * ld;
#pragma unroll
            for (int m = 0; m < RB; ++m) acc[m] = fmaf(Ak[m], bv, acc[m]);
        }
    }
    __syncthreads();
    float *P = fl;
#pragma unroll
    for (int m = 0; m < RB; ++m) P[(ks * RB + m) * NCW + tn] = acc[m];
    __syncthreads();
    const int t = threadIdx.x;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int m = t / NCW + 16 * h, cl = t % NCW;
        float d = 0.f;
        for (int z = 0; z < KS; ++z) d += P[(z * RB + m) * NCW + cl];
        const int slot = r0 + m, c = c0 + cl;
        double mp = 0.0;
        if (slot < nrep && c < g.ncol_e) {
            mp = mcur[(long)slot * g.ncol_tot + (long)e * g.ncol_e + c] + (double)d;
            if (O.acc32) mp = (double)(float)mp;
        }
        tile_epilogue(g, O, nrep, slot, e, ct, c, mp);
    }
    if (O.bump && ct == 0 && e == 0 && threadIdx.x < RB && r0 + (int)threadIdx.x < nrep)
        atomicAdd(reinterpret_cast<unsigned long long *>(O.stepc + r0 + threadIdx.x), 1ull);
}

__global__ __launch_bounds__(256) void k_gemm_delta_f32(Geo g, int nrep, const int *__restrict__ dcounts,
                                                         const float *__restrict__ basis32,
                                                         const double *__restrict__ mcur, TileOut O,
                                                         chm_lp_dev::LPArgs LP) {
    if (lp_riders(LP)) return;
    gemm_delta_f32_body(g, nrep, dcounts, basis32, mcur, O, blockIdx.x, blockIdx.y, blockIdx.z);
}

__global__ void k_to_f32(long n, const double *__restrict__ in, float *__restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// basis_R (nsp, nrow_e, ncol_e) FP64 -> three bf16 planes in MFMA B-operand order:
// Tb[e][ct][kb][plane][lane] = 8 bf16 of column ct*16 + lane%16, rows kb*32 + 8(lane/16) + i.
__global__ void k_split_bf16(Geo g, const double *__restrict__ basisR, unsigned short *__restrict__ Tb) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.nsp * g.nct * g.nkb * 64 * 8;
    if (tid >= n) return;
    int i = (int)(tid % 8);
    long r = tid / 8;
    int lane = (int)(r % 64);
    r /= 64;
    int kb = (int)(r % g.nkb);
    r /= g.nkb;
    int ct = (int)(r % g.nct);
    int e = (int)(r / g.nct);
    int col = ct * 16 + (lane & 15);
    int k = kb * 32 + 8 * (lane >> 4) + i;
    double v = (col < g.ncol_e && k < g.nrow_e) ? basisR[((long)e * g.nrow_e + k) * g.ncol_e + col] : 0.0;
    unsigned short h0 = bf16_rne((float)v);
    double r1 = v - (double)bf16_to_float(h0);
    unsigned short h1 = bf16_rne((float)r1);
    double r2 = r1 - (double)bf16_to_float(h1);
    unsigned short h2 = bf16_rne((float)r2);
    long base = ((((long)e * g.nct + ct) * g.nkb + kb) * 3) * 64 * 8 + (long)lane * 8 + i;
    Tb[base] = h0;
    Tb[base + 64 * 8] = h1;
    Tb[base + 2 * 64 * 8] = h2;
}

// ---------------------------------------------------------------- K3+K1 per slot
// One workgroup (1024 lanes) per replica slot.  Accept part (step t): the chi^2 of the
// trial from the GEMM's per-tile sums, Metropolis with a counter-based uniform
// (method_a_rmc.py:278-285), commit of counts / species / running chi(R), best
// retention (:301-307), move log, FP64 re-anchor.  Propose part (step t+1):
// splitmix64 proposal (kernels.py:362-398) and the gather of every directed pair and
// multiple-scattering path through the two swapped sites (kernels.py:425-468,
// chimera/mslibrary.py:466-486) into a dense per-slot row of count changes.  Neighbour
// and path sites come from integer lattice arithmetic on the site index (no dependent
// table loads); the histogram lives in LDS and is stored with plain coalesced writes.
struct SlotArgs {
    const double *partials; // (nrep, nsp, nct, 4)
    const double *q0;       // (nsp, 2): per edge T0 = sum t^2 and Q0 = sum t^2 w
    const double *mtrial;   // (nrep, ncol_tot)
    const double *target, *wcol;
    const double *basisR;   // FP64, for re-anchoring
    double *mcur;           // (nrep, ncol_tot) running accumulator (incremental mode)
    const double *beta_slot;// beta of each slot's current rung (kept by the exchange)
    const int *prop_acc;    // (nrep, kPropInts) a, b, new a, new b, is-swap, log q ratio
    int *prop_next;         // (nrep, kPropInts) the proposal made here (other parity buffer)
    const double *c2_prev;  // chi^2 after the previous step (read by every workgroup)
    double *c2_next;        // chi^2 after this step (written by the lead workgroup)
    int *counts;            // (nrep, nrow_tot)
    int *dc_acc;            // count changes of the step being accepted (zeroed after)
    int *dc_prop;           // count changes of the step being proposed (zero on entry)
    uint8_t *species, *best_sp;
    double *best_c2;
    const double *best_snap;   // best_c2 as the trial transform before this launch copied it
    const int *rung_of;
    const uint64_t *stepc;  // the pending step (advanced by the transform kernels)
    long long *acc_rung;
    int *anchor;            // accepted moves since the last FP64 re-anchor, per slot
    long long *nanchor;     // re-anchors done, per slot
    int reanchor_step;      // the FP64 re-anchor kernel ran before this step's transform
    const int *ss_sten;     // (ncoord) packed dx,dy,dz,shell (int8 each)
    const int4 *ms_sten;    // (nstencil) 16-byte path records
    int reanchor_every;     // 0: never
    int incremental;
    int *log;               // (log_cap, nrep, 4) or null
    int log_cap;
    int *series;            // (series_cap, nsp*nsp*nshell) or null
    int series_cap, series_rung;
    long long rec_base;     // record index = step - 1 - rec_base
    uint8_t *pend_occ;      // (nrep, noccp) occupancy of each slot's pending proposal, or null
    uint8_t *rec_occ;       // (rec_cap, noccp) move records (training data), or null
    float *rec_y;           //   energy change nfit * (chi2' - chi2) of the proposal
    int *rec_meta;          //   rung | accepted << 16, or -1 for a null move
    int rec_cap, nocc, noccp;
    uint64_t seed;
    int profile;
    int nrep;               // replica slots of the problem
    int accept;             // 0: none, 1: step, 2: init (store chi^2, no decision)
    int propose;            // 1: propose + gather the next step; 4: the same with the learned
                            //    proposal (the scoring kernel's draw for this step's outcome)
    double nfit;
    const int *draws;       // (nrep, kDrawInts) made by the transform just before, or null
    // learned proposal (lp_lq null: off).  The scoring workgroups of this step's transform
    // left per-tile partial sums of its log q ratio and, for the next step's candidates,
    // g / dU at the committed configuration and with this step's swap made, and the tile
    // maxima of their logits for either outcome.
    const double *lp_lq;    // (nrep, MAXT, 6)
    const float *lp_dmax;   // (nrep, MAXT, 2)
    const float *lp_sg;     // (nrep, 2, ncand)
    const int *lp_sdu;
    const int8_t *lp_st;    // (nrep, ncand)
    const int *lp_dir_rd;   // lifted: direction before this step (null: not lifted)
    int *lp_dir_wr;         //   after it (reversed when a learned step is not accepted)
    int *lp_cache;          // (nrep, 2, natoms, HID) per-site layer-1 sums: an accepted swap is
                            //   folded in here by the slot's folding workgroups
    const int *lp_w1q;      //   with the scorer's layer-1 weight rows,
    const int *lp_nbr;      //   the stencil-ordered neighbour table
    const signed char *lp_posmap;   //   and the offset -> stencil position cube (side 2 pm_r + 1)
    int lp_pm_r;
    int *lp_cs;             // this step's outcome for the next scoring workgroups
    uint64_t *lp_lpn;       // the step the next transform's scoring workgroups serve
    int lp_ncand, lp_lift;
    double lp_eps, lp_beta_max, lp_lam, lp_kappa;
};

__device__ double block_sum(double v, double *red) {
    __syncthreads();
    red[threadIdx.x] = v;
    __syncthreads();
    for (int h = blockDim.x / 2; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

// chi^2 from the per-edge residual sums {U1, U2, W1, W2}; q0v: per edge T0, Q0
__device__ __forceinline__ double chi2_of_sums(const Geo &g, const double (*S)[4],
                                               const double (&q0v)[2 * chm::kMaxNsp], int profile,
                                               double nfit) {
    double tot = 0.0;
#pragma unroll
    for (int e = 0; e < chm::kMaxNsp; ++e) {
        if (e >= g.nsp) break;
        const double U1 = S[e][0], U2 = S[e][1], W1 = S[e][2], W2 = S[e][3];
        const double T0 = q0v[2 * e], Q0 = q0v[2 * e + 1];
        double s = 1.0, sm1 = 0.0;
        if (profile) {
            const double P2 = fmax(U2 + 2.0 * U1 + T0, 1e-30);
            const double P1 = U1 + T0;
            s = P1 / P2;
            sm1 = (P1 - P2) / P2;
        }
        tot += s * s * W2 + 2.0 * s * sm1 * W1 + sm1 * sm1 * Q0;
    }
    return tot / nfit;
}

// One replica slot s, workgroup `split` of `nsplit` (the body of k_slot_step and of the
// batched launch over many problems).  Every input that does not depend on the decision
// is requested before the first barrier (one memory round trip); the decision is taken
// by every workgroup, and each commits its own share of the counts, running chi(R) and
// best configuration.
// split < nsplit: a workgroup of the step's commit and gather shares; nsplit <= split <
// nsplit + nfold: a folding workgroup of the learned proposal (decides, folds, exits)
__device__ __forceinline__ void slot_step_body(const Geo &g, const SlotArgs &A, int s, int split, int nsplit,
                                               int nfold = 0) {
    extern __shared__ int lds[];
    int *hist = lds;
    uint8_t *lsp = reinterpret_cast<uint8_t *>(lds + ((g.nrow_tot + 3) & ~3));
    __shared__ double sums[chm::kMaxNsp][4];
    __shared__ int flag;
    __shared__ Prop P;
    const int nwg = nsplit;
    const bool lead = split == 0;            // the workgroup that owns the slot's scalars
    const int t = threadIdx.x;
    const int nt = blockDim.x;
    if (t == 0) flag = 0;
    const long base_sp = (long)s * g.natoms_pad;
    // ------------------------------------------------ every input, in one round trip:
    // all loads are issued (fixed counts, unconditional where the buffer allows) before
    // any of them is consumed
    const uint64_t stepc = A.stepc[s];       // written only by the transform kernels
    const int rung = A.rung_of[s];
    const double beta = A.beta_slot[s];
    const double c2p = A.c2_prev[s];
    const double best = A.best_snap[s];   // the best before this launch (see snap_best)
    const int anchor = A.anchor[s];
    int pp[kPropInts];                       // the proposal being accepted
#pragma unroll
    for (int q = 0; q < kPropInts; ++q) pp[q] = A.prop_acc[kPropInts * s + q];
    // learned proposal: the next step's draw is made for either outcome of this step before
    // the decision, by wave 2 (this step not accepted) and wave 3 (accepted), four candidates
    // per lane: g / dU at the committed configuration and with this step's swap made, and
    // the tiles' maxima of the outcome's logits.  Lanes 0..15 of wave 3 take the tiles'
    // partial sums of this step's log q ratio.
    const int wave = t >> 6, lane = t & 63;
    const int lp_nt = A.lp_ncand / chm_lp_dev::MR;
    const bool drawer = A.lp_lq && A.propose == 4 && wave >= 2;
    float4 dgA = {}, dgB = {};
    int4 ddA = make_int4(INT_MIN, INT_MIN, INT_MIN, INT_MIN), ddB = ddA;
    int dtB = 0, lpd0 = 1;
    float dm[chm_lp_dev::MAXT];
    double lqm = -__builtin_huge_val(), lqz = 0.0, lqd = 0.0;
    if (A.lp_lq) {
        const long cb = (long)s * A.lp_ncand;
        if (drawer && 4 * lane < A.lp_ncand) {
            dgA = reinterpret_cast<const float4 *>(A.lp_sg + 2 * cb)[lane];
            ddA = reinterpret_cast<const int4 *>(A.lp_sdu + 2 * cb)[lane];
            if (wave == 3) {
                dgB = reinterpret_cast<const float4 *>(A.lp_sg + 2 * cb + A.lp_ncand)[lane];
                ddB = reinterpret_cast<const int4 *>(A.lp_sdu + 2 * cb + A.lp_ncand)[lane];
                dtB = reinterpret_cast<const int *>(A.lp_st + cb)[lane];
            }
        }
#pragma unroll
        for (int k = 0; k < chm_lp_dev::MAXT; ++k)
            dm[k] = drawer ? A.lp_dmax[((long)s * chm_lp_dev::MAXT + k) * 2 + (wave == 3 ? 0 : 1)] : 0.f;
        if (wave == 3 && lane < 16 && (lane & 7) < lp_nt) {
            const double *q = A.lp_lq + ((long)s * chm_lp_dev::MAXT + (lane & 7)) * chm_lp_dev::kLqDoubles +
                              3 * (lane >> 3);
            lqm = q[0];
            lqz = q[1];
            lqd = q[2];
        }
        if (A.lp_dir_rd) lpd0 = A.lp_dir_rd[s];
    }
    __shared__ int lp_dnew, lp_jo[2];
    __shared__ double lp_lqv;
    // ... and, to fold an accepted swap into the per-site sums, the offset cube
    __shared__ __align__(16) signed char lp_pm[chm_lp_dev::kPosMapMax];
    const bool lp_fold = A.lp_cache && A.accept == 1;
    const int lp_npm16 = lp_fold ? ((2 * A.lp_pm_r + 1) * (2 * A.lp_pm_r + 1) * (2 * A.lp_pm_r + 1) + 15) / 16 : 0;
    const uint4 lp_pmv = (t < lp_npm16) ? reinterpret_cast<const uint4 *>(A.lp_posmap)[t] : make_uint4(0, 0, 0, 0);
    double q0v[2 * chm::kMaxNsp];            // per edge T0, Q0 (padded with zeros)
#pragma unroll
    for (int q = 0; q < 2 * chm::kMaxNsp; ++q) q0v[q] = A.q0[q];
    // species: one 16-byte vector per lane
    const int nv = g.natoms_pad / 16;
    const uint4 *gsp = reinterpret_cast<const uint4 *>(A.species + base_sp);
    const int4 spv = buf_i128(make_buf(gsp, 16u * nv), t < nv, 16u * t);
    // per-edge sums over the column tiles: (e, j) groups of CT lanes (CT-aligned inside a
    // wave), each lane summing tiles c0, c0 + CT, ... in order, then DPP sums over the
    // group's lanes
    const int G = 4 * g.nsp;
    const int CT = (G <= 16) ? 16 : 8;
    const int gi = t / CT, c0 = t - gi * CT;
    const bool sum_lane = A.accept && t < G * CT;
    const double *Pp = A.partials + (long)s * g.nsp * g.nct * 4;
    const Buf bp = make_buf(Pp, 32u * g.nsp * g.nct);
    const uint32_t poff = 8u * ((gi >> 2) * g.nct * 4 + (gi & 3));
    constexpr int PR = 8;
    double pv[PR];
#pragma unroll
    for (int r = 0; r < PR; ++r) pv[r] = buf_f64(bp, sum_lane && c0 + r * CT < g.nct, poff + 32u * (c0 + r * CT));
    // this workgroup's share of the running chi(R) to copy forward on acceptance
    constexpr int MC = 4;
    double mt[MC];
    const int mchunk = (g.ncol_tot + nwg - 1) / nwg;
    const int mbeg = split * mchunk, mend = min(g.ncol_tot, mbeg + mchunk);
    const bool mcopy = A.accept && A.incremental;
    const double *mts = A.mtrial + (long)s * g.ncol_tot;
    const Buf bmt = make_buf(mts, 8u * g.ncol_tot);
#pragma unroll
    for (int r = 0; r < MC; ++r) {
        const int q = mbeg + t + r * nt;
        mt[r] = buf_f64(bmt, mcopy && q < mend, 8u * q);
    }
    // ... and of the count commit.  With the pair-count series of this slot's rung being
    // recorded, the lead workgroup commits every row (it reads the committed counts).
    // The chunks are in words of the packed count-change row (rows 2w, 2w+1).
    int *cn = A.counts + (long)s * g.nrow_tot;
    int *dca = A.dc_acc + (long)s * g.nrow_w;
    const long long ri = (long long)stepc - 1 - A.rec_base;
    const bool series_on = A.accept == 1 && A.series && rung == A.series_rung && ri >= 0 &&
                           ri < A.series_cap;
    // (the prefetched chunk does not depend on series_on, which needs the step counter)
    const int wchunk = (g.nrow_w + nwg - 1) / nwg;
    const int pbeg = split * wchunk, pend = min(g.nrow_w, pbeg + wchunk);
    const int rbeg = series_on ? 0 : pbeg, rend = series_on ? (lead ? g.nrow_w : 0) : pend;
    const Buf bcn = make_buf(cn, 4u * g.nrow_tot), bdc = make_buf(dca, 4u * g.nrow_w);
    constexpr int RC = 2;
    int dcv[RC];
    int2 cnv[RC];
#pragma unroll
    for (int r = 0; r < RC; ++r) {
        const int w = pbeg + t + r * nt;
        dcv[r] = buf_i32(bdc, A.accept && w < pend, 4u * w);
        cnv[r] = buf_i64(bcn, A.accept == 1 && w < pend, 8u * w);
    }
    // this lane's first gather records (independent of the proposal)
    const int stride = nwg * nt;
    const int first = split * nt + t;
    RecPre recs;
    prefetch_recs(g, A.ss_sten, A.ms_sten, first, stride, A.propose, recs);
    // the draws of this step's acceptance uniform and of the next uniform proposal
    // (wave-uniform: scalar arithmetic)
    // (made by the transform when it ran just before; else here, by wave 0 only: lane 0
    // takes the decision and makes the pick)
    double u_acc = 1.0;
    PropDraws dr{};
    if (A.draws) {
        const int *D = A.draws + (long)s * kDrawInts;
        if (A.accept == 1) u_acc = draw_u(D, 0);
        if (A.propose) dr = draw_p(D, 0);
    } else if (t < 64) {
        if (A.accept == 1) {
            const uint64_t h = splitmix64(A.seed ^ (stepc * kGolden) ^ ((uint64_t)rung << 32) ^ kAccSalt);
            u_acc = u01(splitmix64(h));
        }
        if (A.propose) dr = propose_draws(g.natoms, A.seed, stepc + 1, (uint64_t)rung);
    }
    __builtin_amdgcn_sched_barrier(0);   // every load issued before the first use
    // ------------------------------------------------ consume
    if (t < nv) reinterpret_cast<int4 *>(lsp)[t] = spv;
    for (int q = t + nt; q < nv; q += nt) reinterpret_cast<uint4 *>(lsp)[q] = gsp[q];
    if (A.accept) {
        double psum = 0.0;   // out-of-range tiles loaded as 0.0
#pragma unroll
        for (int r = 0; r < PR; ++r) psum += pv[r];
        for (int c = c0 + PR * CT; sum_lane && c < g.nct; c += CT)
            psum += Pp[(long)((gi >> 2) * g.nct + c) * 4 + (gi & 3)];
        psum = (CT == 16) ? group_sum_d<16>(psum) : group_sum_d<8>(psum);
        if (t < G * CT && c0 == 0) sums[gi >> 2][gi & 3] = psum;
    }
    if (A.propose)
        for (int q = t; q < g.nrow_tot; q += nt) hist[q] = 0;
    if (A.lp_lq && wave == 3) {
        const double v = chm_lp_dev::lq_merge(lqm, lqz, lqd);
        if (lane == 0) lp_lqv = v;
    }
    if (t < lp_npm16) reinterpret_cast<uint4 *>(lp_pm)[t] = lp_pmv;
    __syncthreads();
    // the learned draws of the next step (waves 2 and 3, beside the decision)
    if (drawer) {
        const uint64_t hn = lp_stream(A.seed, stepc + 1, rung);
        if (!lp_coin_uniform(hn, beta, A.lp_eps, A.lp_beta_max)) {
            const bool oacc = wave == 3;
            // the direction after this step: kept if it is accepted, reversed if a learned
            // step is not (a uniform step keeps it)
            const bool learned_n = A.accept == 1 && !lp_coin_uniform(lp_stream(A.seed, stepc, rung), beta,
                                                                     A.lp_eps, A.lp_beta_max);
            const int dout = (oacc || !learned_n) ? lpd0 : -lpd0;
            float m = -__builtin_huge_valf();
#pragma unroll
            for (int k = 0; k < chm_lp_dev::MAXT; ++k)
                if (k < lp_nt) m = fmaxf(m, dm[k]);
            const float ga[4] = {dgA.x, dgA.y, dgA.z, dgA.w}, gb[4] = {dgB.x, dgB.y, dgB.z, dgB.w};
            const int da[4] = {ddA.x, ddA.y, ddA.z, ddA.w}, db[4] = {ddB.x, ddB.y, ddB.z, ddB.w};
            float lg[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const bool useB = oacc && ((dtB >> (8 * q)) & 0xFF);
                lg[q] = (4 * lane + q < A.lp_ncand)
                            ? chm_lp_dev::logit_of(A.lp_lam, A.lp_kappa, A.lp_lift, useB ? gb[q] : ga[q],
                                                   useB ? db[q] : da[q], beta, dout)
                            : -__builtin_huge_valf();
            }
            const int j = chm_lp_dev::wave_draw4(lg, m, u01(splitmix64(hn ^ kSelSalt)));
            if (lane == 0) lp_jo[oacc] = j;
        }
    }
    if (A.accept) {
        // ------------------------------------------------ decision (same in every workgroup)
        if (t == 0) {
            const double c2t = chi2_of_sums(g, sums, q0v, A.profile, A.nfit);
            int acc = 1;
            double cur = c2t;
            if (A.accept == 1) {
                acc = 0;
                // a learned step's log q(x'->x) - log q(x->x') from the scoring tiles' partial
                // sums; a uniform step's (0) in pp[5]
                const bool learned = A.lp_lq && !lp_coin_uniform(lp_stream(A.seed, stepc, rung), beta,
                                                                 A.lp_eps, A.lp_beta_max);
                if (pp[4]) {
                    double lq = (double)__int_as_float(pp[5]);
                    if (learned) lq = lp_lqv;
                    const double x = beta * (c2t - c2p) * A.nfit - lq;
                    acc = metropolis(x, u_acc);
                }
                // a lifted learned step that is not accepted reverses its direction (a
                // uniform step, on the coin or a cold rung, keeps it)
                lp_dnew = (learned && !acc) ? -lpd0 : lpd0;
                if (acc) {
                    lsp[pp[0]] = (uint8_t)pp[2];
                    lsp[pp[1]] = (uint8_t)pp[3];
                }
                cur = acc ? c2t : c2p;
            }
            flag = acc | ((A.accept == 2 || cur < best) ? 2 : 0);
            if (lead) {
                A.c2_next[s] = cur;
                if (A.accept == 2 || cur < best) A.best_c2[s] = cur;
                if (A.accept == 1) {
                    if (acc) {
                        A.species[base_sp + pp[0]] = (uint8_t)pp[2];
                        A.species[base_sp + pp[1]] = (uint8_t)pp[3];
                        atomicAdd(reinterpret_cast<unsigned long long *>(A.acc_rung + rung), 1ull);
                    }
                    if (A.lp_lq) {
                        if (A.lp_dir_wr) A.lp_dir_wr[s] = lp_dnew;
                        A.lp_cs[s] = acc;
                        A.lp_lpn[s] = stepc + 1;
                    }
                    if (A.log && ri >= 0 && ri < A.log_cap) {
                        int *L = A.log + ((long)ri * A.nrep + s) * 4;
                        L[0] = pp[0];
                        L[1] = pp[1];
                        L[2] = acc;
                        L[3] = rung;
                    }
                    // accepted moves since the running chi(R) was last recomputed in FP64
                    if (A.incremental) {
                        if (A.reanchor_step && anchor >= A.reanchor_every) {
                            A.anchor[s] = acc;
                            atomicAdd(reinterpret_cast<unsigned long long *>(A.nanchor + s), 1ull);
                        } else if (acc) {
                            A.anchor[s] = anchor + 1;
                        }
                    }
                    if (A.rec_y) {
                        // the move record of this step: energy change and outcome
                        const long idx = (long)(((stepc - 1) * (uint64_t)A.nrep + (uint64_t)s) %
                                                (uint64_t)A.rec_cap);
                        A.rec_y[idx] = pp[4] ? (float)((c2t - c2p) * A.nfit) : 0.f;
                        A.rec_meta[idx] = pp[4] ? (rung | (acc << 16)) : -1;
                    }
                }
            }
        }
        __syncthreads();
        const int acc = flag & 1;
        const bool add = acc && A.accept == 1;
        if (split >= nwg) {
            // a folding workgroup (learned proposal): an accepted swap folded into the slot's
            // per-site layer-1 sums (a takes pp[2], b takes pp[3]: the species they
            // exchange), the items shared by the slot's folding workgroups
            if (lp_fold && add && pp[4]) {
                int *cache = A.lp_cache + (long)s * 2 * g.natoms * chm_lp_dev::HID;
                for (int it = (split - nwg) * nt + t; it < chm_lp_dev::fold_nitems(g); it += nfold * nt) {
                    chm_lp_dev::FoldItem F;
                    chm_lp_dev::fold_item_load(g, A.lp_pm_r, lp_pm, A.lp_w1q, cache, pp[0], pp[1], pp[2], pp[3],
                                               pp[3], pp[2], it,
                                               chm_lp_dev::fold_item_site(g, A.lp_nbr, pp[0], pp[1], it), F);
                    chm_lp_dev::fold_item_store(F);
                }
            }
            return;
        }
        // ------------------------------------------------ commit of this workgroup's share
        // words rbeg + t + r nt: prefetched when inside this workgroup's chunk (rbeg == pbeg
        // whenever rend > 0: series_on commits all rows in the lead, whose chunk starts at 0)
        int2 *cn2 = reinterpret_cast<int2 *>(cn);
        auto commit_word = [&](int w, int d, int2 c) {
            if (d != 0) {
                if (add) cn2[w] = make_int2(c.x + dw_lo(d), c.y + dw_hi(d));
                dca[w] = 0;
            }
        };
#pragma unroll
        for (int r = 0; r < RC; ++r) {
            const int w = rbeg + t + r * nt;
            if (w < rend) {
                if (w < pend)
                    commit_word(w, dcv[r], cnv[r]);
                else
                    commit_word(w, dca[w], cn2[w]);
            }
        }
        for (int w = rbeg + t + RC * nt; w < rend; w += nt) commit_word(w, dca[w], cn2[w]);
        if (series_on && lead) {
            __syncthreads();
            const int nss = g.nsp * g.nsp * g.nshell;
            for (int q = t; q < nss; q += nt) {
                const int a = q / (g.nsp * g.nshell), r = q % (g.nsp * g.nshell);
                A.series[(long)ri * nss + q] = cn[a * g.rstride + r];
            }
        }
        if (A.incremental && acc) {
            double *mc = A.mcur + (long)s * g.ncol_tot;
#pragma unroll
            for (int r = 0; r < MC; ++r) {
                const int q = mbeg + t + r * nt;
                if (q < mend) mc[q] = mt[r];
            }
            for (int q = mbeg + t + MC * nt; q < mend; q += nt) mc[q] = A.mtrial[(long)s * g.ncol_tot + q];
        }
        if (flag & 2) {
            const int nv = g.natoms_pad / 16;
            const int vchunk = (nv + nwg - 1) / nwg;
            const int vend = min(nv, (split + 1) * vchunk);
            const uint4 *src = reinterpret_cast<const uint4 *>(lsp);
            uint4 *dst = reinterpret_cast<uint4 *>(A.best_sp + base_sp);
            for (int q = split * vchunk + t; q < vend; q += nt) dst[q] = src[q];
        }
        if (lead && A.rec_y && A.accept == 1 && pp[4]) {
            // ... and the occupancy of its proposal
            const long idx = (long)(((stepc - 1) * (uint64_t)A.nrep + (uint64_t)s) % (uint64_t)A.rec_cap);
            for (int q = t; q < A.noccp; q += nt) A.rec_occ[idx * A.noccp + q] = A.pend_occ[(long)s * A.noccp + q];
        }
    }
    if (!A.propose) return;
    // ------------------------------------------------ propose + gather of the next step
    // the learned proposal: the draw made above for this step's outcome (every workgroup
    // draws the same one)
    const uint64_t hn = lp_stream(A.seed, stepc + 1, rung);
    const bool learned_next = A.propose == 4 && !lp_coin_uniform(hn, beta, A.lp_eps, A.lp_beta_max);
    if (!A.accept && learned_next) __syncthreads();   // (no decision barrier since the draws)
    const int jl = learned_next ? lp_jo[(A.accept == 1 && (flag & 1)) ? 1 : 0] : -1;
    if (t == 0) {
        if (learned_next) {
            if (jl >= 0) {
                int a, b;
                lp_candidate(hn, jl, g.natoms, a, b);
                P = Prop{a, b, lsp[b], lsp[a], 1};
            } else {
                P = Prop{0, 0, lsp[0], lsp[0], 0};   // no valid candidate: a null move
            }
        } else {
            P = propose_pick(lsp, dr);
        }
        if (lead && A.propose == 4 && A.accept != 1) {
            A.lp_cs[s] = 0;
            A.lp_lpn[s] = stepc + 1;
        }
        if (lead) {
            int *o = A.prop_next + kPropInts * s;
            o[0] = P.a;
            o[1] = P.b;
            o[2] = P.na;
            o[3] = P.nb;
            o[4] = P.sw;
            o[5] = 0;   // symmetric proposal: log q ratio 0
        }
    }
    __syncthreads();
    const Prop p = P;
    if (lead && A.pend_occ && p.sw) {
        int ax, ay, az, bx, by, bz;
        site_xyz_of(g, p.a, ax, ay, az);
        site_xyz_of(g, p.b, bx, by, bz);
        for (int q = t; q < A.nocc; q += nt)
            A.pend_occ[(long)s * A.noccp + q] =
                lsp[occ_site(g, A.ss_sten, p.a, p.b, ax, ay, az, bx, by, bz, q)];
    }
    if (p.sw)
        gather_swap(g, p, A.ss_sten, A.ms_sten, lsp, hist, first, stride, recs,
                    A.dc_prop + (long)s * g.nrow_w);
}

// Workgroup -> (replica slot s, split) such that all nsplit workgroups of slot s run on
// XCD s mod 8: consecutive workgroup ids are dispatched to the eight XCDs in turn, so ids
// congruent mod 8 share one XCD and its L2, and the slot's species, tile sums, counts and
// path records are fetched into one L2 instead of eight.  1-D grid of
// nsplit * ceil(nrep / 8) * 8 workgroups; the ones past the last slot exit at once.
__device__ __forceinline__ bool xcd_slot(int nrep, int nsplit, int &s, int &split) {
    const int L = blockIdx.x;
    const int j = L >> 3;
    const int m = j / nsplit;
    split = j - m * nsplit;
    s = (L & 7) + 8 * m;
    return s < nrep;
}

__global__ __launch_bounds__(256) void k_slot_step(Geo g, SlotArgs A, int nsplit, int nfold) {
    int s, split;
    if (!xcd_slot(A.nrep, nsplit + nfold, s, split)) return;
    slot_step_body(g, A, s, split, nsplit, nfold);
}

// ---------------------------------------------------------------- paired steps
// Two Metropolis steps per launch pair, by pre-fetching (Brockwell, J. Comput. Graph.
// Stat. 15 (2006) 246): the trial of step n+2 depends only on whether step n+1 is
// accepted, so the transform launch computes the trial of step n+1 (m1 = m + D1 B), of
// step n+2 after a rejection (m2r = m + D2r B) and after an acceptance (m2a = m1 + D2a B),
// each in the arithmetic of a single step; the per-replica launch below takes both
// decisions in order, commits them, and proposes and gathers the next pair: step n+3 on
// the final configuration, step n+4 on it (r) and on it with step n+3's swap made (a).
// Decisions, counts, chi^2 and records are those of the steps run one at a time.
struct PairArgs {
    const double *partials[3];   // (nrep, nsp, nct, 4) of m1, m2r, m2a
    const double *mtrial[3];     // (nrep, ncol_tot) m1, m2r, m2a
    const double *q0;
    double *mcur;
    const double *beta_slot;
    const int *prop_acc[3];      // this round's proposals: step 1, step 2 | rejected, | accepted
    int *dc_acc[3];              //   and count-change rows (zeroed here)
    int *prop_next[3];           // next round's (propose 3), or [0] the single step's (propose 1)
    int *dc_next[3];             //   count-change rows (zero on entry)
    const double *c2_prev;       // chi^2 before the round (read by every workgroup)
    double *c2_next;             // chi^2 after it (written by the lead; another buffer)
    int *counts;
    uint8_t *species, *best_sp;
    double *best_c2;
    const double *best_snap;   // best_c2 as the trial transform before this launch copied it
    const int *rung_of;
    const uint64_t *stepc;       // n + 2 after the round's transform (n when proposing only)
    long long *acc_rung;
    int *anchor;
    long long *nanchor;
    int reanchor_step, reanchor_every;
    const int *ss_sten;
    const int4 *ms_sten;
    int *log;
    int log_cap;
    int *series;
    int series_cap, series_rung;
    long long rec_base;
    uint64_t seed;
    int profile;
    int nrep;
    int accept;                  // 0: propose only, 1: the round's two steps
    int propose;                 // 0: nothing, 1: the single step after the round, 3: next round
    int gr2;                     // largest squared gather offset, units (a0/2)^2
    double nfit;
    const int *draws;            // (nrep, kDrawInts) made by the transform just before, or null
};

__device__ __forceinline__ int min_image(int d, int n) {
    d = (d < 0) ? d + n : d;   // |d| < n
    return (2 * d > n) ? d - n : d;
}
__device__ __forceinline__ void write_prop(int *o, const Prop &P) {
    o[0] = P.a;
    o[1] = P.b;
    o[2] = P.na;
    o[3] = P.nb;
    o[4] = P.sw;
    o[5] = 0;   // symmetric proposal: log q ratio 0
}

__device__ __forceinline__ void slot_pair_body(const Geo &g, const PairArgs &A, int s, int split, int nsplit) {
    extern __shared__ int lds[];
    const int R4 = (g.nrow_tot + 3) & ~3;
    int *hist0 = lds, *hist1 = lds + R4, *hist2 = lds + 2 * R4;
    uint8_t *lsp = reinterpret_cast<uint8_t *>(lds + 3 * R4);
    uint8_t *lsp2 = lsp + g.natoms_pad;
    __shared__ double sums[3][chm::kMaxNsp][4];
    __shared__ int flag;
    __shared__ Prop Ps[3];
    const int nwg = nsplit;
    const bool lead = split == 0;
    const int t = threadIdx.x;
    const int nt = blockDim.x;
    if (t == 0) flag = 0;
    const long base_sp = (long)s * g.natoms_pad;
    // ------------------------------------------------ every input, in one round trip
    const uint64_t stepc = A.stepc[s];
    const int rung = A.rung_of[s];
    const double beta = A.beta_slot[s];
    const double c2p = A.c2_prev[s];
    const double best = A.best_snap[s];   // the best before this launch (see snap_best)
    const int anchor = A.anchor[s];
    int pr[3][kPropInts];
#pragma unroll
    for (int v = 0; v < 3; ++v)
#pragma unroll
        for (int q = 0; q < kPropInts; ++q) pr[v][q] = A.prop_acc[v][kPropInts * s + q];
    double q0v[2 * chm::kMaxNsp];
#pragma unroll
    for (int q = 0; q < 2 * chm::kMaxNsp; ++q) q0v[q] = A.q0[q];
    const int nv = g.natoms_pad / 16;
    const uint4 *gsp = reinterpret_cast<const uint4 *>(A.species + base_sp);
    const int4 spv = buf_i128(make_buf(gsp, 16u * nv), t < nv, 16u * t);
    // tile sums of the three trials: (e, j) groups of CT lanes, tiles c0, c0 + CT, ...
    const int G = 4 * g.nsp;
    const int CT = (G <= 16) ? 16 : 8;
    const int gi = t / CT, c0 = t - gi * CT;
    const bool sum_lane = A.accept && t < G * CT;
    const uint32_t poff = 8u * ((gi >> 2) * g.nct * 4 + (gi & 3));
    constexpr int PR = 6;
    double pv[3][PR];
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const Buf bp = make_buf(A.partials[v] + (long)s * g.nsp * g.nct * 4, 32u * g.nsp * g.nct);
#pragma unroll
        for (int r = 0; r < PR; ++r)
            pv[v][r] = buf_f64(bp, sum_lane && c0 + r * CT < g.nct, poff + 32u * (c0 + r * CT));
    }
    // this workgroup's share of the running chi(R) (one of the three trials is kept)
    constexpr int MC = 2;
    double mt[3][MC];
    const int mchunk = (g.ncol_tot + nwg - 1) / nwg;
    const int mbeg = split * mchunk, mend = min(g.ncol_tot, mbeg + mchunk);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const Buf bm = make_buf(A.mtrial[v] + (long)s * g.ncol_tot, 8u * g.ncol_tot);
#pragma unroll
        for (int r = 0; r < MC; ++r) {
            const int q = mbeg + t + r * nt;
            mt[v][r] = buf_f64(bm, A.accept && q < mend, 8u * q);
        }
    }
    // ... and of the count commit; with the pair-count series of this rung recorded, the
    // lead commits every row (and writes the series from the values it holds)
    int *cn = A.counts + (long)s * g.nrow_tot;
    const long long ri1 = (long long)stepc - 2 - A.rec_base, ri2 = ri1 + 1;
    const bool series_on = A.accept && A.series && rung == A.series_rung &&
                           ((ri1 >= 0 && ri1 < A.series_cap) || (ri2 >= 0 && ri2 < A.series_cap));
    // (chunks in words of the packed count-change rows: rows 2w, 2w+1; the prefetched chunk
    // does not depend on series_on, which needs the step counter)
    const int wchunk = (g.nrow_w + nwg - 1) / nwg;
    const int pbeg = split * wchunk, pend = min(g.nrow_w, pbeg + wchunk);
    const int rbeg = series_on ? 0 : pbeg, rend = series_on ? (lead ? g.nrow_w : 0) : pend;
    constexpr int RC = 2;
    int dcv[3][RC];
    int2 cnv[RC];
    {
        const Buf bcn = make_buf(cn, 4u * g.nrow_tot);
#pragma unroll
        for (int r = 0; r < RC; ++r) cnv[r] = buf_i64(bcn, A.accept && pbeg + t + r * nt < pend, 8u * (pbeg + t + r * nt));
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            const Buf bd = make_buf(A.dc_acc[v] + (long)s * g.nrow_w, 4u * g.nrow_w);
#pragma unroll
            for (int r = 0; r < RC; ++r)
                dcv[v][r] = buf_i32(bd, A.accept && pbeg + t + r * nt < pend, 4u * (pbeg + t + r * nt));
        }
    }
    const int stride = nwg * nt;
    const int first = split * nt + t;
    RecPre recs;
    prefetch_recs(g, A.ss_sten, A.ms_sten, first, stride, A.propose != 0, recs);
    // the round's acceptance uniforms and the next proposals' draws (wave-uniform)
    // (made by the transform when it ran just before; else here, by wave 0 only: lane 0
    // takes the decisions and makes the picks)
    double u1 = 1.0, u2 = 1.0;
    PropDraws dr3{}, dr4{};
    if (A.draws) {
        const int *D = A.draws + (long)s * kDrawInts;
        if (A.accept) {
            u1 = draw_u(D, 0);
            u2 = draw_u(D, 1);
        }
        if (A.propose) dr3 = draw_p(D, 0);
        if (A.propose == 3) dr4 = draw_p(D, 1);
    } else if (t < 64) {
        if (A.accept) {
            uint64_t h = splitmix64(A.seed ^ ((stepc - 1) * kGolden) ^ ((uint64_t)rung << 32) ^ kAccSalt);
            u1 = u01(splitmix64(h));
            h = splitmix64(A.seed ^ (stepc * kGolden) ^ ((uint64_t)rung << 32) ^ kAccSalt);
            u2 = u01(splitmix64(h));
        }
        if (A.propose) dr3 = propose_draws(g.natoms, A.seed, stepc + 1, (uint64_t)rung);
        if (A.propose == 3) dr4 = propose_draws(g.natoms, A.seed, stepc + 2, (uint64_t)rung);
    }
    __builtin_amdgcn_sched_barrier(0);   // every load issued before the first use
    // ------------------------------------------------ consume
    if (t < nv) reinterpret_cast<int4 *>(lsp)[t] = spv;
    for (int q = t + nt; q < nv; q += nt) reinterpret_cast<uint4 *>(lsp)[q] = gsp[q];
    if (A.accept) {
#pragma unroll
        for (int v = 0; v < 3; ++v) {
            double psum = 0.0;   // out-of-range tiles loaded as 0.0
#pragma unroll
            for (int r = 0; r < PR; ++r) psum += pv[v][r];
            const double *Pp = A.partials[v] + (long)s * g.nsp * g.nct * 4;
            for (int c = c0 + PR * CT; sum_lane && c < g.nct; c += CT)
                psum += Pp[(long)((gi >> 2) * g.nct + c) * 4 + (gi & 3)];
            psum = (CT == 16) ? group_sum_d<16>(psum) : group_sum_d<8>(psum);
            if (t < G * CT && c0 == 0) sums[v][gi >> 2][gi & 3] = psum;
        }
    }
    if (A.propose)
        for (int q = t; q < 3 * R4; q += nt) lds[q] = 0;
    __syncthreads();
    int acc1 = 0, acc2 = 0, bestsel = 0;
    if (A.accept) {
        // ------------------------------------------------ both decisions (every workgroup)
        // chi^2 of the three trials on lanes 0..2 at once
        double c2v = 0.0;
        if (t < 3) c2v = chi2_of_sums(g, sums[t], q0v, A.profile, A.nfit);
        double c2_1 = 0.0, c2_2r = 0.0, c2_2a = 0.0;
        if (t < 64) {
            c2_1 = __shfl(c2v, 0);
            c2_2r = __shfl(c2v, 1);
            c2_2a = __shfl(c2v, 2);
        }
        if (t == 0) {
            int p1[kPropInts];
#pragma unroll
            for (int q = 0; q < kPropInts; ++q) p1[q] = pr[0][q];
            if (p1[4]) {
                const double x = beta * (c2_1 - c2p) * A.nfit - (double)__int_as_float(p1[5]);
                acc1 = metropolis(x, u1);
            }
            const double cur1 = acc1 ? c2_1 : c2p;
            int q2[kPropInts];
#pragma unroll
            for (int q = 0; q < kPropInts; ++q) q2[q] = acc1 ? pr[2][q] : pr[1][q];
            const double c2_2 = acc1 ? c2_2a : c2_2r;
            if (q2[4]) {
                const double x = beta * (c2_2 - cur1) * A.nfit - (double)__int_as_float(q2[5]);
                acc2 = metropolis(x, u2);
            }
            const double cur2 = acc2 ? c2_2 : cur1;
            if (acc1) {
                lsp[p1[0]] = (uint8_t)p1[2];
                lsp[p1[1]] = (uint8_t)p1[3];
            }
            if (acc2) {
                lsp[q2[0]] = (uint8_t)q2[2];
                lsp[q2[1]] = (uint8_t)q2[3];
            }
            // best retention after each step, in order (method_a_rmc.py:301-307)
            double bb = best;
            if (cur1 < bb) {
                bb = cur1;
                bestsel = 1;
            }
            if (cur2 < bb) {
                bb = cur2;
                bestsel = 2;
            }
            flag = acc1 | (acc2 << 1) | (bestsel << 2);
            if (lead) {
                A.c2_next[s] = cur2;
                if (bestsel) A.best_c2[s] = bb;
                if (acc1) {
                    A.species[base_sp + p1[0]] = (uint8_t)p1[2];
                    A.species[base_sp + p1[1]] = (uint8_t)p1[3];
                }
                if (acc2) {
                    A.species[base_sp + q2[0]] = (uint8_t)q2[2];
                    A.species[base_sp + q2[1]] = (uint8_t)q2[3];
                }
                if (acc1 + acc2)
                    atomicAdd(reinterpret_cast<unsigned long long *>(A.acc_rung + rung), (unsigned long long)(acc1 + acc2));
                if (A.log) {
                    if (ri1 >= 0 && ri1 < A.log_cap) {
                        int *L = A.log + ((long)ri1 * A.nrep + s) * 4;
                        L[0] = p1[0];
                        L[1] = p1[1];
                        L[2] = acc1;
                        L[3] = rung;
                    }
                    if (ri2 >= 0 && ri2 < A.log_cap) {
                        int *L = A.log + ((long)ri2 * A.nrep + s) * 4;
                        L[0] = q2[0];
                        L[1] = q2[1];
                        L[2] = acc2;
                        L[3] = rung;
                    }
                }
                if (A.reanchor_step && anchor >= A.reanchor_every) {
                    A.anchor[s] = acc1 + acc2;
                    atomicAdd(reinterpret_cast<unsigned long long *>(A.nanchor + s), 1ull);
                } else if (acc1 + acc2) {
                    A.anchor[s] = anchor + acc1 + acc2;
                }
            }
        }
        __syncthreads();
        acc1 = flag & 1;
        acc2 = (flag >> 1) & 1;
        bestsel = flag >> 2;
        const int v2 = acc1 ? 2 : 1;
        // ------------------------------------------------ commit of this workgroup's share
        const int nss = g.nsp * g.nshell;
        int2 *cn2 = reinterpret_cast<int2 *>(cn);
        int *dw0 = A.dc_acc[0] + (long)s * g.nrow_w, *dw1 = A.dc_acc[1] + (long)s * g.nrow_w,
            *dw2 = A.dc_acc[2] + (long)s * g.nrow_w;
        auto commit_row = [&](int q, int c, int d1, int d2r, int d2a) {   // row q, returns cB
            const int d2 = acc1 ? d2a : d2r;
            const int cA = c + (acc1 ? d1 : 0), cB = cA + (acc2 ? d2 : 0);
            if (series_on) {
                const int a = q / g.rstride, r = q - a * g.rstride;
                if (a < g.nsp && r < nss) {
                    const int idx = a * nss + r, ns = g.nsp * nss;
                    if (ri1 >= 0 && ri1 < A.series_cap) A.series[ri1 * ns + idx] = cA;
                    if (ri2 >= 0 && ri2 < A.series_cap) A.series[ri2 * ns + idx] = cB;
                }
            }
            return cB;
        };
        auto commit_word = [&](int w, int2 c, int d1, int d2r, int d2a) {
            const int lo = commit_row(2 * w, c.x, dw_lo(d1), dw_lo(d2r), dw_lo(d2a));
            const int hi = commit_row(2 * w + 1, c.y, dw_hi(d1), dw_hi(d2r), dw_hi(d2a));
            if (lo != c.x || hi != c.y) cn2[w] = make_int2(lo, hi);
            if (d1) dw0[w] = 0;
            if (d2r) dw1[w] = 0;
            if (d2a) dw2[w] = 0;
        };
        // words rbeg + t + r nt: prefetched when inside this workgroup's chunk (rbeg == pbeg
        // whenever rend > 0: series_on commits all rows in the lead, whose chunk starts at 0)
#pragma unroll
        for (int r = 0; r < RC; ++r) {
            const int w = rbeg + t + r * nt;
            if (w < rend) {
                if (w < pend)
                    commit_word(w, cnv[r], dcv[0][r], dcv[1][r], dcv[2][r]);
                else
                    commit_word(w, cn2[w], dw0[w], dw1[w], dw2[w]);
            }
        }
        for (int w = rbeg + t + RC * nt; w < rend; w += nt) commit_word(w, cn2[w], dw0[w], dw1[w], dw2[w]);
        // the running chi(R): the trial of the last accepted step
        const int fv = acc2 ? v2 : (acc1 ? 0 : -1);
        if (fv >= 0) {
            double *mc = A.mcur + (long)s * g.ncol_tot;
#pragma unroll
            for (int r = 0; r < MC; ++r) {
                const int q = mbeg + t + r * nt;
                if (q < mend) mc[q] = (fv == 0) ? mt[0][r] : ((fv == 1) ? mt[1][r] : mt[2][r]);
            }
            for (int q = mbeg + t + MC * nt; q < mend; q += nt) mc[q] = A.mtrial[fv][(long)s * g.ncol_tot + q];
        }
        if (bestsel) {
            // the configuration after step 1 (bestsel 1 with step 2 accepted) is the final
            // one with step 2's swap undone, built in lsp2
            const uint8_t *src = lsp;
            if (bestsel == 1 && acc2) {
                for (int q = t; q < nv; q += nt) reinterpret_cast<uint4 *>(lsp2)[q] = reinterpret_cast<const uint4 *>(lsp)[q];
                __syncthreads();
                if (t == 0) {
                    const int a2 = acc1 ? pr[2][0] : pr[1][0], b2 = acc1 ? pr[2][1] : pr[1][1];
                    const int na2 = acc1 ? pr[2][2] : pr[1][2], nb2 = acc1 ? pr[2][3] : pr[1][3];
                    lsp2[a2] = (uint8_t)nb2;
                    lsp2[b2] = (uint8_t)na2;
                }
                __syncthreads();
                src = lsp2;
            }
            const int vchunk = (nv + nwg - 1) / nwg;
            const int vend = min(nv, (split + 1) * vchunk);
            uint4 *dst = reinterpret_cast<uint4 *>(A.best_sp + base_sp);
            for (int q = split * vchunk + t; q < vend; q += nt) dst[q] = reinterpret_cast<const uint4 *>(src)[q];
        }
    }
    if (!A.propose) return;
    // ------------------------------------------------ propose + gather of the next step(s)
    if (t == 0) {
        const Prop P3 = propose_pick(lsp, dr3);
        Ps[0] = P3;
        if (A.propose == 3) {
            Ps[1] = propose_pick(lsp, dr4);
            // step n+4 after step n+3 is accepted: picked with P3's swap made
            Ps[2] = P3.sw ? propose_pick_sp(SpOver{lsp, P3.a, P3.b, P3.na, P3.nb}, dr4) : Ps[1];
        }
        if (lead) {
            write_prop(A.prop_next[0] + kPropInts * s, P3);
            if (A.propose == 3) {
                write_prop(A.prop_next[1] + kPropInts * s, Ps[1]);
                write_prop(A.prop_next[2] + kPropInts * s, Ps[2]);
            }
        }
    }
    __syncthreads();
    const Prop P3 = Ps[0], P4r = Ps[1], P4a = Ps[2];
    const bool pair = A.propose == 3;
    SwapSite W3, W4r, W4a;
    W3.p = P3;
    W4r.p = P4r;
    W4a.p = P4a;
    site_xyz_of(g, P3.a, W3.ax, W3.ay, W3.az);
    site_xyz_of(g, P3.b, W3.bx, W3.by, W3.bz);
    // the a-variant rows equal the r-variant ones when the picks agree and P3's sites lie
    // outside the gather stencil (squared offsets <= gr2) of P4's two sites
    bool same = true;
    if (pair) {
        site_xyz_of(g, P4r.a, W4r.ax, W4r.ay, W4r.az);
        site_xyz_of(g, P4r.b, W4r.bx, W4r.by, W4r.bz);
        if (P3.sw) {
            same = P4a.a == P4r.a && P4a.b == P4r.b && P4a.na == P4r.na && P4a.nb == P4r.nb &&
                   P4a.sw == P4r.sw;
            if (same && P4r.sw) {
                auto d2 = [&](int ux, int uy, int uz, int vx, int vy, int vz) {
                    const int dx = min_image(vx - ux, g.box[0]), dy = min_image(vy - uy, g.box[1]),
                              dz = min_image(vz - uz, g.box[2]);
                    return dx * dx + dy * dy + dz * dz;
                };
                same = d2(W3.ax, W3.ay, W3.az, W4r.ax, W4r.ay, W4r.az) > A.gr2 &&
                       d2(W3.ax, W3.ay, W3.az, W4r.bx, W4r.by, W4r.bz) > A.gr2 &&
                       d2(W3.bx, W3.by, W3.bz, W4r.ax, W4r.ay, W4r.az) > A.gr2 &&
                       d2(W3.bx, W3.by, W3.bz, W4r.bx, W4r.by, W4r.bz) > A.gr2;
            }
        }
        if (!same) {
            site_xyz_of(g, P4a.a, W4a.ax, W4a.ay, W4a.az);
            site_xyz_of(g, P4a.b, W4a.bx, W4a.by, W4a.bz);
        }
    }
    const SpOver ov{lsp, P3.a, P3.b, P3.na, P3.nb};
    const int nss_items = 2 * g.ncoord;
    const int nitems = nss_items + 2 * g.nstencil;
    auto items = [&](int it, const int4 rec) {
        if (P3.sw) gather_item(g, W3, lsp, hist0, it, nss_items, rec);
        if (pair && P4r.sw) gather_item(g, W4r, lsp, hist1, it, nss_items, rec);
        if (pair && !same && P4a.sw) gather_item_sp(g, W4a, ov, hist2, it, nss_items, rec);
    };
#pragma unroll
    for (int r = 0; r < kRecPrefetch; ++r) {
        const int it = first + r * stride;
        if (it < nitems) items(it, it < nss_items ? make_int4(recs.ss[r], 0, 0, 0) : recs.ms[r]);
    }
    for (int it = first + kRecPrefetch * stride; it < nitems; it += stride)
        items(it, item_rec(g, A.ss_sten, A.ms_sten, it));
    __syncthreads();
    int *d3 = A.dc_next[0] + (long)s * g.nrow_w;
    int *d4r = pair ? A.dc_next[1] + (long)s * g.nrow_w : nullptr;
    int *d4a = pair ? A.dc_next[2] + (long)s * g.nrow_w : nullptr;
    for (int w = t; w < g.nrow_w; w += nt) {
        const int h0 = dw_pack(hist0[2 * w], hist0[2 * w + 1]);
        if (h0 != 0) atomicAdd(&d3[w], h0);
        if (pair) {
            const int h1 = dw_pack(hist1[2 * w], hist1[2 * w + 1]);
            const int h2 = same ? h1 : dw_pack(hist2[2 * w], hist2[2 * w + 1]);
            if (h1 != 0) atomicAdd(&d4r[w], h1);
            if (h2 != 0) atomicAdd(&d4a[w], h2);
        }
    }
}

__global__ __launch_bounds__(256) void k_slot_pair(Geo g, PairArgs A, int nsplit) {
    int s, split;
    if (!xcd_slot(A.nrep, nsplit, s, split)) return;
    slot_pair_body(g, A, s, split, nsplit);
}

constexpr int kPairKB = 3;   // K blocks per wave in the paired transform

struct PairTile {   // the three trials of a round
    const int *dc[3];    // D1, D2r, D2a
    TileOut O[3];        // m1, m2r, m2a (O[0] advances the step counter by two)
};

__global__ __launch_bounds__(256) void k_gemm_pair_bf16(Geo g, int nrep, const bf16x8 *__restrict__ Tb,
                                                        const double *__restrict__ mcur, PairTile P) {
    // one workgroup per tile computes all three trials (B read once); 3 K blocks per wave
    // (nkb <= 12) keep the three operand sets within 2 workgroups per CU
    if ((int)blockIdx.x == g.nct) {   // the extra column of workgroups: draws + step counter
        if (blockIdx.z == 0) compute_draws(P.O[0], nrep, blockIdx.y * RB);
        return;
    }
    gemm_delta_bf16_tile<3, kPairKB>(g, nrep, P.dc, Tb, mcur, P.O, blockIdx.x, blockIdx.y, blockIdx.z);
}

// chi^2 of arbitrary window transforms (one row per workgroup), direct form:
// scale from the sums, then sum |s m - t|^2 w (method_a_rmc.py:211-227)
__global__ __launch_bounds__(256) void k_eval_rows(Geo g, const double *__restrict__ rows,
                                                    const double *__restrict__ target,
                                                    const double *__restrict__ wcol, int profile,
                                                    double nfit, double *__restrict__ out) {
    __shared__ double red[256];
    __shared__ double scale[chm::kMaxNsp];
    const double *m = rows + (long)blockIdx.x * g.ncol_tot;
    for (int e = 0; e < g.nsp; ++e) {
        double num = 0.0, den = 0.0;
        for (int q = threadIdx.x; q < g.ncol_e; q += blockDim.x) {
            const double v = m[(long)e * g.ncol_e + q];
            num += v * target[(long)e * g.ncol_e + q];
            den += v * v;
        }
        num = block_sum(num, red);
        den = block_sum(den, red);
        if (threadIdx.x == 0) scale[e] = profile ? num / fmax(den, 1e-30) : 1.0;
    }
    __syncthreads();
    double res = 0.0;
    for (int q = threadIdx.x; q < g.ncol_tot; q += blockDim.x) {
        const double d = scale[q / g.ncol_e] * m[q] - target[q];
        res += d * d * wcol[q];
    }
    res = block_sum(res, red);
    if (threadIdx.x == 0) out[blockIdx.x] = res / nfit;
}

// ---------------------------------------------------------------- K4
__global__ void k_exchange(int nrep, const double *__restrict__ beta, const double *__restrict__ c2,
                           int *__restrict__ slot_of, int *__restrict__ rung_of,
                           double *__restrict__ beta_slot, uint64_t *__restrict__ iex,
                           long long *__restrict__ att, long long *__restrict__ acc, uint64_t seed,
                           double nfit) {
    exchange_body(nrep, beta, c2, slot_of, rung_of, beta_slot, iex, att, acc, seed, nfit);
}

// chi(k) of every slot: counts @ basis_k per edge (for reading back)
__global__ void k_counts_to_f64(const int *__restrict__ counts, double *__restrict__ out, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (double)counts[t];
}

}  // namespace


namespace {

TileOut tile_out(chm_sampler *S) {
    TileOut O{};
    O.stepc = S->stepc.as<uint64_t>();
    O.bump = 0;
    O.best_c2 = S->best_c2.as<double>();
    O.best_snap = S->best_snap.as<double>();
    O.target = S->target.as<double>();
    O.wcol = S->wcol.as<double>();
    O.mtrial = S->mtrial.as<double>();
    O.partials = S->partials.as<double>();
    O.acc32 = S->acc32;
    return O;
}

SlotArgs slot_args(chm_sampler *S, int accept, int propose) {
    SlotArgs A{};
    A.partials = S->partials.as<double>();
    A.q0 = S->q0.as<double>();
    A.mtrial = S->mtrial.as<double>();
    A.target = S->target.as<double>();
    A.wcol = S->wcol.as<double>();
    A.basisR = S->basisR.as<double>();
    A.mcur = S->mcur.as<double>();
    A.beta_slot = S->beta_slot.as<double>();
    A.counts = S->counts.as<int>();
    A.species = S->species.as<uint8_t>();
    A.best_sp = S->best_sp.as<uint8_t>();
    A.best_c2 = S->best_c2.as<double>();
    A.best_snap = S->best_snap.as<double>();
    A.rung_of = S->rung_of.as<int>();
    A.stepc = S->stepc.as<uint64_t>();
    A.acc_rung = S->acc_rung.as<long long>();
    A.anchor = S->anchor.as<int>();
    A.nanchor = S->nanchor.as<long long>();
    A.ss_sten = S->ss_sten.as<int>();
    A.ms_sten = S->stencil.as<int4>();
    A.reanchor_every = S->reanchor_every;
    A.incremental = S->incremental;
    A.log = S->log_cap ? S->log.as<int>() : nullptr;
    A.log_cap = S->log_cap;
    A.series = S->series_cap ? S->series.as<int>() : nullptr;
    A.series_cap = S->series_cap;
    A.series_rung = S->series_rung;
    A.rec_base = S->rec_base;
    const chm_lp &L = S->lp;
    A.nocc = L.nocc;
    A.noccp = L.noccp;
    A.pend_occ = L.rec_cap ? L.pend_occ.as<uint8_t>() : nullptr;
    A.rec_cap = L.rec_cap;
    A.rec_occ = L.rec_cap ? L.rec_occ.as<uint8_t>() : nullptr;
    A.rec_y = L.rec_cap ? L.rec_y.as<float>() : nullptr;
    A.rec_meta = L.rec_cap ? L.rec_meta.as<int>() : nullptr;
    A.lp_eps = L.eps;
    A.lp_beta_max = L.beta_max;
    A.lp_lam = L.lam;
    A.lp_kappa = L.kappa;
    A.lp_lift = L.lift;
    A.lp_ncand = L.ncand;
    A.seed = S->seed;
    A.profile = S->profile;
    A.nrep = S->nrep;
    A.accept = accept;
    A.propose = propose;
    A.nfit = S->nfit;
    return A;
}

size_t lds_gemm(const Geo &g) {
    size_t a = sizeof(double) * (size_t)g.nrow_e * (RB + 1);
    size_t p = sizeof(double) * (size_t)KS * RB * NCW;
    return std::max(a, p);
}

// grid x and dynamic LDS of a transform launch with the learned proposal's workgroups (lp)
void lp_ride(chm_sampler *S, const chm_lp_dev::LPArgs *lp, dim3 &grid, size_t &lds) {
    if (!lp) return;
    grid.x += (unsigned)lp->ntile_blocks;
    lds = std::max(lds, lp_tile_lds(S));
}

int launch_gemm_abs(chm_sampler *S, const int *dc, int bump, const chm_lp_dev::LPArgs *lp = nullptr) {
    const Geo &g = S->g;
    dim3 grid(g.nct, (S->nrep + RB - 1) / RB, g.nsp);
    TileOut O = tile_out(S);
    O.bump = bump;
    size_t lds = lds_gemm(g);
    lp_ride(S, lp, grid, lds);
    k_gemm_abs_f64<<<grid, 256, lds, S->ctx->stream>>>(g, S->nrep, S->counts.as<int>(), dc,
                                                       S->basisR.as<double>(), O,
                                                       lp ? *lp : chm_lp_dev::LPArgs{});
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

// the draws of the next per-replica launch, made by the transform's extra workgroup column
void set_draws(chm_sampler *S, TileOut &O, int mode) {
    O.draws = S->draws.as<int>();
    O.draw_mode = mode;
    O.rung_of = S->rung_of.as<int>();
    O.seed = S->seed;
    O.natoms = S->g.natoms;
}

int launch_gemm_delta(chm_sampler *S, const int *dc, int bump, const chm_lp_dev::LPArgs *lp = nullptr) {
    const Geo &g = S->g;
    dim3 grid(g.nct + 1, (S->nrep + RB - 1) / RB, g.nsp);
    TileOut O = tile_out(S);
    O.bump = bump;
    set_draws(S, O, 1);
    size_t lds = 0;
    lp_ride(S, lp, grid, lds);
    k_gemm_delta_bf16<<<grid, 256, lds, S->ctx->stream>>>(g, S->nrep, dc, reinterpret_cast<const bf16x8 *>(S->Tb.p),
                                                          S->mcur.as<double>(), O, lp ? *lp : chm_lp_dev::LPArgs{});
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

// the learned proposal's buffers of the per-replica launch that decides step n (accept 1)
// or proposes step n + 1 after n steps (accept 0)
void lp_slot_args(chm_sampler *S, SlotArgs &A, int accept, long long n) {
    chm_lp &L = S->lp;
    if (!L.mode) return;
    const size_t np = (size_t)S->nrep * 2 * chm_lp_dev::MAXC, nt = (size_t)S->nrep * chm_lp_dev::MAXC;
    const long long nx = n + 1;
    A.lp_lq = L.lq.as<double>();
    A.lp_dmax = L.dmax.as<float>();
    A.lp_sg = L.sg.as<float>() + (size_t)(nx & 1) * np;
    A.lp_sdu = L.sdu.as<int>() + (size_t)(nx & 1) * np;
    A.lp_st = L.st8.as<int8_t>() + (size_t)(nx & 1) * nt;
    if (L.lift) {
        // decided step n: direction before it in copy n & 1, after it in copy (n + 1) & 1;
        // a segment's first proposal reads the direction after its n steps
        A.lp_dir_rd = L.dir.as<int>() + (size_t)((accept == 1 ? n : nx) & 1) * S->nrep;
        A.lp_dir_wr = L.dir.as<int>() + (size_t)(nx & 1) * S->nrep;
    }
    A.lp_cache = L.cache.as<int>();
    A.lp_w1q = L.w1q.as<int>();
    A.lp_nbr = S->nbr_st.as<int>();
    A.lp_posmap = L.posmap.as<signed char>();
    A.lp_pm_r = L.pm_r;
    A.lp_cs = L.cs.as<int>();
    A.lp_lpn = L.lpn.as<uint64_t>();
}

// accept: 0 none, 1 step n, 2 init (n = 0).  propose: step n + 1.
int launch_slot(chm_sampler *S, int accept, int propose, long long n, int reanchor_step = 0,
                const double *c2_prev = nullptr) {
    SlotArgs A = slot_args(S, accept, propose);
    if (accept != 2) lp_slot_args(S, A, accept, n);
    A.reanchor_step = reanchor_step;
    A.dc_acc = drow(S, n);
    A.dc_prop = drow(S, n + 1);
    A.prop_acc = prow(S, n);
    A.prop_next = prow(S, n + 1);
    A.c2_prev = c2_prev ? c2_prev : c2row(S, n - 1);
    A.draws = (accept == 1 && S->incremental == 1) ? S->draws.as<int>() : nullptr;   // see launch_gemm_delta
    A.c2_next = c2row(S, n);
    // (a decided learned step: the slot's folding workgroups follow its nsplit ones; the
    // learned proposal's launches take their own split, the uniform sampler's tuned one
    // being too many workgroups beside the folding ones)
    const int nfold = (A.lp_cache && accept == 1) ? chm_lp_dev::kFoldPerSlot : 0;
    const int nsplit = S->lp.mode ? S->lp.nsplit : S->nsplit;
    k_slot_step<<<dim3((nsplit + nfold) * ((S->nrep + 7) & ~7)), 256, S->lds_k1, S->ctx->stream>>>(
        S->g, A, nsplit, nfold);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

// propose step h + 1 after h completed steps (nothing is accepted); the learned proposal
// scores step h + 1's candidates first
int launch_propose(chm_sampler *S, long long h) {
    if (S->lp.mode) {
        int rc = lp_launch_tiles0(S, h);
        return rc ? rc : launch_slot(S, 0, 4, h);
    }
    return launch_slot(S, 0, 1, h);
}

// accept step n, then (propose) propose step n + 1
int launch_accept(chm_sampler *S, long long n, int propose, int reanchor_step = 0) {
    return launch_slot(S, 1, propose ? (S->lp.mode ? 4 : 1) : 0, n, reanchor_step);
}


int launch_reanchor(chm_sampler *S) {
    const Geo &g = S->g;
    k_reanchor_f64<<<dim3(g.nct, S->nrep, g.nsp), 256, 0, S->ctx->stream>>>(
        g, S->nrep, S->counts.as<int>(), S->basisR.as<double>(), S->anchor.as<int>(), S->reanchor_every,
        S->mcur.as<double>(), S->acc32);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int launch_gemm_delta32(chm_sampler *S, const int *dc, int bump, const chm_lp_dev::LPArgs *lp = nullptr) {
    const Geo &g = S->g;
    dim3 grid(g.nct, (S->nrep + RB - 1) / RB, g.nsp);
    TileOut O = tile_out(S);
    O.bump = bump;
    size_t lds = std::max(sizeof(float) * (size_t)g.nrow_e * (RB + 1), sizeof(float) * (size_t)KS * RB * NCW);
    lp_ride(S, lp, grid, lds);
    k_gemm_delta_f32<<<grid, 256, lds, S->ctx->stream>>>(g, S->nrep, dc, S->basisR32.as<float>(),
                                                          S->mcur.as<double>(), O, lp ? *lp : chm_lp_dev::LPArgs{});
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int launch_gemm(chm_sampler *S, const int *dc, int bump, const chm_lp_dev::LPArgs *lp = nullptr) {
    if (S->incremental == 2) return launch_gemm_delta32(S, dc, bump, lp);
    return S->incremental ? launch_gemm_delta(S, dc, bump, lp) : launch_gemm_abs(S, dc, bump, lp);
}

// the transform of step n; with the learned proposal, the same launch carries the scoring
// workgroups (step n's reverse probability, step n+1's candidates)
int launch_trial(chm_sampler *S, long long n, bool more) {
    if (!S->lp.mode) return launch_gemm(S, drow(S, n), 1);
    const chm_lp_dev::LPArgs lp = lp_step_args(S, more ? 1 : 2, n);
    return launch_gemm(S, drow(S, n), 1, &lp);
}

// ------------------------------------------------ paired steps
int *pdc_row(chm_sampler *S, int k, int v) {
    return S->pdc.as<int>() + ((size_t)(k & 1) * 3 + v) * S->nrep * S->g.nrow_tot;
}
int *pprop_row(chm_sampler *S, int k, int v) {
    return S->pprop.as<int>() + ((size_t)(k & 1) * 3 + v) * S->nrep * kPropInts;
}
double *pc2_row(chm_sampler *S, int k) { return S->pc2.as<double>() + (size_t)(k & 1) * S->nrep; }

bool use_pairs(const chm_sampler *S, int L) {
    return S->pairs && S->incremental == 1 && !S->lp.mode && !S->lp.rec_cap && L >= 3 &&
           S->g.nkb <= 4 * kPairKB;
}

// the three trials of round k (bump: step-counter advance, 2 for a real round)
int launch_gemm_pair(chm_sampler *S, int k, int bump) {
    const Geo &g = S->g;
    const size_t mt = (size_t)S->nrep * g.ncol_tot, pt = (size_t)S->nrep * g.nsp * g.nct * 4;
    PairTile P{};
    for (int v = 0; v < 3; ++v) {
        P.dc[v] = pdc_row(S, k, v);
        P.O[v] = tile_out(S);
        P.O[v].mtrial = S->pmt.as<double>() + v * mt;
        P.O[v].partials = S->ppart.as<double>() + v * pt;
        P.O[v].bump = v == 0 ? bump : 0;
    }
    set_draws(S, P.O[0], 2);
    const int nrb = (S->nrep + RB - 1) / RB;
    k_gemm_pair_bf16<<<dim3(g.nct + 1, nrb, g.nsp), 256, 0, S->ctx->stream>>>(
        g, S->nrep, reinterpret_cast<const bf16x8 *>(S->Tb.p), S->mcur.as<double>(), P);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

// round k: accept its two steps (accept 1) and/or propose round k + 1 (propose 3) or the
// single step n1 (propose 1, into the single-step rows of step n1).  k = -1 proposes
// round 0.
int launch_slot_pair(chm_sampler *S, int accept, int propose, int k, const double *c2_prev, double *c2_next,
                     int reanchor_step, long long n1) {
    const Geo &g = S->g;
    const size_t mt = (size_t)S->nrep * g.ncol_tot, pt = (size_t)S->nrep * g.nsp * g.nct * 4;
    PairArgs A{};
    for (int v = 0; v < 3; ++v) {
        A.partials[v] = S->ppart.as<double>() + v * pt;
        A.mtrial[v] = S->pmt.as<double>() + v * mt;
        A.prop_acc[v] = pprop_row(S, k, v);
        A.dc_acc[v] = pdc_row(S, k, v);
        A.prop_next[v] = pprop_row(S, k + 1, v);
        A.dc_next[v] = pdc_row(S, k + 1, v);
    }
    if (propose == 1) {
        A.prop_next[0] = prow(S, n1);
        A.dc_next[0] = drow(S, n1);
    }
    A.q0 = S->q0.as<double>();
    A.mcur = S->mcur.as<double>();
    A.beta_slot = S->beta_slot.as<double>();
    A.c2_prev = c2_prev;
    A.c2_next = c2_next;
    A.counts = S->counts.as<int>();
    A.species = S->species.as<uint8_t>();
    A.best_sp = S->best_sp.as<uint8_t>();
    A.best_c2 = S->best_c2.as<double>();
    A.best_snap = S->best_snap.as<double>();
    A.rung_of = S->rung_of.as<int>();
    A.stepc = S->stepc.as<uint64_t>();
    A.acc_rung = S->acc_rung.as<long long>();
    A.anchor = S->anchor.as<int>();
    A.nanchor = S->nanchor.as<long long>();
    A.reanchor_step = reanchor_step;
    A.reanchor_every = S->reanchor_every;
    A.ss_sten = S->ss_sten.as<int>();
    A.ms_sten = S->stencil.as<int4>();
    A.log = S->log_cap ? S->log.as<int>() : nullptr;
    A.log_cap = S->log_cap;
    A.series = S->series_cap ? S->series.as<int>() : nullptr;
    A.series_cap = S->series_cap;
    A.series_rung = S->series_rung;
    A.rec_base = S->rec_base;
    A.seed = S->seed;
    A.profile = S->profile;
    A.nrep = S->nrep;
    A.accept = accept;
    A.propose = propose;
    A.gr2 = S->gr2;
    A.nfit = S->nfit;
    A.draws = accept ? S->draws.as<int>() : nullptr;   // made by launch_gemm_pair just before
    k_slot_pair<<<dim3(S->nsplit * ((S->nrep + 7) & ~7)), 256, S->lds_pair, S->ctx->stream>>>(g, A,
                                                                                              S->nsplit);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

// L >= 3 steps after h as L/2 paired rounds (and one single step if L is odd).  chi^2
// moves c2row(h) -> pc2(0) -> pc2(1) -> ... and ends in c2row(h + L), where every other
// path expects it; no launch reads the buffer it writes.
int launch_segment_pairs(chm_sampler *S, long long h, int L) {
    const int R = L / 2, odd = L & 1;
    const int anchor = S->reanchor_every > 0;
    int rc = launch_slot_pair(S, 0, 3, -1, c2row(S, h), pc2_row(S, 0), 0, h + L);   // round 0's rows
    if (!rc && anchor) rc = launch_reanchor(S);
    for (int k = 0; k < R && !rc; ++k) {
        rc = launch_gemm_pair(S, k, 2);
        const int prop = (k + 1 < R) ? 3 : (odd ? 1 : 0);
        const double *in = (k == 0) ? c2row(S, h) : pc2_row(S, k - 1);
        double *out = (k == R - 1 && !odd) ? c2row(S, h + L) : pc2_row(S, k);
        if (!rc) rc = launch_slot_pair(S, 1, prop, k, in, out, anchor && k == 0, h + L);
    }
    if (odd && !rc) {
        rc = launch_gemm(S, drow(S, h + L), 1);
        if (!rc) rc = launch_slot(S, 1, 0, h + L, 0, pc2_row(S, R - 1));
    }
    return rc;
}

// L steps after h completed ones, no exchange inside: propose step h+1, then for every
// step n: transform (advances the device step counter), accept [+ propose n+1].
int launch_segment(chm_sampler *S, long long h, int L) {
    if (use_pairs(S, L)) return launch_segment_pairs(S, h, L);
    int rc = launch_propose(S, h);
    // incremental mode: replicas with reanchor_every accepted moves since their last FP64
    // recompute get the running chi(R) recomputed from the integer counts, here
    const int anchor = S->incremental && S->reanchor_every > 0;
    if (!rc && anchor) rc = launch_reanchor(S);
    for (int t = 0; t < L && !rc; ++t) {
        const long long n = h + 1 + t;
        rc = launch_trial(S, n, t + 1 < L);
        if (!rc) rc = launch_accept(S, n, t + 1 < L ? 1 : 0, anchor && t == 0);
    }
    return rc;
}

int launch_exchange(chm_sampler *S, long long n) {
    k_exchange<<<1, 256, 0, S->ctx->stream>>>(
        S->nrep, S->beta.as<double>(), c2row(S, n), S->slot_of.as<int>(),
        S->rung_of.as<int>(), S->beta_slot.as<double>(), S->iex.as<uint64_t>(),
        S->exch_att.as<long long>(),
        S->exch_acc.as<long long>(), S->seed, S->nfit);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}


}  // namespace

namespace chm_smp {
void drop_graph(chm_sampler *S) {
    for (int q = 0; q < 2; ++q) {
        if (S->gexec[q]) hipGraphExecDestroy(S->gexec[q]);
        if (S->graph[q]) hipGraphDestroy(S->graph[q]);
        S->gexec[q] = nullptr;
        S->graph[q] = nullptr;
    }
    S->graph_steps = 0;
}
}  // namespace chm_smp

namespace chm_smp {
// The lattice geometry of a sampler from its host description: derived sizes, the basis
// tables packed into registers, and every site's coordinates checked to map back to its
// index (shared by the count-space and the pair-space samplers).
int geo_from(const chm_geom *gm, const int32_t *site_xyz, Geo &g) {
    g = Geo{};
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
    g.rstride = (g.nrow_e + 31) & ~31;
    g.nrow_tot = g.nsp * g.rstride;
    g.nrow_w = g.nrow_tot / 2;   // rstride is even
    g.nkw = gm->nkw;
    g.nwin = gm->nwin;
    g.ncol_e = gm->nkw * gm->nwin * 2;
    g.ncol_tot = g.nsp * g.ncol_e;
    g.nct = (g.ncol_e + NCW - 1) / NCW;
    g.nkb = (g.nrow_e + 31) / 32;
    g.nbasis = gm->nbasis;
    for (int d = 0; d < 3; ++d) {
        g.box[d] = gm->box[d];
        g.ncell[d] = gm->ncell[d];
    }
    for (int q = 0; q < 8; ++q) g.bpar[q] = gm->basis_of_parity[q];
    CHM_REQUIRE(g.nbasis <= 4, CHM_E_ARG, "at most 4 sites per cell");
    for (int b = 0; b < g.nbasis; ++b) {
        int p = -1;
        for (int q = 0; q < 8; ++q)
            if (g.bpar[q] == b) p = q;
        CHM_REQUIRE(p >= 0, CHM_E_ARG, "basis site without a parity pattern");
        g.bxyz[3 * b] = p & 1;
        g.bxyz[3 * b + 1] = (p >> 1) & 1;
        g.bxyz[3 * b + 2] = (p >> 2) & 1;
    }
    g.bpar4 = 0;
    g.bxyz3 = 0;
    for (int q = 0; q < 8; ++q) {
        CHM_REQUIRE(g.bpar[q] >= -1 && g.bpar[q] < g.nbasis, CHM_E_ARG, "basis_of_parity out of range");
        g.bpar4 |= (uint32_t)(g.bpar[q] & 15) << (4 * q);
    }
    for (int q = 0; q < 3 * g.nbasis; ++q) g.bxyz3 |= (uint32_t)g.bxyz[q] << q;
    CHM_REQUIRE(g.natoms < (1 << 20), CHM_E_ARG, "at most 2^20 sites");
    CHM_REQUIRE(g.nbasis >= 1 && g.ncell[1] >= 1 && g.ncell[2] >= 1, CHM_E_ARG, "empty cell grid");
    g.rnb = 1.0f / (float)g.nbasis;
    g.rnc1 = 1.0f / (float)g.ncell[1];
    g.rnc2 = 1.0f / (float)g.ncell[2];
    for (int q = 0; q < g.natoms; ++q) {
        int v = site_xyz[q];
        int x = v & 1023, y = (v >> 10) & 1023, z = (v >> 20) & 1023;
        CHM_REQUIRE(x < g.box[0] && y < g.box[1] && z < g.box[2], CHM_E_ARG, "site outside box");
        int b = g.bpar[(x & 1) | ((y & 1) << 1) | ((z & 1) << 2)];
        int cell = ((x >> 1) * g.ncell[1] + (y >> 1)) * g.ncell[2] + (z >> 1);
        CHM_REQUIRE(b >= 0 && cell * g.nbasis + b == q, CHM_E_ARG,
                    "site coordinates do not map back to their index");
    }
    return CHM_OK;
}

// Every multiple-scattering stencil record in range: slot count, weight, entry offset
// within the absorber's block, offsets within the box, absorber-role records first.
int check_stencil(const Geo &g, const int8_t *stencil) {
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
    return CHM_OK;
}
}  // namespace chm_smp

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
    Geo g;
    {
        int rc0 = geo_from(gm, site_xyz, g);
        if (rc0) return rc0;
    }
    // host-side validation of every index a kernel will follow
    for (long q = 0; q < (long)g.natoms * g.ncoord; ++q) {
        CHM_REQUIRE(nbr_j[q] >= 0 && nbr_j[q] < g.natoms, CHM_E_ARG, "neighbour out of range");
        CHM_REQUIRE(nbr_shell[q] >= 0 && nbr_shell[q] < g.nshell, CHM_E_ARG, "shell out of range");
    }
    // the neighbour offsets of site 0 form a stencil that must reproduce every site's list
    std::vector<int> ss(g.ncoord), nbr_st;
    int r2max = 0;
    {
        auto xyz_of = [&](int q, int &x, int &y, int &z) {
            int v = site_xyz[q];
            x = v & 1023;
            y = (v >> 10) & 1023;
            z = (v >> 20) & 1023;
        };
        auto centred = [](int d, int n) {
            d %= n;
            if (d > n / 2) d -= n;
            if (d <= -n / 2) d += n;
            return d;
        };
        int x0, y0, z0;
        xyz_of(0, x0, y0, z0);
        std::vector<int> off(3 * g.ncoord), shl(g.ncoord);
        for (int o = 0; o < g.ncoord; ++o) {
            int x, y, z;
            xyz_of(nbr_j[o], x, y, z);
            off[3 * o] = centred(x - x0, g.box[0]);
            off[3 * o + 1] = centred(y - y0, g.box[1]);
            off[3 * o + 2] = centred(z - z0, g.box[2]);
            shl[o] = nbr_shell[o];
            for (int d = 0; d < 3; ++d)
                CHM_REQUIRE(off[3 * o + d] > -128 && off[3 * o + d] < 128 && 2 * std::abs(off[3 * o + d]) < g.box[d],
                            CHM_E_ARG, "neighbour offset too large for the box");
            ss[o] = (off[3 * o] & 0xFF) | ((off[3 * o + 1] & 0xFF) << 8) |
                    ((off[3 * o + 2] & 0xFF) << 16) | ((shl[o] & 0xFF) << 24);
        }
        for (int o = 0; o < g.ncoord; ++o)
            r2max = std::max(r2max, off[3 * o] * off[3 * o] + off[3 * o + 1] * off[3 * o + 1] +
                                        off[3 * o + 2] * off[3 * o + 2]);
        std::vector<long long> want(g.ncoord), got(g.ncoord);
        nbr_st.resize((size_t)g.natoms * g.ncoord);
        for (int q = 0; q < g.natoms; ++q) {
            int x, y, z;
            xyz_of(q, x, y, z);
            for (int o = 0; o < g.ncoord; ++o) {
                int X = ((x + off[3 * o]) % g.box[0] + g.box[0]) % g.box[0];
                int Y = ((y + off[3 * o + 1]) % g.box[1] + g.box[1]) % g.box[1];
                int Z = ((z + off[3 * o + 2]) % g.box[2] + g.box[2]) % g.box[2];
                int b = g.bpar[(X & 1) | ((Y & 1) << 1) | ((Z & 1) << 2)];
                CHM_REQUIRE(b >= 0, CHM_E_ARG, "neighbour stencil leaves the lattice");
                int site = (((X >> 1) * g.ncell[1] + (Y >> 1)) * g.ncell[2] + (Z >> 1)) * g.nbasis + b;
                got[o] = (long long)site * 256 + shl[o];
                nbr_st[(size_t)q * g.ncoord + o] = site;
                want[o] = (long long)nbr_j[(long)q * g.ncoord + o] * 256 + nbr_shell[(long)q * g.ncoord + o];
            }
            std::sort(got.begin(), got.end());
            std::sort(want.begin(), want.end());
            CHM_REQUIRE(got == want, CHM_E_ARG,
                        "the neighbour list is not translation invariant (a Bravais lattice is required)");
        }
    }
    {
        int rc0 = check_stencil(g, stencil);
        if (rc0) return rc0;
    }
    // largest squared offset a gather visits (neighbours and path sites)
    int gr2 = r2max;
    for (int q = 0; q < g.nstencil; ++q) {
        const int8_t *rb = stencil + 16 * q;
        for (int z = 0; z <= rb[12]; ++z)
            gr2 = std::max(gr2, rb[3 * z] * rb[3 * z] + rb[3 * z + 1] * rb[3 * z + 1] + rb[3 * z + 2] * rb[3 * z + 2]);
    }
    CHM_HIP(hipSetDevice(ctx->device));
    auto *S = new chm_sampler();
    S->ctx = ctx;
    S->g = g;
    S->nrep = nrep;
    S->nk = nk;
    S->profile = profile_scale ? 1 : 0;
    S->nfit = (double)g.ncol_tot;  // nsp*nkw*nwin complex points, Re and Im counted
    S->lds_k1 = sizeof(int) * ((g.nrow_tot + 3) & ~3) + (size_t)g.natoms_pad;
    S->lds_pair = sizeof(int) * 3 * ((g.nrow_tot + 3) & ~3) + 2 * (size_t)g.natoms_pad;
    S->gr2 = gr2;
    {
        // workgroups per replica for the gather: ~3 items per lane, at most 16
        const int items = 2 * g.ncoord + 2 * g.nstencil;
        S->nsplit = std::max(1, std::min(16, (items + 767) / 768));
    }
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
    // per-column weight 1/sigma^2 (Re and Im share a point's sigma) and, per edge,
    // Q0 = sum over the window of t^2 w
    std::vector<double> wc(g.ncol_tot), q0(2 * chm::kMaxNsp, 0.0);   // padded: read as one block
    for (int q = 0; q < g.ncol_tot / 2; ++q) {
        if (!(sigma[q] > 0.0)) return fail((chm::set_error("sigma must be positive"), CHM_E_ARG));
        wc[2 * q] = wc[2 * q + 1] = 1.0 / (sigma[q] * sigma[q]);
    }
    for (int q = 0; q < g.ncol_tot; ++q) {
        q0[2 * (q / g.ncol_e)] += target[q] * target[q];
        q0[2 * (q / g.ncol_e) + 1] += target[q] * target[q] * wc[q];
    }
    if ((rc = up(S->ss_sten, ss.data(), sizeof(int) * g.ncoord)) ||
        (rc = up(S->nbr_st, nbr_st.data(), sizeof(int) * nbr_st.size())) ||
        (rc = up(S->site_xyz, site_xyz, sizeof(int) * g.natoms)) ||
        (rc = up(S->nbr_j, nbr_j, sizeof(int) * g.natoms * g.ncoord)) ||
        (rc = up(S->nbr_sh, nbr_shell, sizeof(int) * g.natoms * g.ncoord)) ||
        (rc = up(S->stencil, g.nstencil ? (const void *)stencil : (const void *)"", 16 * (size_t)g.nstencil)) ||
        (rc = up(S->basisR, basis_R, sizeof(double) * (size_t)g.nsp * g.nrow_e * g.ncol_e)) ||
        (rc = up(S->basisK, basis_k, sizeof(double) * (size_t)g.nsp * g.nrow_e * nk)) ||
        (rc = up(S->target, target, sizeof(double) * g.ncol_tot)) ||
        (rc = up(S->wcol, wc.data(), sizeof(double) * wc.size())) ||
        (rc = up(S->q0, q0.data(), sizeof(double) * q0.size())))
        return fail(rc);
    {
        long nb = (long)g.nsp * g.nct * g.nkb * 3 * 64 * 8;
        if ((rc = S->Tb.ensure(sizeof(unsigned short) * nb))) return fail(rc);
        long nt = (long)g.nsp * g.nct * g.nkb * 64 * 8;
        k_split_bf16<<<dim3((unsigned)((nt + 255) / 256)), 256, 0, ctx->stream>>>(
            g, S->basisR.as<double>(), S->Tb.as<unsigned short>());
        hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            chm::set_error(std::string("split_bf16: ") + hipGetErrorString(le));
            return fail(CHM_E_HIP);
        }
    }
    size_t nr = nrep;
    if ((rc = S->species.ensure(nr * g.natoms_pad)) || (rc = S->best_sp.ensure(nr * g.natoms_pad)) ||
        (rc = S->counts.ensure(sizeof(int) * nr * g.nrow_tot)) ||
        (rc = S->dcounts.ensure(sizeof(int) * 2 * nr * g.nrow_tot)) ||
        (rc = S->c2.ensure(sizeof(double) * 2 * nr)) || (rc = S->best_c2.ensure(sizeof(double) * nr)) ||
        (rc = S->best_snap.ensure(sizeof(double) * nr)) ||
        (rc = S->rung_of.ensure(sizeof(int) * nr)) || (rc = S->slot_of.ensure(sizeof(int) * nr)) ||
        (rc = S->beta.ensure(sizeof(double) * nr)) || (rc = S->beta_slot.ensure(sizeof(double) * nr)) ||
        (rc = S->stepc.ensure(sizeof(uint64_t) * nr)) ||
        (rc = S->iex.ensure(sizeof(uint64_t))) || (rc = S->prop.ensure(sizeof(int) * 2 * kPropInts * nr)) ||
        (rc = S->mtrial.ensure(sizeof(double) * nr * g.ncol_tot)) ||
        (rc = S->mcur.ensure(sizeof(double) * nr * g.ncol_tot)) ||
        (rc = S->partials.ensure(sizeof(double) * nr * g.nsp * g.nct * 4)) ||
        (rc = S->anchor.ensure(sizeof(int) * nr)) ||
        (rc = S->nanchor.ensure(sizeof(long long) * nr)) ||
        (rc = S->exch_att.ensure(sizeof(long long) * nr)) ||
        (rc = S->exch_acc.ensure(sizeof(long long) * nr)) ||
        (rc = S->acc_rung.ensure(sizeof(long long) * nr)) ||
        (rc = S->pdc.ensure(sizeof(int) * 6 * nr * g.nrow_tot)) ||
        (rc = S->pprop.ensure(sizeof(int) * 6 * kPropInts * nr)) ||
        (rc = S->pmt.ensure(sizeof(double) * 3 * nr * g.ncol_tot)) ||
        (rc = S->ppart.ensure(sizeof(double) * 3 * nr * g.nsp * g.nct * 4)) ||
        (rc = S->pc2.ensure(sizeof(double) * 2 * nr)) ||
        (rc = S->draws.ensure(sizeof(int) * kDrawInts * nr)))
        return fail(rc);
    hipFuncSetAttribute((const void *)k_gemm_abs_f64, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds_gemm(g));
    hipFuncSetAttribute((const void *)k_slot_step,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)S->lds_k1);
    hipFuncSetAttribute((const void *)k_full_counts, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)S->lds_k1);
    hipFuncSetAttribute((const void *)k_slot_pair, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)S->lds_pair);
    if (S->lds_pair > 160 * 1024) S->pairs = 0;   // single steps only
    if (lds_gemm(g) > 160 * 1024 || S->lds_k1 > 160 * 1024)
        return fail((chm::set_error("problem too large for LDS staging"), CHM_E_ARG));
    S->lp.r2max = r2max;
    if ((rc = lp_prepare(S))) return fail(rc);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        chm::set_error(std::string("sampler upload: ") + hipGetErrorString(e));
        return fail(CHM_E_HIP);
    }
    chm::ctx_retain(ctx);
    *out = S;
    return CHM_OK;
}

extern "C" int chm_sampler_destroy(chm_sampler *S) {
    if (!S) return CHM_OK;
    chm_ctx *ctx = S->ctx;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    drop_graph(S);
    delete S;
    chm::ctx_release(ctx);
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
    S->rec_base = S->hstep;
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
    CHM_HIP(hipMemcpyAsync(S->beta_slot.p, beta, sizeof(double) * nrep, hipMemcpyHostToDevice, st));
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
    CHM_HIP(hipMemsetAsync(S->anchor.p, 0, S->anchor.bytes, st));
    CHM_HIP(hipMemsetAsync(S->c2.p, 0, S->c2.bytes, st));
    CHM_HIP(hipMemsetAsync(S->best_c2.p, 0, S->best_c2.bytes, st));
    CHM_HIP(hipMemsetAsync(S->best_snap.p, 0, S->best_snap.bytes, st));
    CHM_HIP(hipMemsetAsync(S->nanchor.p, 0, S->nanchor.bytes, st));
    CHM_HIP(hipMemsetAsync(S->pdc.p, 0, S->pdc.bytes, st));
    CHM_HIP(hipMemsetAsync(S->pprop.p, 0, S->pprop.bytes, st));
    CHM_HIP(hipMemsetAsync(S->pc2.p, 0, S->pc2.bytes, st));
    if (S->log_cap) CHM_HIP(hipMemsetAsync(S->log.p, 0xff, S->log.bytes, st));
    S->seed = seed;
    S->hstep = 0;
    S->rec_base = 0;
    S->lp.cache_ok = false;   // new species: the learned proposal's per-site sums are stale
    int nsplit = full_counts_split(g, nrep);
    k_full_counts<<<dim3(nsplit, nrep), 256, S->lds_k1, st>>>(
        g, S->species.as<uint8_t>(), S->site_xyz.as<int>(), S->nbr_j.as<int>(), S->nbr_sh.as<int>(),
        S->stencil.as<int4>(), S->counts.as<int>());
    CHM_LAUNCH_CHECK();
    int rc = launch_gemm_abs(S, drow(S, 0), 0);    // both rows are zero here
    if (rc) return rc;
    if ((rc = launch_slot(S, 2, 0, 0))) return rc;
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
    if (S->lp.mode && (rc = lp_ensure_cache(S))) return rc;
    if (E > 0 && S->graph_steps != E) {
        // one exchange period per graph: propose, E x (transform, accept [+ propose]),
        // exchange.  The count-change rows alternate by step parity, so a period that
        // starts after an odd number of steps (odd E) has its own graph.
        drop_graph(S);
        for (int par = 0; par < ((E & 1) ? 2 : 1) && !rc; ++par) {
            CHM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = launch_segment(S, par, E);
            if (!rc) rc = launch_exchange(S, par + E);
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
            S->graph[par] = gr;
            CHM_HIP(hipGraphInstantiate(&S->gexec[par], gr, nullptr, nullptr, 0));
        }
        if (rc) return rc;
        S->graph_steps = E;
    }
    long long left = nsteps;
    while (left > 0) {
        if (E > 0 && S->hstep % E == 0 && left >= E) {
            CHM_HIP(hipGraphLaunch(S->gexec[(E & 1) ? (S->hstep & 1) : 0], st));
            S->hstep += E;
            left -= E;
        } else {
            long long L = left;
            if (E > 0) L = std::min<long long>(left, E - S->hstep % E);
            if ((rc = launch_segment(S, S->hstep, (int)L))) return rc;
            S->hstep += L;
            left -= L;
            if (E > 0 && S->hstep % E == 0 && (rc = launch_exchange(S, S->hstep))) return rc;
        }
    }
    // no host sync: every read and every state change synchronises the stream first
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
            CHM_HIP(hipMemcpy(v.data(), what == 1 ? (void *)c2row(S, S->hstep) : S->best_c2.p, 8 * nrep,
                              hipMemcpyDeviceToHost));
            double *o = static_cast<double *>(host);
            for (int r = 0; r < nrep; ++r) o[r] = v[slot_of[r]];
            return CHM_OK;
        }
        case 2: {
            const int nrc = g.nsp * g.nrow_e;
            if ((rc = need((int64_t)nrep * nrc * 4))) return rc;
            std::vector<int> v((size_t)nrep * g.nrow_tot);
            CHM_HIP(hipMemcpy(v.data(), S->counts.p, 4 * v.size(), hipMemcpyDeviceToHost));
            int *o = static_cast<int *>(host);
            for (int r = 0; r < nrep; ++r)
                for (int e = 0; e < g.nsp; ++e)
                    memcpy(o + (size_t)r * nrc + (size_t)e * g.nrow_e,
                           v.data() + (size_t)slot_of[r] * g.nrow_tot + (size_t)e * g.rstride,
                           4 * g.nrow_e);
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
            rc = chm::gemm_f64(st, g.nsp, nrep, S->nk, g.nrow_e, cf, g.nrow_tot, g.rstride,
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
        case 12:
        case 13: {
            int64_t n = (int64_t)8 * nrep * g.ncol_tot;
            if ((rc = need(n))) return rc;
            CHM_HIP(hipMemcpy(host, what == 12 ? S->mtrial.p : S->mcur.p, n, hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 14: {
            if ((rc = need((int64_t)8 * nrep))) return rc;
            CHM_HIP(hipMemcpy(host, S->nanchor.p, 8 * nrep, hipMemcpyDeviceToHost));
            return CHM_OK;
        }
        case 15: {
            // the per-site layer-1 sums (of the committed configuration)
            CHM_REQUIRE(S->lp.cache_ok, CHM_E_STATE, "the learned proposal's per-site sums are not built");
            const int64_t n = (int64_t)4 * nrep * 2 * g.natoms * 256;
            if ((rc = need(n))) return rc;
            CHM_HIP(hipMemcpy(host, S->lp.cache.p, n, hipMemcpyDeviceToHost));
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
    int nsplit = full_counts_split(g, nconf);
    k_full_counts<<<dim3(nsplit, nconf), 256, S->lds_k1, st>>>(
        g, sp.as<uint8_t>(), S->site_xyz.as<int>(), S->nbr_j.as<int>(), S->nbr_sh.as<int>(),
        S->stencil.as<int4>(), cn.as<int>());
    CHM_LAUNCH_CHECK();
    std::vector<int> v((size_t)nconf * g.nrow_tot);
    CHM_HIP(hipMemcpyAsync(v.data(), cn.p, sizeof(int) * v.size(), hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    const int nrc = g.nsp * g.nrow_e;
    for (int c = 0; c < nconf; ++c)
        for (int e = 0; e < g.nsp; ++e)
            memcpy(counts + (size_t)c * nrc + (size_t)e * g.nrow_e,
                   v.data() + (size_t)c * g.nrow_tot + (size_t)e * g.rstride, 4 * g.nrow_e);
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
    k_eval_rows<<<dim3(nrows), 256, 0, st>>>(g, m.as<double>(), S->target.as<double>(),
                                             S->wcol.as<double>(), S->profile, S->nfit,
                                             o.as<double>());
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(c2, o.p, sizeof(double) * nrows, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_sampler_exchange(chm_sampler *S) {
    CHM_REQUIRE(S && S->inited, CHM_E_STATE, "init_state first");
    CHM_HIP(hipSetDevice(S->ctx->device));
    int rc = launch_exchange(S, S->hstep);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    return CHM_OK;
}

extern "C" int chm_sampler_set_beta(chm_sampler *S, const double *beta) {
    CHM_REQUIRE(S && beta, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(S->inited, CHM_E_STATE, "init_state first");
    const int nrep = S->nrep;
    for (int r = 0; r < nrep; ++r)
        CHM_REQUIRE(std::isfinite(beta[r]) && beta[r] > 0.0, CHM_E_ARG, "beta must be finite and > 0");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    CHM_HIP(hipStreamSynchronize(st));
    // The temperatures travel with the exchange (slot s holds rung rung_of[s]); the
    // new ladder is written per rung and re-gathered into slot order.
    std::vector<int> rung_of(nrep);
    CHM_HIP(hipMemcpy(rung_of.data(), S->rung_of.p, sizeof(int) * nrep, hipMemcpyDeviceToHost));
    std::vector<double> bslot(nrep);
    for (int s = 0; s < nrep; ++s) {
        CHM_REQUIRE(rung_of[s] >= 0 && rung_of[s] < nrep, CHM_E_STATE, "corrupt rung map");
        bslot[s] = beta[rung_of[s]];
    }
    CHM_HIP(hipMemcpyAsync(S->beta.p, beta, sizeof(double) * nrep, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(S->beta_slot.p, bslot.data(), sizeof(double) * nrep, hipMemcpyHostToDevice, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_sampler_chik(chm_sampler *S, int nconf, const int32_t *counts, double *chi) {
    CHM_REQUIRE(S && nconf > 0 && counts && chi, CHM_E_ARG, "bad arguments");
    const Geo &g = S->g;
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    chm::DevBuf cf, ck;
    const int nrc = g.nsp * g.nrow_e;
    std::vector<double> h((size_t)nconf * nrc);
    for (size_t q = 0; q < h.size(); ++q) h[q] = (double)counts[q];
    int rc;
    if ((rc = cf.ensure(sizeof(double) * h.size())) ||
        (rc = ck.ensure(sizeof(double) * (size_t)nconf * g.nsp * S->nk)))
        return rc;
    CHM_HIP(hipMemcpyAsync(cf.p, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice, st));
    rc = chm::gemm_f64(st, g.nsp, nconf, S->nk, g.nrow_e, cf.as<double>(), nrc, g.nrow_e,
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
    hipEvent_t e0, e1;
    CHM_HIP(hipEventCreate(&e0));
    CHM_HIP(hipEventCreate(&e1));
    float t = 0.f;
    long long h = S->hstep;
    if (use_pairs(S, 3)) {
        // paired rounds: the three-trial transform repeated, then real rounds; times
        // are per step (half a round)
        int rc = launch_slot_pair(S, 0, 3, -1, c2row(S, h), pc2_row(S, 0), 0, h + 1);   // round 0's rows
        CHM_HIP(hipEventRecord(e0, st));
        for (int i = 0; i < iters && !rc; ++i) rc = launch_gemm_pair(S, 0, 0);
        CHM_HIP(hipEventRecord(e1, st));
        if (rc) return rc;
        CHM_HIP(hipEventSynchronize(e1));
        CHM_HIP(hipEventElapsedTime(&t, e0, e1));
        ms[1] = t / iters / 2;
        CHM_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < iters && !rc; ++k) {
            rc = launch_gemm_pair(S, k, 2);
            if (!rc) rc = launch_slot_pair(S, 1, 3, k, k ? pc2_row(S, k - 1) : c2row(S, h), pc2_row(S, k), 0, 0);
        }
        CHM_HIP(hipEventRecord(e1, st));
        if (rc) return rc;
        CHM_HIP(hipEventSynchronize(e1));
        CHM_HIP(hipEventElapsedTime(&t, e0, e1));
        ms[3] = t / iters / 2;
        ms[2] = ms[3] - ms[1];
        ms[0] = 0.0;
        // settle: the last proposed round, ending in the canonical chi^2 row
        if (!rc) rc = launch_gemm_pair(S, iters, 2);
        if (!rc) rc = launch_slot_pair(S, 1, 0, iters, pc2_row(S, iters - 1), c2row(S, h + 2 * (iters + 1)), 0, 0);
        if (rc) return rc;
        S->hstep = h + 2 * (iters + 1);
        CHM_HIP(hipStreamSynchronize(st));
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return CHM_OK;
    }
    int rc = launch_propose(S, h);           // step h+1 pending
    // the trial transform of the pending step, repeated (no counter advance)
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters && !rc; ++i) rc = launch_gemm(S, drow(S, h + 1), 0);
    CHM_HIP(hipEventRecord(e1, st));
    if (rc) return rc;
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[1] = t / iters;
    // real steps of the chain: transform, then accept + propose
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters && !rc; ++i) {
        const long long n = h + 1 + i;
        rc = launch_trial(S, n, true);
        if (!rc) rc = launch_accept(S, n, 1);
    }
    CHM_HIP(hipEventRecord(e1, st));
    if (rc) return rc;
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[3] = t / iters;
    ms[2] = ms[3] - ms[1];                   // accept + propose + gather per step
    ms[0] = 0.0;
    const long long n = h + 1 + iters;       // settle: transform and accept the last proposal
    if (!rc) rc = launch_trial(S, n, false);
    if (!rc) rc = launch_accept(S, n, 0);
    if (rc) return rc;
    S->hstep = n;
    CHM_HIP(hipStreamSynchronize(st));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return CHM_OK;
}

extern "C" int chm_sampler_set_precision(chm_sampler *S, int incremental, int reanchor_every) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    const int mode = incremental & 15, acc32 = (incremental & 16) ? 1 : 0;
    CHM_REQUIRE((incremental & ~31) == 0 && mode >= 0 && mode <= 2 && (mode > 0 || !acc32), CHM_E_ARG,
                "precision mode must be 0, 1 or 2 (+16 for an FP32 accumulator with 1, 2)");
    CHM_REQUIRE(reanchor_every >= 0, CHM_E_ARG, "negative re-anchor period");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    if (mode == 2 && !S->basisR32.p) {
        const Geo &g = S->g;
        const long n = (long)g.nsp * g.nrow_e * g.ncol_e;
        int rc = S->basisR32.ensure(sizeof(float) * n);
        if (rc) return rc;
        k_to_f32<<<(unsigned)((n + 255) / 256), 256, 0, S->ctx->stream>>>(n, S->basisR.as<double>(),
                                                                         S->basisR32.as<float>());
        CHM_LAUNCH_CHECK();
        CHM_HIP(hipStreamSynchronize(S->ctx->stream));
        hipFuncSetAttribute((const void *)k_gemm_delta_f32, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)std::max(sizeof(float) * (size_t)g.nrow_e * (RB + 1),
                                          sizeof(float) * (size_t)KS * RB * NCW));
    }
    S->incremental = mode;
    S->acc32 = acc32;
    S->reanchor_every = reanchor_every;
    S->inited = false;   // the accumulator is built by the next init
    return CHM_OK;
}

extern "C" int chm_sampler_set_tuning(chm_sampler *S, int nsplit) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    CHM_REQUIRE(nsplit >= 1 && nsplit <= 64, CHM_E_ARG, "nsplit must be 1..64");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    S->nsplit = nsplit;
    return CHM_OK;
}

extern "C" int chm_sampler_get_tuning(chm_sampler *S, int *nsplit) {
    CHM_REQUIRE(S && nsplit, CHM_E_ARG, "bad arguments");
    *nsplit = S->nsplit;
    return CHM_OK;
}

extern "C" int chm_sampler_set_pairs(chm_sampler *S, int on) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    CHM_REQUIRE(!on || S->lds_pair <= 160 * 1024, CHM_E_ARG, "problem too large for paired steps");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    S->pairs = on ? 1 : 0;
    return CHM_OK;
}

// ======================================================================== run marks
namespace {

// One wavefront: the smallest best chi^2 over the slots, the accepted-move total and the
// device wall clock, for RMCSampler.run's history without a host round trip.
__global__ __launch_bounds__(64) void k_mark(int nrep, const double *__restrict__ best_c2,
                                             const long long *__restrict__ acc_rung,
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

extern "C" int chm_sampler_reserve_marks(chm_sampler *S, int n) {
    CHM_REQUIRE(S && n >= 0, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    if (n > S->mark_cap) {
        CHM_HIP(hipStreamSynchronize(S->ctx->stream));   // a mark may still be in flight
        int rc = S->marks.ensure(sizeof(double) * 3 * (size_t)n);
        if (rc) return rc;
        S->mark_cap = n;
    }
    return CHM_OK;
}

extern "C" int chm_sampler_mark(chm_sampler *S, int idx) {
    CHM_REQUIRE(S && S->inited, CHM_E_STATE, "init_state first");
    CHM_REQUIRE(idx >= 0 && idx < S->mark_cap, CHM_E_ARG, "mark index beyond the reserved capacity");
    CHM_HIP(hipSetDevice(S->ctx->device));
    k_mark<<<1, 64, 0, S->ctx->stream>>>(S->nrep, S->best_c2.as<double>(), S->acc_rung.as<long long>(),
                                         S->marks.as<double>() + 3 * (size_t)idx);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

extern "C" int chm_sampler_read_marks(chm_sampler *S, int n, double *out) {
    CHM_REQUIRE(S && out && n >= 1 && n <= S->mark_cap, CHM_E_ARG, "bad mark count");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    CHM_HIP(hipMemcpy(out, S->marks.p, sizeof(double) * 3 * (size_t)n, hipMemcpyDeviceToHost));
    int khz = 0;
    CHM_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, S->ctx->device));
    CHM_REQUIRE(khz > 0, CHM_E_HIP, "device wall clock rate unavailable");
    const double t0 = out[2];
    for (int q = 0; q < n; ++q) out[3 * q + 2] = (out[3 * q + 2] - t0) / (1e3 * khz);
    return CHM_OK;
}

// ======================================================================== checkpoint
// The complete chain state between two run() calls: the next call proposes from it (the
// proposal and count-change rows are rebuilt at the start of every segment), so restoring
// these buffers and the step count continues the chain bit for bit.
namespace {

constexpr uint32_t kStateMagic = 0x53534D43u;   // "CMSS"
constexpr uint32_t kStateVersion = 1;
constexpr int kStateSect = 17;

struct StateHeader {
    uint32_t magic, version;
    int32_t natoms, nsp, nrow_tot, ncol_tot, nrep, natoms_pad, incremental, nsect;
    int64_t hstep;
    uint64_t seed;
    int64_t off[kStateSect], bytes[kStateSect];
};

// species, counts, chi^2 (both step-parity rows), running chi(R), best chi^2, best
// species, rung_of, slot_of, beta (rung), beta (slot), step counters, exchange counter,
// accepted per rung, accepted since re-anchor, re-anchors, exchange attempts / accepts
void state_sections(chm_sampler *S, chm::DevBuf *b[kStateSect], size_t n[kStateSect]) {
    const size_t r = S->nrep;
    const Geo &g = S->g;
    chm::DevBuf *bb[kStateSect] = {&S->species, &S->counts,   &S->c2,       &S->mcur,    &S->best_c2,
                                   &S->best_sp, &S->rung_of,  &S->slot_of,  &S->beta,    &S->beta_slot,
                                   &S->stepc,   &S->iex,      &S->acc_rung, &S->anchor,  &S->nanchor,
                                   &S->exch_att, &S->exch_acc};
    const size_t nn[kStateSect] = {r * g.natoms_pad, 4 * r * g.nrow_tot, 16 * r, 8 * r * g.ncol_tot, 8 * r,
                                   r * g.natoms_pad, 4 * r, 4 * r, 8 * r, 8 * r,
                                   8 * r, 8, 8 * r, 4 * r, 8 * r,
                                   8 * r, 8 * r};
    for (int q = 0; q < kStateSect; ++q) {
        b[q] = bb[q];
        n[q] = nn[q];
    }
}

StateHeader state_header(chm_sampler *S) {
    StateHeader h{};
    h.magic = kStateMagic;
    h.version = kStateVersion;
    h.natoms = S->g.natoms;
    h.nsp = S->g.nsp;
    h.nrow_tot = S->g.nrow_tot;
    h.ncol_tot = S->g.ncol_tot;
    h.nrep = S->nrep;
    h.natoms_pad = S->g.natoms_pad;
    h.incremental = S->incremental;
    h.nsect = kStateSect;
    h.hstep = S->hstep;
    h.seed = S->seed;
    chm::DevBuf *b[kStateSect];
    size_t n[kStateSect];
    state_sections(S, b, n);
    int64_t off = (sizeof(StateHeader) + 7) & ~(size_t)7;
    for (int q = 0; q < kStateSect; ++q) {
        h.off[q] = off;
        h.bytes[q] = (int64_t)n[q];
        off += ((int64_t)n[q] + 7) & ~(int64_t)7;
    }
    return h;
}

int64_t state_bytes(const StateHeader &h) {
    return h.off[kStateSect - 1] + ((h.bytes[kStateSect - 1] + 7) & ~(int64_t)7);
}

}  // namespace

extern "C" int chm_sampler_state_size(chm_sampler *S, int64_t *nbytes) {
    CHM_REQUIRE(S && nbytes, CHM_E_ARG, "bad arguments");
    *nbytes = state_bytes(state_header(S));
    return CHM_OK;
}

extern "C" int chm_sampler_export(chm_sampler *S, void *host, int64_t nbytes) {
    CHM_REQUIRE(S && host, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(S->inited, CHM_E_STATE, "init_state first");
    const StateHeader h = state_header(S);
    CHM_REQUIRE(nbytes >= state_bytes(h), CHM_E_ARG, "host buffer too small");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    char *o = static_cast<char *>(host);
    std::memset(o, 0, (size_t)state_bytes(h));
    std::memcpy(o, &h, sizeof(h));
    chm::DevBuf *b[kStateSect];
    size_t n[kStateSect];
    state_sections(S, b, n);
    for (int q = 0; q < kStateSect; ++q)
        CHM_HIP(hipMemcpy(o + h.off[q], b[q]->p, n[q], hipMemcpyDeviceToHost));
    return CHM_OK;
}

extern "C" int chm_sampler_import(chm_sampler *S, const void *host, int64_t nbytes) {
    CHM_REQUIRE(S && host, CHM_E_ARG, "bad arguments");
    const StateHeader want = state_header(S);
    CHM_REQUIRE(nbytes >= (int64_t)sizeof(StateHeader), CHM_E_ARG, "state blob too short");
    StateHeader h;
    std::memcpy(&h, host, sizeof(h));
    CHM_REQUIRE(h.magic == kStateMagic && h.version == kStateVersion, CHM_E_ARG,
                "not a sampler state blob of this version");
    CHM_REQUIRE(h.natoms == want.natoms && h.nsp == want.nsp && h.nrow_tot == want.nrow_tot &&
                    h.ncol_tot == want.ncol_tot && h.nrep == want.nrep && h.natoms_pad == want.natoms_pad &&
                    h.nsect == kStateSect,
                CHM_E_ARG, "state blob was saved from a sampler of another shape");
    CHM_REQUIRE(h.incremental == S->incremental, CHM_E_ARG,
                "state blob was saved in another transform precision (set_precision first)");
    for (int q = 0; q < kStateSect; ++q)
        CHM_REQUIRE(h.off[q] == want.off[q] && h.bytes[q] == want.bytes[q], CHM_E_ARG,
                    "state blob layout does not match");
    CHM_REQUIRE(nbytes >= state_bytes(want), CHM_E_ARG, "state blob too short");
    const char *in = static_cast<const char *>(host);
    const Geo &g = S->g;
    const int nrep = S->nrep;
    // host-side validation of every index and value a kernel will follow
    const uint8_t *sp = reinterpret_cast<const uint8_t *>(in + h.off[0]);
    const uint8_t *bsp = reinterpret_cast<const uint8_t *>(in + h.off[5]);
    for (long q = 0; q < (long)nrep * g.natoms_pad; ++q)
        CHM_REQUIRE(sp[q] < g.nsp && bsp[q] < g.nsp, CHM_E_ARG, "species out of range in state blob");
    const int32_t *rung_of = reinterpret_cast<const int32_t *>(in + h.off[6]);
    const int32_t *slot_of = reinterpret_cast<const int32_t *>(in + h.off[7]);
    for (int r = 0; r < nrep; ++r) {
        CHM_REQUIRE(rung_of[r] >= 0 && rung_of[r] < nrep && slot_of[r] >= 0 && slot_of[r] < nrep &&
                        slot_of[rung_of[r]] == r,
                    CHM_E_ARG, "rung / slot maps in state blob are not inverse permutations");
    }
    const double *beta = reinterpret_cast<const double *>(in + h.off[8]);
    for (int r = 0; r < nrep; ++r)
        CHM_REQUIRE(std::isfinite(beta[r]) && beta[r] > 0.0, CHM_E_ARG, "bad beta in state blob");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    CHM_HIP(hipStreamSynchronize(st));
    drop_graph(S);   // captured launches carry the recording base
    chm::DevBuf *b[kStateSect];
    size_t n[kStateSect];
    state_sections(S, b, n);
    for (int q = 0; q < kStateSect; ++q)
        CHM_HIP(hipMemcpy(b[q]->p, in + h.off[q], n[q], hipMemcpyHostToDevice));
    // no step is pending: the count-change and proposal rows are rebuilt by the next segment
    CHM_HIP(hipMemsetAsync(S->dcounts.p, 0, S->dcounts.bytes, st));
    CHM_HIP(hipMemsetAsync(S->pdc.p, 0, S->pdc.bytes, st));
    if (S->log_cap) CHM_HIP(hipMemsetAsync(S->log.p, 0xff, S->log.bytes, st));
    CHM_HIP(hipStreamSynchronize(st));
    S->hstep = h.hstep;
    S->seed = h.seed;
    S->rec_base = h.hstep;
    S->inited = true;
    S->lp.cache_ok = false;   // new species: the learned proposal's per-site sums are stale
    return CHM_OK;
}

// ======================================================================== batches
// Many independent problems (different supercells, edges, rows, columns, replica counts)
// stepped together: every kernel of a step is ONE launch for the whole batch, whose
// workgroups find their problem and tile (or replica slot) in descriptor tables built
// once; an exchange period of every problem is one captured hipGraph.  The problems keep
// their own state (each is a chm_sampler), so they can still be read and run alone.
namespace {

struct GemmProb {
    Geo g;
    int nrep;
    const int *counts;
    const int *dc;            // count-change rows of the launch's step (parity)
    const double *basisR;
    const bf16x8 *Tb;
    const float *basis32;
    const double *mcur;
    TileOut O;
};

// kMode 0 exact FP64, 1 split-BF16 delta, 2 FP32 delta.  KB: K blocks each wave holds in
// registers; a batch whose problems all have nkb <= 4 runs the KB = 1 instance (a quarter
// of the operand registers, so more tiles are in flight per CU).
template <int kMode, int KB = KBW>
__global__ __launch_bounds__(256) void k_gemm_grouped(const GemmProb *__restrict__ probs,
                                                       const int4 *__restrict__ tiles) {
    const int4 tl = tiles[blockIdx.x];
    const GemmProb &P = probs[tl.x];
    if (kMode == 0)
        gemm_abs_body(P.g, P.nrep, P.counts, P.dc, P.basisR, P.O, tl.y, tl.z, tl.w);
    else if (kMode == 1)
        gemm_delta_bf16_body<KB>(P.g, P.nrep, P.dc, P.Tb, P.mcur, P.O, tl.y, tl.z, tl.w);
    else
        gemm_delta_f32_body(P.g, P.nrep, P.dc, P.basis32, P.mcur, P.O, tl.y, tl.z, tl.w);
}

constexpr int kSlotKinds = 5;   // (accept, propose, re-anchor step): 010 110 100 111 101
constexpr int kKindAccept[kSlotKinds] = {0, 1, 1, 1, 1};
constexpr int kKindPropose[kSlotKinds] = {1, 1, 0, 1, 0};
constexpr int kKindReanchor[kSlotKinds] = {0, 0, 0, 1, 1};

struct SlotProb {
    Geo g;
    int nsplit;
    SlotArgs A[2 * kSlotKinds];   // [parity of the step][kind]
};

__global__ __launch_bounds__(256) void k_slot_grouped(const SlotProb *__restrict__ probs,
                                                       const int2 *__restrict__ slots, int variant) {
    const int2 ps = slots[blockIdx.y];
    const SlotProb &P = probs[ps.x];
    if ((int)blockIdx.x >= P.nsplit) return;
    slot_step_body(P.g, P.A[variant], ps.y, blockIdx.x, P.nsplit);
}

struct ExchProb {
    int nrep;
    const double *beta;
    const double *c2[2];
    int *slot_of, *rung_of;
    double *beta_slot;
    uint64_t *iex;
    long long *att, *acc;
    uint64_t seed;
    double nfit;
};

__global__ void k_exchange_grouped(const ExchProb *__restrict__ probs, int parity) {
    const ExchProb &P = probs[blockIdx.x];
    exchange_body(P.nrep, P.beta, P.c2[parity], P.slot_of, P.rung_of, P.beta_slot, P.iex, P.att, P.acc,
                  P.seed, P.nfit);
}

}  // namespace

struct chm_batch {
    chm_ctx *ctx = nullptr;
    std::vector<chm_sampler *> S;
    int mode = 0, anchor = 0, nslot = 0, nsplit_max = 1, ntile = 0, max_nkb = 0;
    int kb_wave = 0;   // split-BF16 transform: K blocks per wave (0: 1 when every nkb <= 4, else 4)
    size_t lds_gemm_max = 0, lds_slot_max = 0;
    chm::DevBuf gemm[2], gemm_nb, reanchor, tiles, slot, slots, exch;
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    int graph_steps = 0;
    long long hstep = 0;
};

namespace {

void batch_drop_graph(chm_batch *B) {
    for (int q = 0; q < 2; ++q) {
        if (B->gexec[q]) hipGraphExecDestroy(B->gexec[q]);
        if (B->graph[q]) hipGraphDestroy(B->graph[q]);
        B->gexec[q] = nullptr;
        B->graph[q] = nullptr;
    }
    B->graph_steps = 0;
}

int batch_gemm(chm_batch *B, const GemmProb *probs, int mode) {
    hipStream_t st = B->ctx->stream;
    const int4 *tiles = B->tiles.as<int4>();
    if (mode == 0)
        k_gemm_grouped<0><<<B->ntile, 256, B->lds_gemm_max, st>>>(probs, tiles);
    else if (mode == 1 && B->max_nkb <= 4 && B->kb_wave != KBW)
        k_gemm_grouped<1, 1><<<B->ntile, 256, 0, st>>>(probs, tiles);
    else if (mode == 1)
        k_gemm_grouped<1><<<B->ntile, 256, 0, st>>>(probs, tiles);
    else
        k_gemm_grouped<2><<<B->ntile, 256, B->lds_gemm_max, st>>>(probs, tiles);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int batch_slot(chm_batch *B, long long n, int kind) {
    const int variant = (int)(n & 1) * kSlotKinds + kind;
    k_slot_grouped<<<dim3(B->nsplit_max, B->nslot), 256, B->lds_slot_max, B->ctx->stream>>>(
        B->slot.as<SlotProb>(), B->slots.as<int2>(), variant);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int batch_exchange(chm_batch *B, long long n) {
    k_exchange_grouped<<<(unsigned)B->S.size(), 256, 0, B->ctx->stream>>>(B->exch.as<ExchProb>(),
                                                                          (int)(n & 1));
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

// as launch_segment, for every problem at once
int batch_segment(chm_batch *B, long long h, int L) {
    int rc = batch_slot(B, h, 0);
    if (!rc && B->anchor) rc = batch_gemm(B, B->reanchor.as<GemmProb>(), 0);
    for (int t = 0; t < L && !rc; ++t) {
        const long long n = h + 1 + t;
        rc = batch_gemm(B, B->gemm[n & 1].as<GemmProb>(), B->mode);
        const bool prop = t + 1 < L, re = B->anchor && t == 0;
        const int kind = re ? (prop ? 3 : 4) : (prop ? 1 : 2);
        if (!rc) rc = batch_slot(B, n, kind);
    }
    return rc;
}

}  // namespace

extern "C" int chm_batch_create(chm_ctx *ctx, int n, chm_sampler *const *samplers, chm_batch **out) {
    CHM_REQUIRE(ctx && samplers && out && n >= 1, CHM_E_ARG, "bad arguments");
    auto *B = new chm_batch();
    B->ctx = ctx;
    auto fail = [&](int code, const char *msg) {
        delete B;
        chm::set_error(msg);
        return code;
    };
    for (int p = 0; p < n; ++p) {
        chm_sampler *S = samplers[p];
        if (!S) return fail(CHM_E_ARG, "null sampler");
        if (S->ctx->device != ctx->device) return fail(CHM_E_ARG, "every sampler must live on the batch's device");
        if (!S->inited) return fail(CHM_E_STATE, "init every sampler first");
        if (S->lp.mode) return fail(CHM_E_ARG, "the learned proposal is not batched");
        if (p && (S->incremental != samplers[0]->incremental || S->hstep != samplers[0]->hstep))
            return fail(CHM_E_ARG, "batched samplers need one precision mode and one step count");
        B->S.push_back(S);
    }
    chm_sampler *S0 = samplers[0];
    B->mode = S0->incremental;
    B->hstep = S0->hstep;
    for (chm_sampler *S : B->S) B->anchor |= (S->incremental && S->reanchor_every > 0);
    CHM_HIP(hipSetDevice(ctx->device));
    // tiles of every problem's transform, slots of every problem's replicas
    std::vector<int4> tiles;
    std::vector<int2> slots;
    std::vector<GemmProb> gp[2], rp, gnb;
    std::vector<SlotProb> sp(n);
    std::vector<ExchProb> ep(n);
    for (int p = 0; p < n; ++p) {
        chm_sampler *S = B->S[p];
        const Geo &g = S->g;
        B->max_nkb = std::max(B->max_nkb, g.nkb);
        for (int e = 0; e < g.nsp; ++e)
            for (int rb = 0; rb < (S->nrep + RB - 1) / RB; ++rb)
                for (int ct = 0; ct < g.nct; ++ct) tiles.push_back(make_int4(p, ct, rb, e));
        for (int s = 0; s < S->nrep; ++s) slots.push_back(make_int2(p, s));
        B->lds_gemm_max = std::max(B->lds_gemm_max, lds_gemm(g));
        B->lds_slot_max = std::max(B->lds_slot_max, S->lds_k1);
        B->nsplit_max = std::max(B->nsplit_max, S->nsplit);
        for (int par = 0; par < 2; ++par) {
            GemmProb G{};
            G.g = g;
            G.nrep = S->nrep;
            G.counts = S->counts.as<int>();
            G.dc = drow(S, par);
            G.basisR = S->basisR.as<double>();
            G.Tb = reinterpret_cast<const bf16x8 *>(S->Tb.p);
            G.basis32 = S->basisR32.as<float>();
            G.mcur = S->mcur.as<double>();
            G.O = tile_out(S);
            G.O.bump = 1;
            gp[par].push_back(G);
            if (par == 1) {   // for timing: the transform without advancing the step counter
                G.O.bump = 0;
                gnb.push_back(G);
            }
        }
        GemmProb R{};
        R.g = g;
        R.nrep = S->nrep;
        R.counts = S->counts.as<int>();
        R.dc = nullptr;
        R.basisR = S->basisR.as<double>();
        R.O = tile_out(S);
        R.O.anchor = S->anchor.as<int>();
        R.O.anchor_min = (S->incremental && S->reanchor_every > 0) ? S->reanchor_every : (1 << 30);
        R.O.anchor_out = S->mcur.as<double>();
        rp.push_back(R);
        SlotProb &P = sp[p];
        P.g = g;
        P.nsplit = S->nsplit;
        for (int par = 0; par < 2; ++par)
            for (int k = 0; k < kSlotKinds; ++k) {
                SlotArgs A = slot_args(S, kKindAccept[k], kKindPropose[k]);
                const long long nn = par;   // only the parity of the step matters
                A.reanchor_step = kKindReanchor[k];
                A.dc_acc = drow(S, nn);
                A.dc_prop = drow(S, nn + 1);
                A.prop_acc = prow(S, nn);
                A.prop_next = prow(S, nn + 1);
                A.c2_prev = c2row(S, nn + 1);   // step n - 1 has the other parity
                A.c2_next = c2row(S, nn);
                P.A[par * kSlotKinds + k] = A;
            }
        ExchProb &E = ep[p];
        E.nrep = S->nrep;
        E.beta = S->beta.as<double>();
        E.c2[0] = c2row(S, 0);
        E.c2[1] = c2row(S, 1);
        E.slot_of = S->slot_of.as<int>();
        E.rung_of = S->rung_of.as<int>();
        E.beta_slot = S->beta_slot.as<double>();
        E.iex = S->iex.as<uint64_t>();
        E.att = S->exch_att.as<long long>();
        E.acc = S->exch_acc.as<long long>();
        E.seed = S->seed;
        E.nfit = S->nfit;
    }
    if (B->mode == 2)
        for (chm_sampler *S : B->S)
            if (!S->basisR32.p) return fail(CHM_E_STATE, "FP32 basis missing");
    B->ntile = (int)tiles.size();
    B->nslot = (int)slots.size();
    if (B->lds_gemm_max > 160 * 1024 || B->lds_slot_max > 160 * 1024)
        return fail(CHM_E_ARG, "problem too large for LDS staging");
    auto up = [&](chm::DevBuf &b, const void *h, size_t bytes) -> int {
        int rc = b.ensure(bytes);
        if (rc) return rc;
        CHM_HIP(hipMemcpy(b.p, h, bytes, hipMemcpyHostToDevice));
        return CHM_OK;
    };
    int rc;
    if ((rc = up(B->tiles, tiles.data(), sizeof(int4) * tiles.size())) ||
        (rc = up(B->slots, slots.data(), sizeof(int2) * slots.size())) ||
        (rc = up(B->gemm[0], gp[0].data(), sizeof(GemmProb) * n)) ||
        (rc = up(B->gemm[1], gp[1].data(), sizeof(GemmProb) * n)) ||
        (rc = up(B->gemm_nb, gnb.data(), sizeof(GemmProb) * n)) ||
        (rc = up(B->reanchor, rp.data(), sizeof(GemmProb) * n)) ||
        (rc = up(B->slot, sp.data(), sizeof(SlotProb) * n)) ||
        (rc = up(B->exch, ep.data(), sizeof(ExchProb) * n))) {
        delete B;
        return rc;
    }
    hipFuncSetAttribute((const void *)k_gemm_grouped<0>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)B->lds_gemm_max);
    hipFuncSetAttribute((const void *)k_gemm_grouped<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)B->lds_gemm_max);
    hipFuncSetAttribute((const void *)k_slot_grouped, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)B->lds_slot_max);
    for (chm_sampler *S : B->S) drop_graph(S);
    chm::ctx_retain(ctx);
    *out = B;
    return CHM_OK;
}

extern "C" int chm_batch_destroy(chm_batch *B) {
    if (!B) return CHM_OK;
    chm_ctx *ctx = B->ctx;
    hipSetDevice(ctx->device);
    hipStreamSynchronize(ctx->stream);
    batch_drop_graph(B);
    delete B;
    chm::ctx_release(ctx);
    return CHM_OK;
}

extern "C" int chm_batch_run(chm_batch *B, int nsteps, int exchange_every) {
    CHM_REQUIRE(B, CHM_E_ARG, "null batch");
    CHM_REQUIRE(nsteps >= 0 && exchange_every >= 0, CHM_E_ARG, "negative step count");
    for (chm_sampler *S : B->S)
        CHM_REQUIRE(S->hstep == B->hstep && S->inited, CHM_E_STATE,
                    "a sampler of the batch was re-initialised or run on its own");
    CHM_HIP(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    for (chm_sampler *S : B->S)   // work the samplers queued on their own streams comes first
        if (S->ctx != B->ctx) CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    const int E = exchange_every;
    int rc = 0;
    if (E > 0 && B->graph_steps != E) {
        batch_drop_graph(B);
        for (int par = 0; par < ((E & 1) ? 2 : 1) && !rc; ++par) {
            CHM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            rc = batch_segment(B, par, E);
            if (!rc) rc = batch_exchange(B, par + E);
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
            B->graph[par] = gr;
            CHM_HIP(hipGraphInstantiate(&B->gexec[par], gr, nullptr, nullptr, 0));
        }
        B->graph_steps = E;
    }
    long long left = nsteps;
    while (left > 0) {
        if (E > 0 && B->hstep % E == 0 && left >= E) {
            CHM_HIP(hipGraphLaunch(B->gexec[(E & 1) ? (B->hstep & 1) : 0], st));
            B->hstep += E;
            left -= E;
        } else {
            long long L = left;
            if (E > 0) L = std::min<long long>(left, E - B->hstep % E);
            if ((rc = batch_segment(B, B->hstep, (int)L))) return rc;
            B->hstep += L;
            left -= L;
            if (E > 0 && B->hstep % E == 0 && (rc = batch_exchange(B, B->hstep))) return rc;
        }
    }
    CHM_HIP(hipStreamSynchronize(st));
    for (chm_sampler *S : B->S) S->hstep = B->hstep;
    return CHM_OK;
}

extern "C" int chm_batch_set_transform(chm_batch *B, int kb_per_wave) {
    CHM_REQUIRE(B, CHM_E_ARG, "null batch");
    CHM_REQUIRE(kb_per_wave == 0 || kb_per_wave == 1 || kb_per_wave == KBW, CHM_E_ARG,
                "K blocks per wave must be 0 (automatic), 1 or 4");
    CHM_REQUIRE(kb_per_wave != 1 || B->max_nkb <= 4, CHM_E_ARG,
                "one K block per wave needs every problem's rows in at most 4 K blocks");
    CHM_HIP(hipSetDevice(B->ctx->device));
    CHM_HIP(hipStreamSynchronize(B->ctx->stream));
    batch_drop_graph(B);
    B->kb_wave = kb_per_wave;
    return CHM_OK;
}

extern "C" int chm_batch_get_shape(chm_batch *B, int *max_nkb) {
    CHM_REQUIRE(B && max_nkb, CHM_E_ARG, "bad arguments");
    *max_nkb = B->max_nkb;
    return CHM_OK;
}

extern "C" int chm_batch_profile(chm_batch *B, int iters, double *ms) {
    CHM_REQUIRE(B && iters > 0 && ms, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(B->ctx->device));
    hipStream_t st = B->ctx->stream;
    hipEvent_t e0, e1;
    CHM_HIP(hipEventCreate(&e0));
    CHM_HIP(hipEventCreate(&e1));
    float t = 0.f;
    const long long h = B->hstep;
    int rc = batch_slot(B, h, 0);                 // step h + 1 pending
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters && !rc; ++i) rc = batch_gemm(B, B->gemm_nb.as<GemmProb>(), B->mode);
    CHM_HIP(hipEventRecord(e1, st));
    if (rc) return rc;
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[0] = t / iters;                              // transform of every problem
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters && !rc; ++i) {
        const long long n = h + 1 + i;
        rc = batch_gemm(B, B->gemm[n & 1].as<GemmProb>(), B->mode);
        if (!rc) rc = batch_slot(B, n, 1);
    }
    CHM_HIP(hipEventRecord(e1, st));
    if (rc) return rc;
    CHM_HIP(hipEventSynchronize(e1));
    CHM_HIP(hipEventElapsedTime(&t, e0, e1));
    ms[2] = t / iters;                              // whole step
    ms[1] = ms[2] - ms[0];                          // per-replica kernel
    const long long n = h + 1 + iters;
    if (!rc) rc = batch_gemm(B, B->gemm[n & 1].as<GemmProb>(), B->mode);
    if (!rc) rc = batch_slot(B, n, 2);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(st));
    B->hstep = n;
    for (chm_sampler *S : B->S) S->hstep = n;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return CHM_OK;
}
