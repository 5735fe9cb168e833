// Device-side pieces shared by the sampler kernels (sampler.hip) and the learned
// proposal kernels (learned.hip): the counter-based RNG, the lattice geometry and site
// arithmetic, the uniform swap proposal and the gather of a swap's count changes.
#pragma once

#include <cstdint>

#include "common.h"

namespace chm_smp {

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

// Learned proposal: the per-step stream of a rung and the state-independent coin that
// makes a step use the uniform proposal instead (probability eps, and every slot colder
// than beta_max).  Shared by the proposal kernels and the accept kernel (which flips the
// direction of a lifted learned step that was not accepted).
constexpr uint64_t kLpSalt = 0x2545F4914F6CDD1Dull;
constexpr uint64_t kCoinSalt = 0x5851F42D4C957F2Dull;
__device__ __forceinline__ uint64_t lp_stream(uint64_t seed, uint64_t step, int rung) {
    return splitmix64(seed ^ (step * kGolden) ^ ((uint64_t)rung << 32) ^ kLpSalt);
}
__device__ __forceinline__ bool lp_coin_uniform(uint64_t h0, double beta, double eps, double beta_max) {
    return u01(splitmix64(h0 ^ kCoinSalt)) < eps || beta > beta_max;
}
// candidate pair i of a step's candidate set (stream h0), and the uniform that selects one
constexpr uint64_t kSelSalt = 0x14057B7EF767814Full;
constexpr uint64_t kCandMul = 0xD6E8FEB86659FD93ull;
__device__ __forceinline__ void lp_candidate(uint64_t h0, int i, int natoms, int &a, int &b) {
    const uint64_t h = splitmix64(h0 ^ ((uint64_t)(i + 1) * kCandMul));
    a = (int)(((h & 0xFFFFFFFFull) * (uint64_t)natoms) >> 32);
    b = (int)(((h >> 32) * (uint64_t)natoms) >> 32);
}

// Metropolis test of methods/method_a_rmc.py:280-285 for x = beta * d(chi^2) * nfit (minus
// the log proposal ratio): accept if x <= 0 or u < exp(-min(x, 60)).  A NaN x is rejected,
// as the reference's cp.minimum / cp.exp propagate it (fmin alone would drop the NaN and
// accept with probability e^-60).
__device__ __forceinline__ bool metropolis(double x, double u) {
    return (x <= 0.0) || (!__builtin_isnan(x) && u < exp(-fmin(x, 60.0)));
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float f) {
    unsigned int u = __builtin_bit_cast(unsigned int, f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
// two floats rounded to bf16 (nearest even, as bf16_rne for finite values) in one packed
// conversion instruction: a in the low half, b in the high half
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
    typedef float f2_t __attribute__((ext_vector_type(2)));
    typedef __bf16 b2_t __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(unsigned, __builtin_convertvector((f2_t){a, b}, b2_t));
}
__device__ __forceinline__ float bf16_to_float(unsigned short h) {
    return __builtin_bit_cast(float, (unsigned int)h << 16);
}

// DPP moves (every lane of the row pattern active): neighbour swap 0xB1, pair swap 0x4E,
// half-row mirror 0x141, row mirror 0x140
template <int kCtrl>
__device__ __forceinline__ float dpp_f(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), kCtrl, 0xF, 0xF, false));
}
template <int kCtrl>
__device__ __forceinline__ double dpp_d(double v) {
    const int2 w = __builtin_bit_cast(int2, v);
    return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_update_dpp(0, w.x, kCtrl, 0xF, 0xF, false),
                                                __builtin_amdgcn_update_dpp(0, w.y, kCtrl, 0xF, 0xF, false)));
}
// sum over aligned groups of W lanes (W = 8 or 16: within the DPP rows of 16), in every lane
// of the group: neighbours, pairs, quads (half-row mirror), halves (row mirror)
template <int W>
__device__ __forceinline__ double group_sum_d(double v) {
    static_assert(W == 8 || W == 16, "groups of 8 or 16 lanes");
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    if (W == 16) v += dpp_d<0x140>(v);
    return v;
}

// Bounds-checked loads through a buffer resource (raw buffer, stride 0): a lane whose
// byte offset is at or beyond the resource's size gets zero and makes no memory access,
// without a branch, so a block of such loads is issued back to back and waited on once.
struct Buf {
    __amdgpu_buffer_rsrc_t r;
    uint32_t bytes;
};
__device__ __forceinline__ Buf make_buf(const void *p, uint32_t bytes) {
    return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000),
               bytes};
}
__device__ __forceinline__ double buf_f64(const Buf &b, bool ok, uint32_t off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(b.r, ok ? off : b.bytes, 0, 0));
}
__device__ __forceinline__ int buf_i32(const Buf &b, bool ok, uint32_t off) {
    return (int)__builtin_amdgcn_raw_buffer_load_b32(b.r, ok ? off : b.bytes, 0, 0);
}
__device__ __forceinline__ int2 buf_i64(const Buf &b, bool ok, uint32_t off) {
    return __builtin_bit_cast(int2, __builtin_amdgcn_raw_buffer_load_b64(b.r, ok ? off : b.bytes, 0, 0));
}
__device__ __forceinline__ int4 buf_i128(const Buf &b, bool ok, uint32_t off) {
    return __builtin_bit_cast(int4, __builtin_amdgcn_raw_buffer_load_b128(b.r, ok ? off : b.bytes, 0, 0));
}

struct Geo {
    int natoms, natoms_pad, nsp, ncoord, nshell, nblk, nstencil, nabsrec;
    int nrow_e, rstride, nrow_tot, nss_rows, ncol_e, ncol_tot, nkw, nwin;  // rstride: padded rows per edge
    int box[3], ncell[3], nbasis;
    int bpar[8];
    int has_const;
    int nct, nkb;   // 16-column tiles per edge, 32-row K blocks per edge
    int nrow_w;     // 32-bit words per count-change row (two rows per word)
    int bxyz[12];   // coordinates (units a0/2) of each basis site in the cell
    // the same two tables packed into one register each, so a per-lane lookup is a shift
    // (an indexed read of the arrays above would be a memory load per lane):
    // bpar4 nibble q = bpar[q] (4-bit two's complement), bxyz3 bit 3b + d = bxyz[3b + d]
    uint32_t bpar4, bxyz3;
    float rnb, rnc1, rnc2;   // 1/nbasis, 1/ncell[1], 1/ncell[2] (float) for div_small
};

__device__ __forceinline__ int site_index(const Geo &g, int x, int y, int z) {
    const int par = (x & 1) | ((y & 1) << 1) | ((z & 1) << 2);
    const int b = (int)(g.bpar4 << (28 - 4 * par)) >> 28;
    int cell = ((x >> 1) * g.ncell[1] + (y >> 1)) * g.ncell[2] + (z >> 1);
    return cell * g.nbasis + b;
}

__device__ __forceinline__ int wrap(int v, int n) {
    v = (v < 0) ? v + n : v;
    return (v >= n) ? v - n : v;
}

// Count-change rows hold two rows per 32-bit word: row 2w in the low half and row 2w+1 in
// the high half, each a 16-bit two's-complement value (w = lo + 65536 hi, modulo 2^32).
// Adding packed words with 32-bit integer atomics adds both halves exactly, so a merged
// word decodes exactly as long as both final values fit in 16 bits: a swap changes a
// count by at most 2 x (gather items) x 2 < 2^15.  The transform reads half the bytes.
__device__ __forceinline__ int dw_lo(int w) { return (int)((unsigned)w << 16) >> 16; }
__device__ __forceinline__ int dw_hi(int w) { return (int)((unsigned)w - (unsigned)dw_lo(w)) >> 16; }
__device__ __forceinline__ int dw_pack(int lo, int hi) { return (int)((unsigned)lo + ((unsigned)hi << 16)); }
__device__ __forceinline__ int dw_row(const int *dw, int q) {   // row q of a packed row
    const int w = dw[q >> 1];
    return (q & 1) ? dw_hi(w) : dw_lo(w);
}

struct Prop {
    int a, b, na, nb, sw;
};

// kernels.py:378-398: h = splitmix64(seed ^ step*golden ^ key<<32), then swap flag,
// a, and up to 8 candidates.  swap_frac = 1: the flag is always set.  All 8 candidates
// repeating a's species makes a null move.
__device__ inline Prop propose(const uint8_t *sp, int natoms, uint64_t seed, uint64_t stepc,
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

// The same proposal in two halves: the draws (a and the 8 candidates) do not depend on
// the configuration, so they can be made before it is final; the pick reads it.
struct PropDraws {
    int a;
    int c[8];
};
__device__ __forceinline__ PropDraws propose_draws(int natoms, uint64_t seed, uint64_t stepc, uint64_t key) {
    uint64_t h = splitmix64(seed ^ (stepc * kGolden) ^ (key << 32));
    h = splitmix64(h);
    h = splitmix64(h);
    PropDraws d;
    d.a = min((int)(u01(h) * natoms), natoms - 1);
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        h = splitmix64(h);
        d.c[t] = min((int)(u01(h) * natoms), natoms - 1);
    }
    return d;
}
// species lookups: as stored, or with one swap x made (x.a holds x.na, x.b holds x.nb)
struct SpPlain {
    const uint8_t *sp;
    __device__ __forceinline__ int operator()(int q) const { return sp[q]; }
};
struct SpOver {
    const uint8_t *sp;
    int a, b, na, nb;
    __device__ __forceinline__ int operator()(int q) const { return q == a ? na : (q == b ? nb : (int)sp[q]); }
};
template <class SP>
__device__ __forceinline__ Prop propose_pick_sp(const SP &sp, const PropDraws &d) {
    const int sa = sp(d.a);
    int sc[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) sc[t] = sp(d.c[t]);
    Prop p{d.a, d.a, sa, sa, 0};
#pragma unroll
    for (int t = 7; t >= 0; --t)
        if (sc[t] != sa) p = Prop{d.a, d.c[t], sc[t], sa, 1};
    return p;
}
__device__ __forceinline__ Prop propose_pick(const uint8_t *sp, const PropDraws &d) {
    return propose_pick_sp(SpPlain{sp}, d);
}

__device__ __forceinline__ void load_species(uint8_t *lsp, const uint8_t *gsp, int natoms_pad) {
    const uint4 *src = reinterpret_cast<const uint4 *>(gsp);
    uint4 *dst = reinterpret_cast<uint4 *>(lsp);
    for (int q = threadIdx.x; q < natoms_pad / 16; q += blockDim.x) dst[q] = src[q];
}

// n / d for 0 <= n < 2^20 and d >= 1, from a float reciprocal of d: the truncated estimate
// is within one of the quotient (relative error below 2^-21), and one correction each way
// makes it exact.  A few VALU operations instead of the generic integer division.
__device__ __forceinline__ int div_small(int n, int d, float rcp) {
    int q = (int)((float)n * rcp);
    const int r = n - q * d;
    q += (r >= d) ? 1 : 0;
    q -= (r < 0) ? 1 : 0;
    return q;
}

// integer coordinates (units a0/2) of a site, from its index (natoms < 2^20)
__device__ __forceinline__ void site_xyz_of(const Geo &g, int q, int &x, int &y, int &z) {
    const int cell = div_small(q, g.nbasis, g.rnb), b = q - cell * g.nbasis;
    const int c12 = div_small(cell, g.ncell[2], g.rnc2);
    const int cz = cell - c12 * g.ncell[2];
    const int cx = div_small(c12, g.ncell[1], g.rnc1);
    const int cy = c12 - cx * g.ncell[1];
    x = 2 * cx + (int)((g.bxyz3 >> (3 * b)) & 1u);
    y = 2 * cy + (int)((g.bxyz3 >> (3 * b + 1)) & 1u);
    z = 2 * cz + (int)((g.bxyz3 >> (3 * b + 2)) & 1u);
}

struct SwapSite {
    Prop p;
    int ax, ay, az, bx, by, bz;
};

// One gather item of a swap: item < nss is a neighbour of a (first ncoord) or of b (rec.x
// holds its packed stencil entry); the rest are path records `rec` of a's then b's
// incidence stencil.  Count changes go to `hist`.
template <class SP>
__device__ __forceinline__ void gather_item_sp(const Geo &g, const SwapSite &W, const SP &lsp,
                                               int *hist, int it, int nss, const int4 rec) {
    const Prop &p = W.p;
    if (it < nss) {
        // ordered pairs x->y and y->x for every neighbour y of x in {a, b}
        const bool on_a = it < g.ncoord;
        const int sv = rec.x;   // the neighbour stencil entry of this item
        const int dx = (int)(int8_t)(sv & 0xFF), dy = (int)(int8_t)((sv >> 8) & 0xFF);
        const int dz = (int)(int8_t)((sv >> 16) & 0xFF), sh = (sv >> 24) & 0xFF;
        const int x = on_a ? p.a : p.b;
        const int y = site_index(g, wrap((on_a ? W.ax : W.bx) + dx, g.box[0]),
                                 wrap((on_a ? W.ay : W.by) + dy, g.box[1]),
                                 wrap((on_a ? W.az : W.bz) + dz, g.box[2]));
        if (x == p.b && y == p.a) return;  // covered from a's side
        const int sx = lsp(x), sy = lsp(y);
        const int nx = (x == p.a) ? p.na : p.nb;
        const int ny = (y == p.a) ? p.na : ((y == p.b) ? p.nb : sy);
        atomicAdd(&hist[sx * g.rstride + sy * g.nshell + sh], -1);
        atomicAdd(&hist[nx * g.rstride + ny * g.nshell + sh], 1);
        atomicAdd(&hist[sy * g.rstride + sx * g.nshell + sh], -1);
        atomicAdd(&hist[ny * g.rstride + nx * g.nshell + sh], 1);
        return;
    }
    const int m = it - nss;
    const bool on_a = m < g.nstencil;
    const int8_t *rb = reinterpret_cast<const int8_t *>(&rec);
    const int nslot = rb[12], wgt = rb[13];
    const int cls_off = (int)(uint8_t)rb[14] | ((int)rb[15] << 8);
    const int x0 = on_a ? W.ax : W.bx, y0 = on_a ? W.ay : W.by, z0 = on_a ? W.az : W.bz;
    int sites[4];
    bool has_a = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q > nslot) {
            sites[q] = -1;
            continue;
        }
        sites[q] = site_index(g, wrap(x0 + rb[3 * q], g.box[0]), wrap(y0 + rb[3 * q + 1], g.box[1]),
                              wrap(z0 + rb[3 * q + 2], g.box[2]));
        has_a |= (sites[q] == p.a);
    }
    if (!on_a && has_a) return;  // a path through both sites counts once
    int eo = 0, en = 0, pw = 1;
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        if (q > nslot) break;
        const int st = sites[q];
        const int so = lsp(st);
        const int sn = (st == p.a) ? p.na : ((st == p.b) ? p.nb : so);
        eo += so * pw;
        en += sn * pw;
        pw *= g.nsp;
    }
    const int sa0 = lsp(sites[0]);
    const int sa1 = (sites[0] == p.a) ? p.na : ((sites[0] == p.b) ? p.nb : sa0);
    eo += sa0 * g.rstride + g.nss_rows + cls_off;
    en += sa1 * g.rstride + g.nss_rows + cls_off;
    if (eo != en) {
        atomicAdd(&hist[eo], -wgt);
        atomicAdd(&hist[en], wgt);
    }
}


__device__ __forceinline__ void gather_item(const Geo &g, const SwapSite &W, const uint8_t *lsp,
                                            int *hist, int it, int nss, const int4 rec) {
    gather_item_sp(g, W, SpPlain{lsp}, hist, it, nss, rec);
}

constexpr int kRecPrefetch = 4;   // gather records per lane loaded ahead

// The record of gather item `it` (< 2 ncoord + 2 nstencil): the packed neighbour stencil
// entry (in .x) or the path record.
__device__ __forceinline__ int4 item_rec(const Geo &g, const int *__restrict__ ss_sten,
                                         const int4 *__restrict__ ms_sten, int it) {
    const int nss_items = 2 * g.ncoord;
    if (it < nss_items) return make_int4(ss_sten[it < g.ncoord ? it : it - g.ncoord], 0, 0, 0);
    const int m = it - nss_items;
    return ms_sten[m < g.nstencil ? m : m - g.nstencil];
}

// Gather items of a swap split over the workgroups of one replica slot: item `first +
// r * stride`.  The first kRecPrefetch records of this lane do not depend on the
// proposal, so they are requested before it is known: both candidate records (neighbour
// stencil entry and path record) are loaded with clamped, in-bounds indexes and no
// branch; the gather picks the one its item needs.
struct RecPre {
    int4 ms[kRecPrefetch];
    int ss[kRecPrefetch];
};
__device__ __forceinline__ void prefetch_recs(const Geo &g, const int *__restrict__ ss_sten,
                                              const int4 *__restrict__ ms_sten, int first,
                                              int stride, bool on, RecPre &R) {
    const int nss_items = 2 * g.ncoord;
    const int nitems = nss_items + 2 * g.nstencil;
    const Buf bs = make_buf(ss_sten, 4u * g.ncoord);
    const Buf bm = make_buf(g.nstencil > 0 ? (const void *)ms_sten : (const void *)ss_sten, 16u * g.nstencil);
#pragma unroll
    for (int r = 0; r < kRecPrefetch; ++r) {
        const int it = first + r * stride;
        const int m = it - nss_items;
        R.ss[r] = buf_i32(bs, on && it < nss_items, 4u * (it < g.ncoord ? it : it - g.ncoord));
        R.ms[r] = buf_i128(bm, on && m >= 0 && it < nitems, 16u * (m < g.nstencil ? m : m - g.nstencil));
    }
}

// The count changes of swap p (a real swap) into `hist` (zeroed by the caller, visible
// to the workgroup), then this workgroup's nonzero bins added to the slot's packed row
// `dc` with integer atomics (order-independent, so the merged row is deterministic).
__device__ __forceinline__ void gather_swap(const Geo &g, const Prop &p, const int *__restrict__ ss_sten,
                                            const int4 *__restrict__ ms_sten, const uint8_t *lsp,
                                            int *hist, int first, int stride, const RecPre &R, int *dc) {
    SwapSite W;
    W.p = p;
    site_xyz_of(g, p.a, W.ax, W.ay, W.az);
    site_xyz_of(g, p.b, W.bx, W.by, W.bz);
    const int nss_items = 2 * g.ncoord;
    const int nitems = nss_items + 2 * g.nstencil;
#pragma unroll
    for (int r = 0; r < kRecPrefetch; ++r) {
        const int it = first + r * stride;
        if (it < nitems)
            gather_item(g, W, lsp, hist, it, nss_items, it < nss_items ? make_int4(R.ss[r], 0, 0, 0) : R.ms[r]);
    }
    for (int it = first + kRecPrefetch * stride; it < nitems; it += stride)
        gather_item(g, W, lsp, hist, it, nss_items, item_rec(g, ss_sten, ms_sten, it));
    __syncthreads();
    for (int w = threadIdx.x; w < g.nrow_w; w += blockDim.x) {
        const int v = dw_pack(hist[2 * w], hist[2 * w + 1]);
        if (v != 0) atomicAdd(&dc[w], v);
    }
}

// Site of occupancy entry q of the candidate swap (a, b): q < ncoord the q-th stencil
// neighbour of a, then b's, then a and b themselves.  (ax..bz: their coordinates.)
__device__ __forceinline__ int occ_site(const Geo &g, const int *__restrict__ ss_sten, int a, int b,
                                        int ax, int ay, int az, int bx, int by, int bz, int q) {
    if (q >= 2 * g.ncoord) return q == 2 * g.ncoord ? a : b;
    const bool on_a = q < g.ncoord;
    const int sv = ss_sten[on_a ? q : q - g.ncoord];
    const int dx = (int)(int8_t)(sv & 0xFF), dy = (int)(int8_t)((sv >> 8) & 0xFF);
    const int dz = (int)(int8_t)((sv >> 16) & 0xFF);
    return site_index(g, wrap((on_a ? ax : bx) + dx, g.box[0]), wrap((on_a ? ay : by) + dy, g.box[1]),
                      wrap((on_a ? az : bz) + dz, g.box[2]));
}

// species at a site with the sites xa and xb exchanged (xa < 0: as stored)
__device__ __forceinline__ int species_swapped(const uint8_t *lsp, int site, int xa, int xb) {
    if (xa >= 0) {
        if (site == xa) return lsp[xb];
        if (site == xb) return lsp[xa];
    }
    return lsp[site];
}

// One replica-exchange sweep (methods/method_a_rmc.py:332-373): exchange number iex + 1
// picks the parity of the first rung from its stream; neighbour rungs i, i + 1 of that
// parity swap temperatures with probability min(1, e^d), d = (beta_i - beta_i+1)
// (chi2_i - chi2_i+1) nfit: the rung <-> slot maps and the slots' betas change.
__device__ __forceinline__ void exchange_body(int nrep, const double *__restrict__ beta, const double *__restrict__ c2,
                           int *__restrict__ slot_of, int *__restrict__ rung_of,
                           double *__restrict__ beta_slot, uint64_t *__restrict__ iex,
                           long long *__restrict__ att, long long *__restrict__ acc, uint64_t seed,
                           double nfit) {
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
        bool ok = metropolis(-d, u);
        att[i] += 1;
        if (ok) {
            acc[i] += 1;
            slot_of[i] = sj;
            slot_of[i + 1] = si;
            rung_of[sj] = i;
            rung_of[si] = i + 1;
            beta_slot[sj] = beta[i];
            beta_slot[si] = beta[i + 1];
        }
    }
}


}  // namespace chm_smp
