// Device side of the learned swap proposal (no counterpart in the reference: its swap
// proposal is uniform, exafs_gpu/kernels.py:362-398), shared by the learned-proposal
// kernels (learned.hip) and the sampler's transform and per-replica kernels (sampler.hip).
//
// Network of a candidate swap (a, b): one-hot species occupancy of a's and b's stencil
// neighbours and of a and b (nocc = 2 ncoord + 2 entries, nsp classes each):
//     Z1 = b1 + X W1q,  H1 = bf16(silu(Z1)),  Z2 = b2 + H1 W2,  H2 = silu(Z2),
//     g = b3 + H2 . w3          (W2, w3 in bf16, sums in FP32; W2 on MFMA)
// g predicts the swap's energy change nfit (chi2(x') - chi2(x)); a candidate's logit is
// -lam beta g / 2 + tilt(dU) (dU: change of the weighted first-shell bond count), -inf when
// a and b hold one species.
//
// Layer 1 is linear in the one-hot occupancy: X W1q = A[a] + B[b], the two per-site sums
// A[c] = sum over c's stencil positions p of W1q[(p, species)] + W1q[(own a, species of c)],
// B[c] with b's weight blocks.  W1q is bf16(W1) rounded to a multiple of 2^-16 and clamped
// to |w| <= 2, held as integers, so every sum is exact (int32, and below 2^24, exact in FP32
// too): the device keeps A and B of every site of every replica slot and updates them by
// the weight rows a swap changes, with no drift.  The per-replica kernel that accepts a
// swap folds it in (its workgroups share the rows), so the next transform's scoring
// workgroups read the sums of the committed configuration; nothing reads them while they
// are written.
//
// Speculative scoring.  The scores of step n+1's candidates depend on whether step n is
// accepted; the scoring workgroups of step n's transform launch score them at the
// configuration after step n - 1 and, where step n's swap reaches them, with that swap
// made; they also score step n's candidates near its swap with the swap made (its
// reverse probability).  Each 32-candidate tile writes its g / dU and partial softmax
// sums; the per-replica kernel that decides step n merges them into step n's Hastings
// ratio and draws step n+1's proposal for the outcome it decided.
#pragma once

#include <climits>
#include <cstdint>

#include "common.h"
#include "sampler_dev.h"
#include "sampler_state.h"

namespace chm_lp_dev {

using namespace chm_smp;

constexpr int HID = 256;      // hidden width
constexpr int MR = 32;        // candidates per tile
constexpr int MAXC = 256;     // candidates per replica and step, at most
constexpr int MAXT = MAXC / MR;
constexpr int LDH = HID + 8;  // H1 row stride in LDS (bf16), staggers the banks
constexpr int TB = 256;       // threads of a scoring / folding workgroup
constexpr int TW = TB / 64;   // its waves
constexpr int kMaxSpW = 8;    // species dimension of the bond-weight table
constexpr int kMaxFs = 16;    // first-shell stencil entries, at most
constexpr int kNCorr = 4;     // layer-1 corrections of a row: 2 candidate sites x 2 swapped sites
constexpr int kLqDoubles = 6; // per tile: forward (max, Z, D) and reverse (max, Z, D) of log q
constexpr int kFoldPerSlot = 8;   // folding workgroups per replica slot (per-replica kernel)
constexpr int MP = 48;        // rows of one network pass: a tile's candidates, then re-scored ones
constexpr int kPosMapMax = 4096;  // bytes of the offset -> stencil position cube staged in LDS
constexpr float kW1Inv = 1.f / 65536.f;
constexpr int kW1QMax = 131072;

struct LPW {
    const __bf16 *w2t;   // (HID, HID): row n = output unit
    const float *b1, *b2, *w3q, *b3;
};

// Everything the scoring workgroups of one launch read; built by the host for one step
// (mode 1: step n with step n+1's candidates; 2: step n only, the last of a segment; 0: a
// segment's first candidates after n steps).
struct LPArgs {
    Geo g;
    const uint8_t *species;
    const int *rung_of;
    const double *beta_slot;
    const int *ss_sten;
    const int *nbr;            // (natoms, ncoord) neighbour sites in stencil order
    const signed char *posmap; // stencil position of an offset (cube of side 2 pm_r + 1), or -1
    int pm_r;
    signed char fs[kMaxFs];    // the first-shell stencil positions (n1 of them)
    int n1;
    uint64_t seed;
    int ncand, nrep, r2max;
    double eps, lam, kappa, beta_max;
    signed char wpair[64];     // first-shell bond weights w[s][t] (8 x 8) of the tilt's dU
    int lift;                  // lifted proposal: tilt kappa d dU with d = dir[slot]
    const int *dir;            // (nrep) +-1
    LPW W;
    const int *w1q;            // (kp, HID) layer-1 weights, integers in units of 2^-16
    int mode;
    // the step number, read on the device (the launch may sit in a captured graph that
    // replays it at other steps): steps done (mode 0) or the step the per-replica launch
    // before this one set up (mode 1, 2)
    const uint64_t *stepc;
    const uint64_t *lpn;
    const int *cache;          // (nrep, 2, natoms, HID): the sums of the committed configuration
    const int *prop_n;         // (nrep, kPropInts) step n's proposal
    const int *cs;             // (nrep) 1: step n - 1's swap was accepted (mode 1 / 2)
    const float *sg_in;        // step n's candidates as drawn: (nrep, 2, ncand) g at x (0) and
    const int *sdu_in;         //   with step n-1's swap (1), their dU, and (nrep, ncand) whether
    const int8_t *st_in;       //   that swap reached them
    float *sg_out;             // the same for step n+1's candidates
    int *sdu_out;
    int8_t *st_out;
    double *lq;                // (nrep, MAXT, kLqDoubles) partial sums of step n's log q
    float *dmax;               // (nrep, MAXT, 2) tile maxima of step n+1's logits: accepted,
                               //   not accepted
    int ntile_blocks;
};

__device__ __forceinline__ float silu_fast(float z) { return z * __frcp_rn(1.f + __expf(-z)); }

__device__ __forceinline__ int centred(int d, int n) {
    d = (d < 0) ? d + n : d;
    return (2 * d > n) ? d - n : d;
}
__device__ __forceinline__ int dist2_sites(const Geo &g, int p, int q) {
    int px, py, pz, qx, qy, qz;
    site_xyz_of(g, p, px, py, pz);
    site_xyz_of(g, q, qx, qy, qz);
    const int dx = centred(qx - px, g.box[0]), dy = centred(qy - py, g.box[1]), dz = centred(qz - pz, g.box[2]);
    return dx * dx + dy * dy + dz * dz;
}
// stencil position of site x around site c, or -1 (x is not one of c's stencil neighbours);
// pm: the offset -> position cube (in LDS)
__device__ __forceinline__ int stencil_pos(const Geo &g, int pm_r, const signed char *pm, int c, int x) {
    int cx, cy, cz, xx, xy, xz;
    site_xyz_of(g, c, cx, cy, cz);
    site_xyz_of(g, x, xx, xy, xz);
    const int dx = centred(xx - cx, g.box[0]), dy = centred(xy - cy, g.box[1]), dz = centred(xz - cz, g.box[2]);
    const int R = pm_r, side = 2 * R + 1;
    if (dx < -R || dx > R || dy < -R || dy > R || dz < -R || dz > R) return -1;
    return pm[((dx + R) * side + (dy + R)) * side + (dz + R)];
}

// the tilt of a candidate's logit: kappa |dU| (symmetric), or kappa d dU for a lifted
// proposal moving in direction d = +-1 (the reverse move is scored with -d)
__device__ __forceinline__ float logit_of(double lam, double kappa, int lift, float g, int du, double beta, int d) {
    if (du == INT_MIN) return -__builtin_huge_valf();
    const double tl = lift ? kappa * (double)(d * du) : kappa * fabs((double)du);
    return (float)(-lam * beta * 0.5 * (double)g + tl);
}

__device__ __forceinline__ int feat_nbr(const Geo &g, int role, int p, int s) {
    return (role * g.ncoord + p) * g.nsp + s;
}
__device__ __forceinline__ int feat_own(const Geo &g, int role, int s) { return (2 * g.ncoord + role) * g.nsp + s; }

// A swap record: sites a, b and their species before it (v = 0: no swap)
__device__ __forceinline__ int2 corr_entry(const Geo &g, int pm_r, const signed char *pm, int c, int role, int x,
                                           int s_tgt, int s_old) {
    if (s_tgt == s_old) return make_int2(-1, -1);
    if (x == c) return make_int2(feat_own(g, role, s_tgt), feat_own(g, role, s_old));
    const int p = stencil_pos(g, pm_r, pm, c, x);
    if (p < 0) return make_int2(-1, -1);
    return make_int2(feat_nbr(g, role, p, s_tgt), feat_nbr(g, role, p, s_old));
}

__device__ __forceinline__ void add_row16(int (&acc)[16], const int *row, int sign) {
    const int4 *p = reinterpret_cast<const int4 *>(row);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int4 v = p[k];
        acc[4 * k] += sign * v.x;
        acc[4 * k + 1] += sign * v.y;
        acc[4 * k + 2] += sign * v.z;
        acc[4 * k + 3] += sign * v.w;
    }
}

// Sum over the 16 lanes of each DPP row (lanes 16 k .. 16 k + 15), in every lane: swaps of
// neighbours, of pairs, of quads (half-row mirror) and of halves (row mirror).
__device__ __forceinline__ float row16_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));
    return v;
}

// max / sum over lanes 0..31 (all active), in each of them: the rows of 16 by DPP, then
// the two rows by one exchange
__device__ __forceinline__ float max32(float v) {
    v = fmaxf(v, dpp_f<0xB1>(v));
    v = fmaxf(v, dpp_f<0x4E>(v));
    v = fmaxf(v, dpp_f<0x141>(v));
    v = fmaxf(v, dpp_f<0x140>(v));
    return fmaxf(v, __shfl_xor(v, 16, 32));
}
__device__ __forceinline__ double sum32(double v) {
    v += dpp_d<0xB1>(v);
    v += dpp_d<0x4E>(v);
    v += dpp_d<0x141>(v);
    v += dpp_d<0x140>(v);
    return v + __shfl_xor(v, 16, 32);
}

// One K half of a wave's W2 fragments (hidden units [64 w, 64 w + 64), four 16-column
// tiles), with b2 and w3 of those units
struct W2Half {
    bf16x8 B[4][4];
};
__device__ __forceinline__ void load_w2_half(const LPW &W, int half, W2Half &F) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = 8 * (lane >> 4);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const __bf16 *w2 = W.w2t + (long)(64 * wave + 16 * c + lr) * HID + lk + 128 * half;
#pragma unroll
        for (int k = 0; k < 4; ++k) F.B[k][c] = *reinterpret_cast<const bf16x8 *>(w2 + 32 * k);
    }
}

__device__ __forceinline__ void load_b2w3(const LPW &W, float (&b2)[4], float (&w3)[4]) {
    const int t = threadIdx.x, wave = t >> 6, lr = t & 15;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        b2[c] = W.b2[64 * wave + 16 * c + lr];
        w3[c] = W.w3q[64 * wave + 16 * c + lr];
    }
}

// g of the first 16 nrt rows (nrt <= MP / 16) whose H1 (bf16, MP x LDH) is in LDS: Z2 = b2 +
// H1 W2 on v_mfma_f32_16x16x32_bf16 (wave w: hidden units [64 w, 64 w + 64)), then b3 +
// silu(Z2) . w3; gp (TW x MP) is LDS scratch, the result goes to gout (LDS).  F holds the
// first K half of the W2 fragments (requested by the caller ahead of time); the second is
// loaded into the same registers, so the scoring workgroups fit beside the transform's in
// one launch.  Every thread of the TB-thread workgroup calls it.
__device__ inline void layer2(const LPW &W, const __bf16 *H, int nrt, W2Half &F, const float (&b2)[4],
                              const float (&w3)[4], float *gp, float *gout) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = 8 * (lane >> 4);
    constexpr int RT = MP / 16;
    f32x4 acc[RT][4];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int half = 0; half < 2; ++half) {
        if (half == 1) load_w2_half(W, 1, F);
        const W2Half &FH = F;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int kt = 4 * half + k;
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                if (r >= nrt) break;
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(H + (16 * r + lr) * LDH + 32 * kt + lk);
#pragma unroll
                for (int c = 0; c < 4; ++c)
                    acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, FH.B[k][c], acc[r][c], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        if (r >= nrt) break;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v = 0.f;
#pragma unroll
            for (int c = 0; c < 4; ++c) v += silu_fast(acc[r][c][i] + b2[c]) * w3[c];
            v = row16_sum(v);
            if (lr == 0) gp[wave * MP + 16 * r + 4 * (lane >> 4) + i] = v;
        }
    }
    __syncthreads();
    if (t < 16 * nrt) {
        float v = 0.f;
#pragma unroll
        for (int w = 0; w < TW; ++w) v += gp[w * MP + t];
        gout[t] = W.b3[0] + v;
    }
    __syncthreads();
}

// H1 of row r, columns [16 ch, 16 ch + 16): silu(b1 + 2^-16 sum) rounded to bf16, into LDS
__device__ __forceinline__ void h1_store(const LPW &W, __bf16 *H, int r, int ch, const int (&acc)[16], bool on) {
    bf16x8 v0, v1;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int col = 16 * ch + i;
        const float z = (float)acc[i] * kW1Inv + W.b1[col];
        const __bf16 h = __builtin_bit_cast(__bf16, on ? bf16_rne(silu_fast(z)) : (unsigned short)0);
        if (i < 8)
            v0[i] = h;
        else
            v1[i - 8] = h;
    }
    *reinterpret_cast<bf16x8 *>(H + r * LDH + 16 * ch) = v0;
    *reinterpret_cast<bf16x8 *>(H + r * LDH + 16 * ch + 8) = v1;
}

// Rows of one network pass, in LDS: candidate (a, b) at the committed state (v = 0) or with
// the proposal swapped (v = 1); dU and g per row, and the flattened list of layer-1
// corrections (row, new weight row, old weight row).
struct Rows {
    int a[MP], b[MP], v[MP];
    int du[MP];
    float gout[MP];
    float gp[TW * MP];
    int crow[MP * kNCorr], cnew[MP * kNCorr], cold[MP * kNCorr];
    int ncl;
};

// The row list of a tile: entry q < n0 is candidate i0 + q of step n+1 at the committed state;
// entry q >= n0 is li[q - n0] (an entry e of the touched test: e < MR candidate i0 + e of step
// n+1, else candidate i0 + e - MR of step n) with the proposal made.  li is the calling wave's
// own copy of the list, so a wave can name any row without a workgroup barrier.
struct RowSrc {
    uint64_t h_next, h_cur;
    int i0, n0, natoms;
    const int *li;
    __device__ __forceinline__ void get(int q, int &a, int &b, int &v) const {
        if (q < n0) {
            lp_candidate(h_next, i0 + q, natoms, a, b);
            v = 0;
        } else {
            const int e = li[q - n0];
            // (a select by mask: a plain conditional of the two wave-uniform values compiled
            // to a stack array)
            const uint64_t m = (e < MR) ? ~0ull : 0ull;
            lp_candidate(h_cur ^ ((h_next ^ h_cur) & m), i0 + (e & (MR - 1)), natoms, a, b);
            v = 1;
        }
    }
};

// g and dU of rows [q0, q0 + nrows) (nrows <= MP) of the list RS of slot s.  The per-site
// sums (`cache`) and the species (lsp) are those of the committed configuration; rows with
// v = 1 have the proposal (pa, pb) made.  fs: first-shell stencil
// positions (n1); wl: bond weights; pm: offset -> stencil position cube; zl: MP x HID int32
// of LDS (the layer-1 sums; H1 is written over them).  The per-site rows are requested
// before the first barrier.
__device__ inline void score_rows(const LPArgs &A, const uint8_t *lsp, const signed char *pm, int pa, int pb,
                                  const int *fs, int n1, const signed char *wl, int s, const RowSrc &RS,
                                  int q0, int nrows, Rows &R, int *zl) {
    const Geo &g = A.g;
    const int t = threadIdx.x, lane = t & 63;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    __bf16 *H = reinterpret_cast<__bf16 *>(zl);
    if (t < nrows) {
        int a, b, v;
        RS.get(q0 + t, a, b, v);
        R.a[t] = a;
        R.b[t] = b;
        R.v[t] = v;
    }
    // stage A: the two per-site rows of each row, one wave per row (rows wave, wave + 4, ...),
    // every load of a lane issued before any is used (a missing row reads nothing: zero)
    constexpr int RW = MP / TW;
    const Buf bc = make_buf(A.cache + (long)s * 2 * g.natoms * HID, 4u * 2u * (uint32_t)g.natoms * HID);
    // (lane k names row wave + TW k; its sites are read back as wave-uniform values)
    int ral = 0, rbl = 0;
    if (lane < RW && wave + TW * lane < nrows) {
        int v;
        RS.get(q0 + wave + TW * lane, ral, rbl, v);
    }
    int4 va[RW], vb[RW];
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const bool on = wave + TW * k < nrows;
        const int ra = __builtin_amdgcn_readlane(ral, k), rb = __builtin_amdgcn_readlane(rbl, k);
        va[k] = buf_i128(bc, on, 16u * ((uint32_t)ra * HID / 4 + lane));
        vb[k] = buf_i128(bc, on, 16u * ((uint32_t)(g.natoms + rb) * HID / 4 + lane));
    }
    const float4 b1 = reinterpret_cast<const float4 *>(A.W.b1)[lane];   // (used in stage C)
    __builtin_amdgcn_sched_barrier(0);
    if (t == 0) R.ncl = 0;
    __syncthreads();
    // dU at the row's state: 4 lanes per row over the first-shell entries of a, then b; the
    // neighbour-table loads are issued before the per-site rows are consumed
    constexpr int NE = 2 * kMaxFs / 4;
    const int ur = t >> 2, l4 = t & 3;
    int y[NE];
#pragma unroll
    for (int j = 0; j < NE; ++j) {
        const int e = l4 + 4 * j;
        y[j] = -1;
        if (ur < nrows && e < 2 * n1)
            y[j] = A.nbr[(long)(e < n1 ? R.a[ur] : R.b[ur]) * g.ncoord + fs[e < n1 ? e : e - n1]];
    }
#pragma unroll
    for (int k = 0; k < RW; ++k) {
        const int r = wave + TW * k;
        if (r < MP)
            reinterpret_cast<int4 *>(zl + r * HID)[lane] =
                make_int4(va[k].x + vb[k].x, va[k].y + vb[k].y, va[k].z + vb[k].z, va[k].w + vb[k].w);
    }
    // corrections of the rows with the proposal made: (row, candidate site, swapped site),
    // four per row, appended to the list
    {
        const int r = t >> 2, role = (t >> 1) & 1, x = (t & 1) ? pb : pa;
        if (r < nrows && R.v[r] && pa >= 0) {
            const int c = role ? R.b[r] : R.a[r];
            const int2 e = corr_entry(g, A.pm_r, pm, c, role, x, species_swapped(lsp, x, pa, pb), lsp[x]);
            if (e.x >= 0) {
                const int slot = atomicAdd(&R.ncl, 1);
                R.crow[slot] = r;
                R.cnew[slot] = e.x;
                R.cold[slot] = e.y;
            }
        }
    }
    {
        int cnt = 0, sa = 0, sb = 0;
        if (ur < nrows) {
            const int a = R.a[ur], b = R.b[ur], v = R.v[ur];
            sa = v ? species_swapped(lsp, a, pa, pb) : lsp[a];
            sb = v ? species_swapped(lsp, b, pa, pb) : lsp[b];
#pragma unroll
            for (int j = 0; j < NE; ++j) {
                const int e = l4 + 4 * j;
                const bool on_a = e < n1;
                if (e >= 2 * n1 || y[j] == (on_a ? b : a)) continue;   // the a-b bond keeps its pair
                const int sy = v ? species_swapped(lsp, y[j], pa, pb) : lsp[y[j]];
                // bond (own species, sy) becomes (other species, sy)
                cnt += on_a ? (wl[sb * kMaxSpW + sy] - wl[sa * kMaxSpW + sy])
                            : (wl[sa * kMaxSpW + sy] - wl[sb * kMaxSpW + sy]);
            }
        }
        cnt += __shfl_xor(cnt, 1, 4);
        cnt += __shfl_xor(cnt, 2, 4);
        if (ur < nrows && l4 == 0) R.du[ur] = (sa != sb) ? cnt : INT_MIN;
    }
    __syncthreads();
    // stage B: the corrections, one wave per list entry (entries wave, wave + 4, ...), the
    // loads of up to CB entries of a lane issued together; integer LDS atomics add them
    {
        constexpr int CB = 8;
        const int ncl = R.ncl;
        const Buf bw = make_buf(A.w1q, 4u * (uint32_t)(2 * g.ncoord + 2) * (uint32_t)g.nsp * HID);
        for (int e0 = 0; e0 < ncl; e0 += TW * CB) {
            int4 wn[CB], wo[CB];
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const int e = e0 + wave + TW * k;
                const bool on = e < ncl;
                wn[k] = buf_i128(bw, on, 16u * ((uint32_t)(on ? R.cnew[e] : 0) * HID / 4 + lane));
                wo[k] = buf_i128(bw, on, 16u * ((uint32_t)(on ? R.cold[e] : 0) * HID / 4 + lane));
            }
#pragma unroll
            for (int k = 0; k < CB; ++k) {
                const int e = e0 + wave + TW * k;
                if (e < ncl) {
                    int *z = zl + R.crow[e] * HID + 4 * lane;
                    atomicAdd(z, wn[k].x - wo[k].x);
                    atomicAdd(z + 1, wn[k].y - wo[k].y);
                    atomicAdd(z + 2, wn[k].z - wo[k].z);
                    atomicAdd(z + 3, wn[k].w - wo[k].w);
                }
            }
        }
    }
    __syncthreads();
    // the first K half of layer 2's weights and b2 / w3, requested before stage C's work
    W2Half F;
    float b2[4], w3[4];
    load_w2_half(A.W, 0, F);
    load_b2w3(A.W, b2, w3);
    // stage C: H1 = bf16(silu(b1 + 2^-16 z)), written over the sums (all read first)
    {
        int4 zv[RW];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int r = wave + TW * k;
            zv[k] = (r < nrows) ? reinterpret_cast<const int4 *>(zl + r * HID)[lane] : make_int4(0, 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int r = wave + TW * k;   // (wave-uniform)
            if (r >= MP) continue;
            uint2 pk = make_uint2(0u, 0u);
            if (r < nrows) {
                pk.x = pk_bf16(silu_fast((float)zv[k].x * kW1Inv + b1.x), silu_fast((float)zv[k].y * kW1Inv + b1.y));
                pk.y = pk_bf16(silu_fast((float)zv[k].z * kW1Inv + b1.z), silu_fast((float)zv[k].w * kW1Inv + b1.w));
            }
            *reinterpret_cast<uint2 *>(H + r * LDH + 4 * lane) = pk;
        }
    }
    __syncthreads();
    layer2(A.W, H, (nrows + 15) >> 4, F, b2, w3, R.gp, R.gout);
}

// log q(x' -> x) - log q(x -> x') of a step from the scoring tiles' partial sums (per tile:
// max m of the logits, sum of e^(l - m), the same sum over the proposed pair): lanes 0..7 of
// the calling wave hold tile `lane` of the forward set, lanes 8..15 the reverse set (tiles
// past the last: m = -inf, z = d = 0).  Every lane of the wave calls it; lane 0 gets the
// result.
__device__ inline double lq_merge(double m, double z, double d) {
    double M = m;
#pragma unroll
    for (int o = 4; o >= 1; o >>= 1) M = fmax(M, __shfl_xor(M, o, 8));
    const double f = (m == -__builtin_huge_val()) ? 0.0 : exp(m - M);
    double Z = z * f, D = d * f;
#pragma unroll
    for (int o = 4; o >= 1; o >>= 1) {
        Z += __shfl_xor(Z, o, 8);
        D += __shfl_xor(D, o, 8);
    }
    const double r = log(D) - log(Z);
    return __shfl(r, 8, 64) - r;
}

// The softmax draw over candidates held four per lane by one wave (candidate 4 lane + q has
// logit lg[q]; m: the largest logit): the candidate whose bin [lo, hi) of the cumulative
// weights (index order) holds u Z, else the last candidate with weight (u Z rounded past the
// end), -1 when none is valid.  Bins meet exactly: a lane's last edge is its inclusive scan
// value, which is the next lane's first.  Every lane of the wave calls it and gets the answer.
__device__ inline int wave_draw4(const float (&lg)[4], float m, double u) {
    const int lane = threadIdx.x & 63;
    const bool any = m > -__builtin_huge_valf();
    double p[4], c[4], run = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        p[q] = (any && lg[q] > -__builtin_huge_valf()) ? exp((double)(lg[q] - m)) : 0.0;
        run += p[q];
        c[q] = run;
    }
    double x = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const double y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    const double prev = __shfl_up(x, 1, 64);
    const double base = lane == 0 ? 0.0 : prev;
    const double target = u * __shfl(x, 63, 64);
    int hit = -1, last = -1;
    double lo = base;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double hi = (q == 3) ? x : base + c[q];
        if (p[q] > 0.0) {
            if (hit < 0 && lo <= target && target < hi) hit = 4 * lane + q;
            last = 4 * lane + q;
        }
        lo = hi;
    }
    const unsigned long long bh = __ballot(hit >= 0);
    if (bh) return __shfl(hit, __ffsll(bh) - 1, 64);
    const unsigned long long bl = __ballot(last >= 0);
    if (bl) return __shfl(last, 63 - __clzll((long long)bl), 64);
    return -1;
}

// partial softmax sums of a tile's lanes (lanes 0..MR-1 = 0..31, all active): max, sum
// e^(l - max), sum over `same`
__device__ __forceinline__ void tile_partial(float l, bool same, double &m, double &Z, double &D) {
    static_assert(MR == 32, "tile reductions span 32 lanes");
    const float mf = max32(l);
    const double p = (mf > -__builtin_huge_valf() && l > -__builtin_huge_valf()) ? exp((double)(l - mf)) : 0.0;
    m = (double)mf;
    Z = sum32(p);
    D = sum32(same ? p : 0.0);
}

// dynamic LDS of a scoring workgroup: the slot's species, H1, the offset -> position cube
__host__ __device__ inline size_t tile_lds(int natoms_pad) {
    return (size_t)((natoms_pad + 15) & ~15) + (size_t)MP * HID * 4 + kPosMapMax;
}

// One scoring workgroup: tile k of slot s (TB threads).  Candidates [k MR, (k+1) MR):
//   (i)   of step n+1 at the committed configuration x (x = the copy with the lag swap made),
//   (ii)  of step n+1 within the radius of step n's swap P, at x with P made,
//   (iii) of step n within the radius of P, at x with P made (the reverse set),
// in one network pass of (i) followed by the rows of (ii) and (iii) (more passes only when
// more than MR of those); writes g / dU of (i), (ii), the partial sums of step n's log q
// ratio and the tile maxima of step n+1's logits for either outcome of step n.
__device__ inline void lp_tile(const LPArgs &A, int s, int tile, char *lds_raw) {
    __shared__ Rows R;
    __shared__ int fs[kMaxFs];
    __shared__ signed char wl[kMaxSpW * kMaxSpW];
    __shared__ int li4[TW][2 * MR];
    __shared__ float gAv[MR], gBv[MR], gRv[MR];
    __shared__ int duAv[MR], duBv[MR], duRv[MR], tBv[MR], tRv[MR];
    const Geo &g = A.g;
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    uint8_t *lsp = reinterpret_cast<uint8_t *>(lds_raw);
    int *zl = reinterpret_cast<int *>(lds_raw + ((g.natoms_pad + 15) & ~15));
    signed char *pm = reinterpret_cast<signed char *>(lds_raw + ((g.natoms_pad + 15) & ~15) + (size_t)MP * HID * 4);
    const int i0 = tile * MR;
    const long cbase = (long)s * A.ncand;
    // (the scoring waves are the launch's critical path: issued ahead of the transform's
    // waves that share their SIMDs)
    __builtin_amdgcn_s_setprio(2);
    // requested before anything is known of the step (one round trip), the small loads the
    // row list needs first: the step, the slot's rung and beta, step n's proposal and the
    // direction; then the species and the offset cube (into registers) and this lane's
    // candidate of step n as step n - 1 left it
    const int rung = A.rung_of[s];
    const double beta = A.beta_slot[s];
    const uint64_t n = (A.mode == 0) ? A.stepc[s] : A.lpn[s], nx = n + 1;
    Prop P{-1, -1, 0, 0, 0};
    if (A.mode != 0) {
        const int *pp = A.prop_n + kPropInts * s;
        P = Prop{pp[0], pp[1], pp[2], pp[3], pp[4]};
    }
    const int d0 = A.lift ? A.dir[s] : 1;
    const int nsv = g.natoms_pad / 16;
    const uint4 *gsp = reinterpret_cast<const uint4 *>(A.species + (long)s * g.natoms_pad);
    const uint4 spv = (t < nsv) ? gsp[t] : make_uint4(0, 0, 0, 0);
    const int side = 2 * A.pm_r + 1, npm16 = (side * side * side + 15) / 16;   // (<= TB)
    const uint4 pmv = (t < npm16) ? reinterpret_cast<const uint4 *>(A.posmap)[t] : make_uint4(0, 0, 0, 0);
    const int wave_u = __builtin_amdgcn_readfirstlane(wave);
    float fgA = 0.f, fgB = 0.f;
    int fdA = INT_MIN, fdB = INT_MIN, ftb = 0, csv = 0;
    if (A.mode != 0 && t < MR) {
        const long c = i0 + t;
        fgA = A.sg_in[cbase * 2 + c];
        fdA = A.sdu_in[cbase * 2 + c];
        fgB = A.sg_in[cbase * 2 + A.ncand + c];
        fdB = A.sdu_in[cbase * 2 + A.ncand + c];
        ftb = A.st_in[cbase + c];
        csv = A.cs[s];
    }
    __builtin_amdgcn_sched_barrier(0);
    const uint64_t h_next = lp_stream(A.seed, nx, rung), h_cur = lp_stream(A.seed, n, rung);
    const bool learn_next = A.mode != 2 && !lp_coin_uniform(h_next, beta, A.eps, A.beta_max);
    const bool learn_cur = A.mode != 0 && !lp_coin_uniform(h_cur, beta, A.eps, A.beta_max);
    const bool has_p = P.sw != 0;
    const bool need_i = learn_next, need_ii = learn_next && has_p, need_iii = learn_cur && has_p;
    if (!need_i && !need_iii) return;   // (the same for every tile of the slot)
    const int n1 = A.n1;
    if (t < kMaxFs) fs[t] = A.fs[t];
    if (t < kMaxSpW * kMaxSpW) wl[t] = A.wpair[t];
    // rows (ii) / (iii): candidates within the stencil radius of P (a superset of the ones a
    // correction reaches; the others score as at x), in candidate order; every wave makes
    // the list itself (lane = entry: (ii) for lane < MR, (iii) else)
    int nl;
    {
        const int kind = lane < MR ? 1 : 2;
        const int c = lane & (MR - 1);
        const bool want = kind == 1 ? need_ii : need_iii;
        bool touched = false;
        if (want) {
            int ca, cb;
            lp_candidate(kind == 1 ? h_next : h_cur, i0 + c, g.natoms, ca, cb);
            touched = dist2_sites(g, ca, P.a) <= A.r2max || dist2_sites(g, ca, P.b) <= A.r2max ||
                      dist2_sites(g, cb, P.a) <= A.r2max || dist2_sites(g, cb, P.b) <= A.r2max;
        }
        const uint64_t m = __ballot(touched);
        if (touched) li4[wave_u][__popcll(m & ((1ull << lane) - 1ull))] = lane;
        nl = __popcll(m);
        if (wave == 0) {
            if (kind == 1) {
                tBv[c] = touched;
                if (need_i) A.st_out[cbase + i0 + c] = (int8_t)touched;
            } else {
                tRv[c] = touched;
            }
        }
    }
    // passes: (i) rows first (when step n+1 is learned), then the re-scored rows
    const int n0 = need_i ? MR : 0, ntot = n0 + nl;
    const RowSrc RS{h_next, h_cur, i0, n0, g.natoms, li4[wave_u]};
    // the species and the offset cube into LDS (read after the first pass's first barrier)
    if (t < nsv) reinterpret_cast<uint4 *>(lsp)[t] = spv;
    for (int q = t + TB; q < nsv; q += TB) reinterpret_cast<uint4 *>(lsp)[q] = gsp[q];
    if (t < npm16) reinterpret_cast<uint4 *>(pm)[t] = pmv;
    for (int q0 = 0; q0 < ntot; q0 += MP) {
        const int nr = min(MP, ntot - q0);
        score_rows(A, lsp, pm, P.a, P.b, fs, n1, wl, s, RS, q0, nr, R, zl);
        if (t < nr) {
            const int q = q0 + t;
            const float gv = R.gout[t];
            const int dv = R.du[t];
            if (q < n0) {
                gAv[q] = gv;
                duAv[q] = dv;
                A.sg_out[cbase * 2 + i0 + q] = gv;
                A.sdu_out[cbase * 2 + i0 + q] = dv;
            } else {
                const int e = li4[0][q - n0], c = e & (MR - 1);
                if (e < MR) {
                    gBv[c] = gv;
                    duBv[c] = dv;
                    A.sg_out[cbase * 2 + A.ncand + i0 + c] = gv;
                    A.sdu_out[cbase * 2 + A.ncand + i0 + c] = dv;
                } else {
                    gRv[c] = gv;
                    duRv[c] = dv;
                }
            }
        }
        __syncthreads();
    }
    if (t >= MR) return;
    // lanes 0..31 of wave 0: candidate i0 + t of step n (partial log q) and of step n+1
    // (tile maxima of its logits for either outcome)
    double *lq = A.lq + ((long)s * MAXT + tile) * kLqDoubles;
    if (need_iii) {
        const int c = i0 + t;
        const bool tb = csv && ftb;
        const float fg = tb ? fgB : fgA;
        const int fdu = tb ? fdB : fdA;
        const float rg = tRv[t] ? gRv[t] : fg;
        const int rdu = tRv[t] ? duRv[t] : fdu;
        int ca, cb;
        lp_candidate(h_cur, c, g.natoms, ca, cb);
        const bool same = (ca == P.a && cb == P.b) || (ca == P.b && cb == P.a);
        double m, Z, D;
        tile_partial(logit_of(A.lam, A.kappa, A.lift, fg, fdu, beta, d0), same, m, Z, D);
        if (t == 0) {
            lq[0] = m;
            lq[1] = Z;
            lq[2] = D;
        }
        tile_partial(logit_of(A.lam, A.kappa, A.lift, rg, rdu, beta, -d0), same, m, Z, D);
        if (t == 0) {
            lq[3] = m;
            lq[4] = Z;
            lq[5] = D;
        }
    }
    if (need_i) {
        // the direction after step n: kept if it is accepted, reversed if a learned step n
        // is not
        const int dR = (A.mode != 0 && learn_cur) ? -d0 : d0;
        const float lR = logit_of(A.lam, A.kappa, A.lift, gAv[t], duAv[t], beta, dR);
        const bool tb = need_ii && tBv[t];
        const float lA = logit_of(A.lam, A.kappa, A.lift, tb ? gBv[t] : gAv[t], tb ? duBv[t] : duAv[t], beta, d0);
        const float mA = max32(lA), mR = max32(lR);
        if (t == 0) {
            A.dmax[((long)s * MAXT + tile) * 2] = mA;
            A.dmax[((long)s * MAXT + tile) * 2 + 1] = mR;
        }
    }
}

// Folding swap (a, b) into a slot's per-site layer-1 sums (cache: the slot's 2 x natoms x HID
// block): site a takes species na (it held oa), b takes nb (it held ob).  The swap changes
// the rows of a, b and their stencil neighbours.  Work item it = (site, role, 16 columns)
// over a's and b's stencil (a's first): k = it >> 5 names the site (the stencil centre
// ci = k / (ncoord + 1), position q = k mod (ncoord + 1), 0: the centre itself), each site
// taken under the first of a, b that reaches it (so no row is written twice); the item adds
// the weight-row changes of a and b where they are the site itself or one of its stencil
// neighbours.  An item is done in three steps: its site (a neighbour-table load), its rows
// (pm: the offset cube in LDS), and the store.
__device__ __forceinline__ int fold_nitems(const Geo &g) { return 2 * (g.ncoord + 1) * 32; }
__device__ __forceinline__ int fold_item_site(const Geo &g, const int *__restrict__ nbr, int a, int b, int it) {
    const int nc = g.ncoord, k = it >> 5;
    const int ci = k / (nc + 1), q = k - ci * (nc + 1);
    const int xc = ci ? b : a;
    return (q == 0) ? xc : nbr[(long)xc * nc + q - 1];
}
struct FoldItem {
    int4 v[4], an[4], ao[4], bn[4], bo[4];   // the row's 16 columns; weight rows of a, b (new, old)
    int *dst;
    bool on;
};
__device__ inline void fold_item_load(const Geo &g, int pm_r, const signed char *pm, const int *w1q, int *cache,
                                      int a, int b, int na, int nb, int oa, int ob, int it, int c, FoldItem &F) {
    const int nc = g.ncoord, k = it >> 5, role = (it >> 4) & 1, ch = it & 15;
    const int ci = k / (nc + 1);
    // positions of a and b around c (the site itself: -2, outside its stencil: -1)
    const int pa = (c == a) ? -2 : stencil_pos(g, pm_r, pm, c, a);
    const int pb = (c == b) ? -2 : stencil_pos(g, pm_r, pm, c, b);
    const bool owner = !(ci == 1 && pa != -1);   // a's items own the sites both reach
    const bool ua = owner && pa != -1 && na != oa, ub = owner && pb != -1 && nb != ob;
    const int fan = ua ? (pa == -2 ? feat_own(g, role, na) : feat_nbr(g, role, pa, na)) : 0;
    const int fao = ua ? (pa == -2 ? feat_own(g, role, oa) : feat_nbr(g, role, pa, oa)) : 0;
    const int fbn = ub ? (pb == -2 ? feat_own(g, role, nb) : feat_nbr(g, role, pb, nb)) : 0;
    const int fbo = ub ? (pb == -2 ? feat_own(g, role, ob) : feat_nbr(g, role, pb, ob)) : 0;
    F.on = ua || ub;
    const long ro = ((long)role * g.natoms + c) * HID + 16 * ch;
    F.dst = cache + ro;
    const Buf bc = make_buf(cache, 4u * 2u * (uint32_t)g.natoms * HID);
    const Buf bw = make_buf(w1q, 4u * (uint32_t)(2 * nc + 2) * (uint32_t)g.nsp * HID);
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
        F.v[qq] = buf_i128(bc, F.on, 4u * (uint32_t)ro + 16u * qq);
        F.an[qq] = buf_i128(bw, ua, 4u * ((uint32_t)fan * HID + 16 * ch) + 16u * qq);
        F.ao[qq] = buf_i128(bw, ua, 4u * ((uint32_t)fao * HID + 16 * ch) + 16u * qq);
        F.bn[qq] = buf_i128(bw, ub, 4u * ((uint32_t)fbn * HID + 16 * ch) + 16u * qq);
        F.bo[qq] = buf_i128(bw, ub, 4u * ((uint32_t)fbo * HID + 16 * ch) + 16u * qq);
    }
}
__device__ __forceinline__ void fold_item_store(const FoldItem &F) {
    if (!F.on) return;
    int4 *dst = reinterpret_cast<int4 *>(F.dst);
#pragma unroll
    for (int qq = 0; qq < 4; ++qq)
        dst[qq] = make_int4(F.v[qq].x + F.an[qq].x - F.ao[qq].x + F.bn[qq].x - F.bo[qq].x,
                            F.v[qq].y + F.an[qq].y - F.ao[qq].y + F.bn[qq].y - F.bo[qq].y,
                            F.v[qq].z + F.an[qq].z - F.ao[qq].z + F.bn[qq].z - F.bo[qq].z,
                            F.v[qq].w + F.an[qq].w - F.ao[qq].w + F.bn[qq].w - F.bo[qq].w);
}

// Scoring workgroup b of a launch: slot (b & 7) + 8 m, tile j - m ntile (j = b >> 3, m = j /
// ntile), so the tiles of a slot run on the XCD of that slot's per-replica workgroups
__device__ inline void lp_block(const LPArgs &A, int b, char *lds_raw) {
    const int ntile = A.ncand / MR;
    if (b >= A.ntile_blocks) return;
    const int j = b >> 3, m = j / ntile;
    const int tile = j - m * ntile, s = (b & 7) + 8 * m;
    if (s < A.nrep) lp_tile(A, s, tile, lds_raw);
}

}  // namespace chm_lp_dev

struct chm_sampler;
namespace chm_smp {
// learned.hip: the scoring arguments of the transform launch of step n (mode 1; 2 for a
// segment's last step) or of a segment's first scoring launch after n steps (0)
chm_lp_dev::LPArgs lp_step_args(chm_sampler *S, int mode, long long n);
}  // namespace chm_smp
