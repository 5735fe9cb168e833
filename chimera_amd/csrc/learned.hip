// Learned swap proposal for the device sampler (gfx950).  No counterpart in the
// reference: its swap proposal is uniform (exafs_gpu/kernels.py:362-398).
//
// Per replica and step, `ncand` candidate pairs (a, b) are drawn from a counter-based
// stream that does not depend on the state.  Each is scored by an MLP over the one-hot
// species occupancy of a's and b's stencil neighbours and of a and b themselves
// (nocc = 2 ncoord + 2 entries, nsp classes each):
//     Z1 = b1 + X W1,  H1 = bf16(silu(Z1)),  Z2 = b2 + H1 W2,  H2 = silu(Z2),
//     g = b3 + H2 . w3          (W1, W2, w3 in bf16, every sum in FP32)
// on v_mfma_f32_16x16x32_bf16, 32 candidates per workgroup.  g predicts the energy change
// nfit (chi2(x') - chi2(x)) of the swap; logit = -lam beta g / 2 + kappa |dU| (dU: change of
// unlike first-shell bonds), -inf when a and b hold the same species.  One candidate is
// drawn from the softmax; its reverse probability comes from the same candidate set
// scored at x' (only candidates whose occupancy the swap changes are re-scored: those
// within the stencil radius of a or b), and the acceptance uses
//     x = beta dE - log(q(x'->x) / q(x->x')),
// with q summing every candidate that names the same unordered pair.  With probability
// eps (a state-independent coin) a step uses the uniform proposal instead.
//
// Training: the sampler writes a record (occupancy, dE, rung, accepted) of every swap
// it proposes; `chm_lp_train` runs Adam on minibatches of the ring with an FP32 forward
// and backward pass (same bf16 weights and H1 rounding as the scorer).
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/chimera_hip.h"
#include "common.h"
#include "sampler_dev.h"
#include "sampler_state.h"

namespace {

using namespace chm_smp;

constexpr uint64_t kSelSalt = 0x14057B7EF767814Full;
constexpr uint64_t kCandMul = 0xD6E8FEB86659FD93ull;
constexpr int HID = 256;      // hidden width
constexpr int MR = 32;        // candidate rows per MLP tile
constexpr int MAXC = 256;     // candidates per replica and step, at most
constexpr int LDH = HID + 8;  // H1 row stride in LDS (bf16), staggers the banks

struct LPW {
    const __bf16 *w1t;   // (HID, kp): row n = hidden unit
    const __bf16 *w2t;   // (HID, HID): row n = output unit
    const float *b1, *b2, *w3q, *b3;
    int kp;
};

struct LPArgs {
    Geo g;
    const uint8_t *species;
    const int *rung_of;
    const double *beta_slot;
    const uint64_t *stepc;
    const int *ss_sten;
    const int *nbr;      // (natoms, ncoord) neighbour sites in stencil order
    uint64_t seed;
    int ncand, nocc, noccp, nfeat, kp, r2max;
    double eps, lam, kappa, beta_max;
    int lift;            // lifted proposal: tilt kappa d dU with d = dir[slot]
    const int *dir;      // (nrep) +-1
    LPW W;
    float *logit;        // (nrep, ncand)
    float *gcand;        // (nrep, ncand) network output of each candidate (lifted proposal)
    int *ducand;         // (nrep, ncand) its first-shell change dU
    int *prop_next;      // (nrep, kPropInts)
    uint8_t *pend_occ;   // (nrep, noccp)
};


__device__ __forceinline__ void candidate(uint64_t h0, int i, int natoms, int &a, int &b) {
    const uint64_t h = splitmix64(h0 ^ ((uint64_t)(i + 1) * kCandMul));
    a = (int)(((h & 0xFFFFFFFFull) * (uint64_t)natoms) >> 32);
    b = (int)(((h >> 32) * (uint64_t)natoms) >> 32);
}

__device__ __forceinline__ float silu(float z) { return z / (1.f + expf(-z)); }
// the scorer's SiLU: hardware exp2 and reciprocal (a few ulp; the trainer keeps silu)
__device__ __forceinline__ float silu_fast(float z) { return z * __frcp_rn(1.f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.f / (1.f + expf(-z));
    return s * (1.f + z * (1.f - s));
}
__device__ __forceinline__ float bf16_round(float f) { return bf16_to_float(bf16_rne(f)); }
__device__ __forceinline__ float bf(__bf16 v) { return bf16_to_float(__builtin_bit_cast(unsigned short, v)); }

__device__ __forceinline__ int centred(int d, int n) {
    d = (d < 0) ? d + n : d;
    return (2 * d > n) ? d - n : d;
}
__device__ __forceinline__ int dist2(const Geo &g, const int *p, const int *q) {
    const int dx = centred(p[0] - q[0], g.box[0]);
    const int dy = centred(p[1] - q[1], g.box[1]);
    const int dz = centred(p[2] - q[2], g.box[2]);
    return dx * dx + dy * dy + dz * dz;
}

// the uniform proposal on this step: a state-independent coin, and every slot colder
// than beta_max (there the network's error would be multiplied by a large beta)
__device__ __forceinline__ bool lp_uniform(const LPArgs &A, uint64_t h0, double beta) {
    return lp_coin_uniform(h0, beta, A.eps, A.beta_max);
}

// the first-shell tilt of a candidate's logit: kappa |dU| (symmetric), or kappa d dU for a
// lifted proposal moving in direction d = +-1 (the reverse move is scored with -d)
__device__ __forceinline__ double tilt(const LPArgs &A, int du, int d) {
    return A.lift ? A.kappa * (double)(d * du) : A.kappa * fabs((double)du);
}

// site of occupancy entry q of candidate (a, b): the stencil neighbours of a, of b, then
// a and b themselves
__device__ __forceinline__ int occ_site_tbl(const LPArgs &A, int a, int b, int q) {
    const int nc = A.g.ncoord;
    if (q < nc) return A.nbr[(long)a * nc + q];
    if (q < 2 * nc) return A.nbr[(long)b * nc + q - nc];
    return q == 2 * nc ? a : b;
}

// Occupancy rows of nrows candidates (rows[r] = (a, b)) with sites xa, xb exchanged
// virtually (xa < 0: as stored) into occ (MR x noccp), the one-hot features into X
// (MR x ldx bf16, zero beyond nfeat and beyond nrows) and dU into du.  shell0[q] marks
// the first-shell stencil entries.  The neighbour-table loads of a lane are all issued
// before any is used.  Ends with the workgroup synchronised.
constexpr int IPT = 8;   // occupancy entries per lane (MR * nocc <= IPT * blockDim.x)

// idx / d for 0 <= idx < 2^20, 0 < d < 2^12 (a float reciprocal; the quotient is exact)
__device__ __forceinline__ int small_div(int idx, int d, float inv_d) {
    int r = (int)(((float)idx + 0.5f) * inv_d);
    return r;
}

// the neighbour-table loads of this lane's occupancy entries (issued, not waited for)
__device__ __forceinline__ void row_sites(const LPArgs &A, const int2 *rows, int nrows, int (&site)[IPT]) {
    const int n = nrows * A.nocc;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int idx = k * blockDim.x + threadIdx.x;
        site[k] = -1;
        if (idx < n) {
            const int r = small_div(idx, A.nocc, 1.f / (float)A.nocc), q = idx - r * A.nocc;
            site[k] = occ_site_tbl(A, rows[r].x, rows[r].y, q);
        }
    }
}

// Occupancy rows of nrows candidates (rows[r] = (a, b), sites from row_sites) with sites
// xa, xb exchanged virtually (xa < 0: as stored) into occ (MR x noccp), the one-hot
// features into X (MR x ldx bf16, zero beyond nfeat and beyond nrows) and dU into du
// (zeroed by the caller).  shell0[q] marks the first-shell stencil entries.  Ends with
// the workgroup synchronised.
__device__ void build_rows(const LPArgs &A, const uint8_t *lsp, const uint8_t *shell0, const int2 *rows,
                           int nrows, const int (&site)[IPT], int xa, int xb, uint8_t *occ, __bf16 *X,
                           int ldx, int *du) {
    const Geo &g = A.g;
    const int t = threadIdx.x;
    const int n = nrows * A.nocc;
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int idx = k * blockDim.x + t;
        if (idx >= n) continue;
        const int r = small_div(idx, A.nocc, 1.f / (float)A.nocc), q = idx - r * A.nocc;
        const int a = rows[r].x, b = rows[r].y;
        const int sp = species_swapped(lsp, site[k], xa, xb);
        occ[r * A.noccp + q] = (uint8_t)sp;
        if (q < 2 * g.ncoord) {
            const bool on_a = q < g.ncoord;
            if (shell0[on_a ? q : q - g.ncoord]) {
                const int sa = species_swapped(lsp, a, xa, xb), sb = species_swapped(lsp, b, xa, xb);
                int cnt = 0;
                if (on_a && site[k] != b) cnt = (sp == sa) - (sp == sb);
                if (!on_a && site[k] != a) cnt = (sp == sb) - (sp == sa);
                if (cnt) atomicAdd(&du[r], cnt);
            }
        }
    }
    __syncthreads();
    const int nchunk = A.kp / 8;
    for (int idx = t; idx < MR * nchunk; idx += blockDim.x) {
        const int r = idx / nchunk, c = idx - r * nchunk;
        bf16x8 v;
        int q = (8 * c) / g.nsp, s = (8 * c) - q * g.nsp;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * c + i;
            const bool on = r < nrows && k < A.nfeat && occ[r * A.noccp + q] == s;
            v[i] = __builtin_bit_cast(__bf16, (unsigned short)(on ? 0x3F80 : 0));
            if (++s == g.nsp) {
                s = 0;
                ++q;
            }
        }
        *reinterpret_cast<bf16x8 *>(X + r * ldx + 8 * c) = v;
    }
    __syncthreads();
}

// first-shell flags of the stencil entries into LDS (at most 256 entries)
__device__ __forceinline__ void load_shell0(const LPArgs &A, uint8_t *shell0) {
    for (int q = threadIdx.x; q < A.g.ncoord; q += blockDim.x) shell0[q] = ((A.ss_sten[q] >> 24) & 0xFF) == 0;
}

// This lane's share of the network: 8 waves x 32 hidden units (two 16-column tiles).  The
// weight fragments are loaded once, into registers, at the start of a kernel, so their
// latency overlaps the occupancy gathers.
constexpr int NT = 512;    // threads per workgroup of the scorer and the selection
constexpr int KT1 = 12;    // layer-1 K tiles held in registers (kp <= 384 in one pass)
struct Frag {
    bf16x8 w1[KT1][2], w2[HID / 32][2];
    float b1[2], b2[2], w3[2];
};

__device__ __forceinline__ void load_frag(const LPW &W, Frag &F) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int lr = lane & 15, lk = 8 * (lane >> 4);
    const int nk1 = W.kp / 32;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int n = 32 * wave + 16 * c + lr;
        const __bf16 *w1 = W.w1t + (long)n * W.kp + lk;
        const __bf16 *w2 = W.w2t + (long)n * HID + lk;
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt)
            if (kt < nk1) F.w1[kt][c] = *reinterpret_cast<const bf16x8 *>(w1 + 32 * kt);
#pragma unroll
        for (int kt = 0; kt < HID / 32; ++kt) F.w2[kt][c] = *reinterpret_cast<const bf16x8 *>(w2 + 32 * kt);
        F.b1[c] = W.b1[n];
        F.b2[c] = W.b2[n];
        F.w3[c] = W.w3q[n];
    }
}

// g of the MR rows whose features are in X (LDS); H (MR x LDH bf16) and gp (8 x MR) are LDS
// scratch; result in gout (LDS).  Every thread of the (NT-thread) workgroup calls it.
// With F the weight fragments come from registers (loaded earlier by load_frag);
// without, each K tile's fragments are loaded one tile ahead (fewer registers).
template <bool kHeld>
__device__ void mlp_rows(const LPW &W, const Frag *F, const __bf16 *X, int ldx, __bf16 *H, float *gp,
                         float *gout) {
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63;
    const int lr = lane & 15, lk = 8 * (lane >> 4);
    const int n0 = 32 * wave + lr, n1 = n0 + 16;
    const __bf16 *w1a = W.w1t + (long)n0 * W.kp + lk, *w1b = W.w1t + (long)n1 * W.kp + lk;
    const __bf16 *w2a = W.w2t + (long)n0 * HID + lk, *w2b = W.w2t + (long)n1 * HID + lk;
    float b1[2], b2[2], w3[2];
    if (kHeld) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            b1[c] = F->b1[c];
            b2[c] = F->b2[c];
            w3[c] = F->w3[c];
        }
    } else {
        b1[0] = W.b1[n0];
        b1[1] = W.b1[n1];
        b2[0] = W.b2[n0];
        b2[1] = W.b2[n1];
        w3[0] = W.w3q[n0];
        w3[1] = W.w3q[n1];
    }
    f32x4 acc[2][2];
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nk1 = W.kp / 32;
    auto step = [&](const bf16x8 &b0, const bf16x8 &b1v, int kt) {
        const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(X + lr * ldx + 32 * kt + lk);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(X + (16 + lr) * ldx + 32 * kt + lk);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1v, acc[0][1], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1v, acc[1][1], 0, 0, 0);
    };
    bf16x8 nb0, nb1;
    if (kHeld) {
#pragma unroll
        for (int kt = 0; kt < KT1; ++kt)
            if (kt < nk1) step(F->w1[kt][0], F->w1[kt][1], kt);
        for (int kt = KT1; kt < nk1; ++kt)   // features beyond the register-held tiles
            step(*reinterpret_cast<const bf16x8 *>(w1a + 32 * kt),
                 *reinterpret_cast<const bf16x8 *>(w1b + 32 * kt), kt);
    } else {
        nb0 = *reinterpret_cast<const bf16x8 *>(w1a);
        nb1 = *reinterpret_cast<const bf16x8 *>(w1b);
        for (int kt = 0; kt < nk1; ++kt) {
            const bf16x8 b0 = nb0, b1v = nb1;
            if (kt + 1 < nk1) {
                nb0 = *reinterpret_cast<const bf16x8 *>(w1a + 32 * (kt + 1));
                nb1 = *reinterpret_cast<const bf16x8 *>(w1b + 32 * (kt + 1));
            }
            step(b0, b1v, kt);
        }
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int n = c ? n1 : n0;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int row = 16 * r + 4 * (lane >> 4) + i;
                H[row * LDH + n] = __builtin_bit_cast(__bf16, bf16_rne(silu_fast(acc[r][c][i] + b1[c])));
            }
            acc[r][c] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    __syncthreads();
    if (!kHeld) {
        nb0 = *reinterpret_cast<const bf16x8 *>(w2a);
        nb1 = *reinterpret_cast<const bf16x8 *>(w2b);
    }
#pragma unroll
    for (int kt = 0; kt < HID / 32; ++kt) {
        bf16x8 b0, b1v;
        if (kHeld) {
            b0 = F->w2[kt][0];
            b1v = F->w2[kt][1];
        } else {
            b0 = nb0;
            b1v = nb1;
            if (kt + 1 < HID / 32) {
                nb0 = *reinterpret_cast<const bf16x8 *>(w2a + 32 * (kt + 1));
                nb1 = *reinterpret_cast<const bf16x8 *>(w2b + 32 * (kt + 1));
            }
        }
        const bf16x8 a0 = *reinterpret_cast<const bf16x8 *>(H + lr * LDH + 32 * kt + lk);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8 *>(H + (16 + lr) * LDH + 32 * kt + lk);
        acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b0, acc[0][0], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b0, acc[1][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, b1v, acc[0][1], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, b1v, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 2; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v = silu_fast(acc[r][0][i] + b2[0]) * w3[0] + silu_fast(acc[r][1][i] + b2[1]) * w3[1];
#pragma unroll
            for (int m = 8; m >= 1; m >>= 1) v += __shfl_xor(v, m, 16);
            if (lr == 0) gp[wave * MR + 16 * r + 4 * (lane >> 4) + i] = v;
        }
    __syncthreads();
    if (t < MR) {
        float v = 0.f;
#pragma unroll
        for (int w = 0; w < NT / 64; ++w) v += gp[w * MR + t];
        gout[t] = W.b3[0] + v;
    }
    __syncthreads();
}

struct RowLds {
    uint8_t *lsp, *occ;
    __bf16 *X, *H;
    int ldx;
};

__device__ RowLds carve(const LPArgs &A, char *base) {
    RowLds L;
    char *p = base;
    L.lsp = reinterpret_cast<uint8_t *>(p);
    p += (A.g.natoms_pad + 15) & ~15;
    L.ldx = A.kp + 8;
    L.X = reinterpret_cast<__bf16 *>(p);
    p += (size_t)MR * L.ldx * 2;
    L.H = reinterpret_cast<__bf16 *>(p);
    p += (size_t)MR * LDH * 2;
    L.occ = reinterpret_cast<uint8_t *>(p);
    return L;
}

// ------------------------------------------------------------------ scoring
// grid (ncand / MR, nrep), NT threads: logits of the candidates of step stepc + 1 at the
// current state
__global__ __launch_bounds__(NT) void k_lp_score(LPArgs A) {
    extern __shared__ __align__(16) char lds_raw[];
    __shared__ int2 rows[MR];
    __shared__ int du[MR];
    __shared__ float gp[(NT / 64) * MR], gout[MR];
    __shared__ uint8_t shell0[256];
    const Geo &g = A.g;
    const int s = blockIdx.y, t = threadIdx.x;
    const uint64_t n1 = A.stepc[s] + 1;
    const int rung = A.rung_of[s];
    const double beta = A.beta_slot[s];
    const uint64_t h0 = lp_stream(A.seed, n1, rung);
    if (lp_uniform(A, h0, beta)) return;   // a uniform step: nothing to score
    RowLds L = carve(A, lds_raw);
    load_species(L.lsp, A.species + (long)s * g.natoms_pad, g.natoms_pad);
    load_shell0(A, shell0);
    const int i0 = blockIdx.x * MR;
    if (t < MR) {
        int a, b;
        candidate(h0, i0 + t, g.natoms, a, b);
        rows[t] = make_int2(a, b);
        du[t] = 0;
    }
    __syncthreads();
    // neighbour loads first, then the weights: the occupancy work waits only for the former
    int site[IPT];
    row_sites(A, rows, MR, site);
    Frag F;
    load_frag(A.W, F);
    build_rows(A, L.lsp, shell0, rows, MR, site, -1, -1, L.occ, L.X, L.ldx, du);
    mlp_rows<true>(A.W, &F, L.X, L.ldx, L.H, gp, gout);
    if (t < MR) {
        const bool valid = L.lsp[rows[t].x] != L.lsp[rows[t].y];
        const float lg = (float)(-A.lam * beta * 0.5 * (double)gout[t] + tilt(A, du[t], A.lift ? A.dir[s] : 1));
        A.logit[(long)s * A.ncand + i0 + t] = valid ? lg : -__builtin_huge_valf();
        if (A.lift) {
            A.gcand[(long)s * A.ncand + i0 + t] = gout[t];
            A.ducand[(long)s * A.ncand + i0 + t] = du[t];
        }
    }
}

// ------------------------------------------------------------------ selection
constexpr int NW = NT / 64;

__device__ __forceinline__ float block_max_f(float v, float *red) {
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float r = red[0];
    for (int w = 1; w < NW; ++w) r = fmaxf(r, red[w]);
    __syncthreads();
    return r;
}

__device__ __forceinline__ double block_sum_d(double v, double *red) {
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = red[0];
    for (int w = 1; w < NW; ++w) r += red[w];
    __syncthreads();
    return r;
}

// inclusive scan over the workgroup's lanes (index order); returns (exclusive, inclusive)
__device__ __forceinline__ void block_scan_d(double v, double *red, double &excl, double &incl) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double x = v;
    for (int m = 1; m < 64; m <<= 1) {
        const double y = __shfl_up(x, m, 64);
        if (lane >= m) x += y;
    }
    __syncthreads();
    if (lane == 63) red[wave] = x;
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < wave; ++w) off += red[w];
    incl = off + x;
    const double prev = __shfl_up(x, 1, 64);
    excl = off + (lane == 0 ? 0.0 : prev);
    __syncthreads();
}

// grid (1, nrep), NT threads: the proposal of step stepc + 1 (learned, or uniform on the
// coin / above beta_max), its log q ratio and its occupancy record; the slot kernel then
// gathers its count changes
__global__ __launch_bounds__(NT) void k_lp_select(LPArgs A) {
    extern __shared__ __align__(16) char lds_raw[];
    __shared__ int2 rows[MR];
    __shared__ int du[MR];
    __shared__ float gp[NW * MR], gout[MR];
    __shared__ float lrev[MAXC];
    __shared__ int alist[MAXC];
    __shared__ int nalist, jsel;
    __shared__ double redd[NW];
    __shared__ float redf[NW];
    __shared__ uint8_t shell0[256];
    __shared__ Prop P;
    const Geo &g = A.g;
    const int s = blockIdx.y, t = threadIdx.x;
    RowLds L = carve(A, lds_raw);
    const uint64_t n1 = A.stepc[s] + 1;
    const int rung = A.rung_of[s];
    const double beta = A.beta_slot[s];
    const uint64_t h0 = lp_stream(A.seed, n1, rung);
    const bool uniform = lp_uniform(A, h0, beta);
    const float lg = (!uniform && t < A.ncand) ? A.logit[(long)s * A.ncand + t] : -__builtin_huge_valf();
    load_species(L.lsp, A.species + (long)s * g.natoms_pad, g.natoms_pad);
    load_shell0(A, shell0);
    if (t == 0) jsel = -1;
    __syncthreads();
    float logqr = 0.f;
    if (uniform) {
        if (t == 0) P = propose(L.lsp, g.natoms, A.seed, n1, (uint64_t)rung);
        __syncthreads();
    } else {
        const float m = block_max_f(lg, redf);
        const bool any = m > -__builtin_huge_valf();
        const double p = (any && lg > -__builtin_huge_valf()) ? exp((double)(lg - m)) : 0.0;
        double excl, incl;
        block_scan_d(p, redd, excl, incl);
        const double Z = block_sum_d(p, redd);
        const double target = u01(splitmix64(h0 ^ kSelSalt)) * Z;
        if (any && p > 0.0 && excl <= target && target < incl) jsel = t;
        __syncthreads();
        if (any && jsel < 0) {   // target rounded past the last bin: the last candidate with p > 0
            if (p > 0.0) atomicMax(&jsel, t);
            __syncthreads();
        }
        const int j = jsel;
        if (t == 0) {
            if (j < 0) {
                P = Prop{0, 0, L.lsp[0], L.lsp[0], 0};
            } else {
                int a, b;
                candidate(h0, j, g.natoms, a, b);
                P = Prop{a, b, L.lsp[b], L.lsp[a], 1};
            }
        }
        __syncthreads();
        if (j >= 0) {
            // Hastings ratio: the candidate set scored at x' = swap(x, a, b); only the
            // candidates within the stencil radius of a or b change their occupancy
            const int a = P.a, b = P.b;
            int pj[3], qj[3];
            site_xyz_of(g, a, pj[0], pj[1], pj[2]);
            site_xyz_of(g, b, qj[0], qj[1], qj[2]);
            if (t == 0) nalist = 0;
            bool same = false;
            // a candidate the swap does not touch keeps its network output and dU; its
            // reverse logit differs from the forward one only by the lifted tilt's sign
            // (the same expression as the scorer's at the proposed state with -d)
            float lu = lg;
            if (A.lift && t < A.ncand && lg > -__builtin_huge_valf())
                lu = (float)(-A.lam * beta * 0.5 * (double)A.gcand[(long)s * A.ncand + t] +
                             tilt(A, A.ducand[(long)s * A.ncand + t], -A.dir[s]));
            if (t < MAXC) lrev[t] = lu;
            __syncthreads();
            if (t < A.ncand) {
                int ca, cb, pc[3], qc[3];
                candidate(h0, t, g.natoms, ca, cb);
                same = (ca == a && cb == b) || (ca == b && cb == a);
                site_xyz_of(g, ca, pc[0], pc[1], pc[2]);
                site_xyz_of(g, cb, qc[0], qc[1], qc[2]);
                const bool touched = dist2(g, pc, pj) <= A.r2max || dist2(g, pc, qj) <= A.r2max ||
                                     dist2(g, qc, pj) <= A.r2max || dist2(g, qc, qj) <= A.r2max;
                if (touched) alist[atomicAdd(&nalist, 1)] = t;
            }
            __syncthreads();
            const int na = nalist;
            for (int c0 = 0; c0 < na; c0 += MR) {
                const int nr = min(MR, na - c0);
                if (t < nr) {
                    int ca, cb;
                    candidate(h0, alist[c0 + t], g.natoms, ca, cb);
                    rows[t] = make_int2(ca, cb);
                }
                if (t < MR) du[t] = 0;
                __syncthreads();
                int site[IPT];
                row_sites(A, rows, nr, site);
                build_rows(A, L.lsp, shell0, rows, nr, site, a, b, L.occ, L.X, L.ldx, du);
                mlp_rows<false>(A.W, nullptr, L.X, L.ldx, L.H, gp, gout);
                if (t < nr) {
                    const int ca = rows[t].x, cb = rows[t].y;
                    const bool valid = species_swapped(L.lsp, ca, a, b) != species_swapped(L.lsp, cb, a, b);
                    const float l2 = (float)(-A.lam * beta * 0.5 * (double)gout[t] +
                                             tilt(A, du[t], A.lift ? -A.dir[s] : 1));
                    lrev[alist[c0 + t]] = valid ? l2 : -__builtin_huge_valf();
                }
                __syncthreads();
            }
            const float lr2 = (t < A.ncand) ? lrev[t] : -__builtin_huge_valf();
            const float m2 = block_max_f(lr2, redf);
            const double p2 = (lr2 > -__builtin_huge_valf()) ? exp((double)(lr2 - m2)) : 0.0;
            const double Z2 = block_sum_d(p2, redd);
            const double D1 = block_sum_d(same ? p : 0.0, redd);
            const double D2 = block_sum_d(same ? p2 : 0.0, redd);
            logqr = (float)((log(D2) - log(Z2)) - (log(D1) - log(Z)));
        }
    }
    const Prop p = P;
    if (t == 0) {
        int *o = A.prop_next + kPropInts * s;
        o[0] = p.a;
        o[1] = p.b;
        o[2] = p.na;
        o[3] = p.nb;
        o[4] = p.sw;
        o[5] = __float_as_int(logqr);
    }
    // occupancy of the proposal at x, for its move record
    if (A.pend_occ && p.sw)
        for (int q = t; q < A.nocc; q += blockDim.x)
            A.pend_occ[(long)s * A.noccp + q] = L.lsp[occ_site_tbl(A, p.a, p.b, q)];
}

// ------------------------------------------------------------------ evaluation (tests)
// grid ceil(n / MR), NT threads: g of given occupancy rows through the scorer's MFMA path
__global__ __launch_bounds__(NT) void k_lp_eval(LPArgs A, int n, const uint8_t *__restrict__ occ_in,
                                                float *__restrict__ out) {
    extern __shared__ __align__(16) char lds_raw[];
    __shared__ float gp[NW * MR], gout[MR];
    Frag F;
    load_frag(A.W, F);
    RowLds L = carve(A, lds_raw);
    const int r0 = blockIdx.x * MR, nr = min(MR, n - r0), t = threadIdx.x;
    for (int idx = t; idx < MR * A.noccp; idx += blockDim.x) {
        const int r = idx / A.noccp, q = idx - r * A.noccp;
        L.occ[idx] = (r < nr && q < A.nocc) ? occ_in[(long)(r0 + r) * A.nocc + q] : 0;
    }
    __syncthreads();
    const int nchunk = A.kp / 8;
    for (int idx = t; idx < MR * nchunk; idx += blockDim.x) {
        const int r = idx / nchunk, c = idx - r * nchunk;
        bf16x8 v;
        int q = (8 * c) / A.g.nsp, s = (8 * c) - q * A.g.nsp;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int k = 8 * c + i;
            const bool on = r < nr && k < A.nfeat && L.occ[r * A.noccp + q] == s;
            v[i] = __builtin_bit_cast(__bf16, (unsigned short)(on ? 0x3F80 : 0));
            if (++s == A.g.nsp) {
                s = 0;
                ++q;
            }
        }
        *reinterpret_cast<bf16x8 *>(L.X + r * L.ldx + 8 * c) = v;
    }
    __syncthreads();
    mlp_rows<true>(A.W, &F, L.X, L.ldx, L.H, gp, gout);
    if (t < nr) out[r0 + t] = gout[t];
}

// ------------------------------------------------------------------ training
struct TrainArgs {
    int B, nocc, noccp, nsp, kp;
    const uint8_t *occ;     // (B, noccp)
    const float *y, *w;     // (B)
    const int *wsum;        // number of rows with w = 1 (weights are 0 / 1)
    const __bf16 *w1;       // (kp, HID) feature-major
    const __bf16 *w2;       // (HID, HID) input-major
    const __bf16 *w2t;      // (HID, HID) output-major
    const float *b1, *b2, *w3q, *b3;
    float *Z1, *H1, *Z2, *H2, *gval, *dZ2, *dZ1;   // (B, HID) / (B)
    float *grad;            // flat, same layout as the parameters
    float *loss;
    int off_b1, off_w2, off_b2, off_w3, off_b3;
};

constexpr int TR = 8;   // rows per workgroup in the row kernels

// forward, FP32: thread = hidden unit, TR rows per workgroup
__global__ __launch_bounds__(256) void k_lp_fwd_train(TrainArgs T) {
    __shared__ float h1s[TR][HID];
    __shared__ float red[TR][4];
    const int n = threadIdx.x, r0 = blockIdx.x * TR;
    for (int r = 0; r < TR; ++r) {
        const int row = r0 + r;
        float z = 0.f;
        if (row < T.B) {
            z = T.b1[n];
            const uint8_t *o = T.occ + (long)row * T.noccp;
            for (int q = 0; q < T.nocc; ++q) {
                const int sq = o[q] < T.nsp ? o[q] : 0;   // (rows are validated; never index past W1)
                z += bf(T.w1[(long)(T.nsp * q + sq) * HID + n]);
            }
            T.Z1[(long)row * HID + n] = z;
        }
        const float h = bf16_round(silu(z));
        h1s[r][n] = h;
        if (row < T.B) T.H1[(long)row * HID + n] = h;
    }
    __syncthreads();
    float z2[TR];
#pragma unroll
    for (int r = 0; r < TR; ++r) z2[r] = T.b2[n];
    for (int k = 0; k < HID; ++k) {
        const float wv = bf(T.w2[(long)k * HID + n]);
#pragma unroll
        for (int r = 0; r < TR; ++r) z2[r] += h1s[r][k] * wv;
    }
    const float w3 = T.w3q[n];
    for (int r = 0; r < TR; ++r) {
        const int row = r0 + r;
        const float h2 = silu(z2[r]);
        if (row < T.B) {
            T.Z2[(long)row * HID + n] = z2[r];
            T.H2[(long)row * HID + n] = h2;
        }
        float v = h2 * w3;
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
        if ((n & 63) == 0) red[r][n >> 6] = v;
    }
    __syncthreads();
    if (n < TR && r0 + n < T.B) T.gval[r0 + n] = T.b3[0] + ((red[n][0] + red[n][1]) + (red[n][2] + red[n][3]));
}

// per row: dZ2 = dg w3 silu'(Z2), dH1 = dZ2 W2^T, dZ1 = dH1 silu'(Z1); loss
__global__ __launch_bounds__(256) void k_lp_bwd_rows(TrainArgs T) {
    __shared__ float dz2s[TR][HID];
    const int n = threadIdx.x, r0 = blockIdx.x * TR;
    const float W = fmaxf((float)T.wsum[0], 1.f);
    const float w3 = T.w3q[n];
    for (int r = 0; r < TR; ++r) {
        const int row = r0 + r;
        float d = 0.f;
        if (row < T.B) {
            const float res = T.gval[row] - T.y[row];
            const float dg = T.w[row] * res / W;
            d = dg * w3 * dsilu(T.Z2[(long)row * HID + n]);
            T.dZ2[(long)row * HID + n] = d;
            if (n == 0 && T.w[row] != 0.f) atomicAdd(T.loss, 0.5f * T.w[row] * res * res / W);
        }
        dz2s[r][n] = d;
    }
    __syncthreads();
    float dh[TR];
#pragma unroll
    for (int r = 0; r < TR; ++r) dh[r] = 0.f;
    // thread n is input unit k = n here: dH1[k] = sum_o dZ2[o] W2[k][o]
    for (int o = 0; o < HID; ++o) {
        const float wv = bf(T.w2[(long)n * HID + o]);
#pragma unroll
        for (int r = 0; r < TR; ++r) dh[r] += dz2s[r][o] * wv;
    }
    for (int r = 0; r < TR; ++r) {
        const int row = r0 + r;
        if (row < T.B) T.dZ1[(long)row * HID + n] = dh[r] * dsilu(T.Z1[(long)row * HID + n]);
    }
}

// weight gradients, reduced over the batch in a fixed order: workgroup f < kp: dW1 row f;
// kp + k: dW2 row k; kp + HID: b1, b2, w3, b3
__global__ __launch_bounds__(256) void k_lp_bwd_w(TrainArgs T) {
    const int n = threadIdx.x, f = blockIdx.x;
    if (f < T.kp) {
        const int q = f / T.nsp, s = f - q * T.nsp;
        float acc = 0.f;
        if (q < T.nocc)
            for (int r = 0; r < T.B; ++r)
                if (T.occ[(long)r * T.noccp + q] == s) acc += T.dZ1[(long)r * HID + n];
        T.grad[(long)f * HID + n] = acc;
    } else if (f < T.kp + HID) {
        const int k = f - T.kp;
        float acc = 0.f;
        for (int r = 0; r < T.B; ++r) acc += T.H1[(long)r * HID + k] * T.dZ2[(long)r * HID + n];
        T.grad[T.off_w2 + (long)k * HID + n] = acc;
    } else {
        const float W = fmaxf((float)T.wsum[0], 1.f);
        float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f;
        for (int r = 0; r < T.B; ++r) {
            const float dg = T.w[r] * (T.gval[r] - T.y[r]) / W;
            gb1 += T.dZ1[(long)r * HID + n];
            gb2 += T.dZ2[(long)r * HID + n];
            gw3 += dg * T.H2[(long)r * HID + n];
            gb3 += dg;
        }
        T.grad[T.off_b1 + n] = gb1;
        T.grad[T.off_b2 + n] = gb2;
        T.grad[T.off_w3 + n] = gw3;
        if (n == 0) T.grad[T.off_b3] = gb3;
    }
}

// minibatch of the record ring: row r takes record h(seed, it, r) mod nvalid; weight 1
// for a swap of a rung <= max_rung, else 0; targets clipped to [-clip, clip]
__global__ void k_lp_batch(int B, long long nvalid, int noccp, uint64_t seed, long long it, int max_rung,
                           float clip, const uint8_t *__restrict__ rec_occ, const float *__restrict__ rec_y,
                           const int *__restrict__ rec_meta, uint8_t *__restrict__ occ, float *__restrict__ y,
                           float *__restrict__ w, int *__restrict__ wsum) {
    const int r = blockIdx.x, t = threadIdx.x;
    const uint64_t h = splitmix64(seed ^ ((uint64_t)it * kGolden) ^ ((uint64_t)(r + 1) << 40) ^ kLpSalt);
    const long long idx = (long long)(h % (uint64_t)nvalid);
    // a null move's entry has no occupancy: its row gets species 0 (and weight 0)
    const bool swap = rec_meta[idx] >= 0;
    for (int q = t; q < noccp; q += blockDim.x) occ[(long)r * noccp + q] = swap ? rec_occ[idx * noccp + q] : 0;
    if (t == 0) {
        const int meta = rec_meta[idx];
        const bool use = meta >= 0 && (meta & 0xFFFF) <= max_rung;
        w[r] = use ? 1.f : 0.f;
        y[r] = fminf(fmaxf(rec_y[idx], -clip), clip);
        if (use) atomicAdd(wsum, 1);
    }
}

__global__ void k_lp_adam(int n, float *__restrict__ p, const float *__restrict__ gr, float *__restrict__ m,
                          float *__restrict__ v, float lr, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = gr[i];
    const float mi = 0.9f * m[i] + 0.1f * gi;
    const float vi = 0.999f * v[i] + 0.001f * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
}

// bf16 copies of W1 (both orders), W2 (both orders) and the rounded w3
__global__ void k_lp_quant(int kp, const float *__restrict__ p, int off_w2, int off_w3, __bf16 *__restrict__ w1,
                           __bf16 *__restrict__ w1t, __bf16 *__restrict__ w2, __bf16 *__restrict__ w2t,
                           float *__restrict__ w3q) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n1 = (long)kp * HID, n2 = (long)HID * HID;
    if (i < n1) {
        const int f = (int)(i / HID), n = (int)(i % HID);
        const __bf16 v = __builtin_bit_cast(__bf16, bf16_rne(p[i]));
        w1[i] = v;
        w1t[(long)n * kp + f] = v;
    } else if (i < n1 + n2) {
        const long j = i - n1;
        const int k = (int)(j / HID), n = (int)(j % HID);
        const __bf16 v = __builtin_bit_cast(__bf16, bf16_rne(p[off_w2 + j]));
        w2[j] = v;
        w2t[(long)n * HID + k] = v;
    } else if (i < n1 + n2 + HID) {
        const int n = (int)(i - n1 - n2);
        w3q[n] = bf16_round(p[off_w3 + n]);
    }
}

// ------------------------------------------------------------------ host helpers
struct Offs {
    int w1, b1, w2, b2, w3, b3, n;
};
Offs offsets(int kp) {
    Offs o;
    o.w1 = 0;
    o.b1 = kp * HID;
    o.w2 = o.b1 + HID;
    o.b2 = o.w2 + HID * HID;
    o.w3 = o.b2 + HID;
    o.b3 = o.w3 + HID;
    o.n = o.b3 + 1;
    return o;
}

LPArgs lp_args(chm_sampler *S) {
    const chm_lp &L = S->lp;
    LPArgs A{};
    A.g = S->g;
    A.species = S->species.as<uint8_t>();
    A.rung_of = S->rung_of.as<int>();
    A.beta_slot = S->beta_slot.as<double>();
    A.stepc = S->stepc.as<uint64_t>();
    A.ss_sten = S->ss_sten.as<int>();
    A.nbr = S->nbr_st.as<int>();
    A.seed = S->seed;
    A.ncand = L.ncand;
    A.nocc = L.nocc;
    A.noccp = L.noccp;
    A.nfeat = L.nfeat;
    A.kp = L.kp;
    A.r2max = L.r2max;
    A.eps = L.eps;
    A.lam = L.lam;
    A.kappa = L.kappa;
    A.beta_max = L.beta_max;
    A.lift = L.lift;
    A.dir = L.dir.as<int>();
    const Offs o = offsets(L.kp);
    A.W.w1t = L.w1t.as<__bf16>();
    A.W.w2t = L.w2t.as<__bf16>();
    A.W.b1 = L.params.as<float>() + o.b1;
    A.W.b2 = L.params.as<float>() + o.b2;
    A.W.b3 = L.params.as<float>() + o.b3;
    A.W.w3q = L.w3q.as<float>();
    A.W.kp = L.kp;
    A.logit = L.logit.as<float>();
    A.gcand = L.gcand.as<float>();
    A.ducand = L.ducand.as<int>();
    A.pend_occ = L.pend_occ.as<uint8_t>();
    return A;
}

size_t lds_rows(const chm_sampler *S) {
    const chm_lp &L = S->lp;
    return (size_t)((S->g.natoms_pad + 15) & ~15) + (size_t)MR * (L.kp + 8) * 2 + (size_t)MR * LDH * 2 +
           (size_t)MR * L.noccp;
}

int quantise(chm_sampler *S) {
    chm_lp &L = S->lp;
    const Offs o = offsets(L.kp);
    const long n = (long)L.kp * HID + HID * HID + HID;
    k_lp_quant<<<(unsigned)((n + 255) / 256), 256, 0, S->ctx->stream>>>(
        L.kp, L.params.as<float>(), o.w2, o.w3, L.w1.as<__bf16>(), L.w1t.as<__bf16>(), L.w2.as<__bf16>(),
        L.w2t.as<__bf16>(), L.w3q.as<float>());
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int ensure_train(chm_sampler *S, int B) {
    chm_lp &L = S->lp;
    if (B <= L.tr_cap) return CHM_OK;
    int rc;
    const size_t b = (size_t)B;
    if ((rc = L.tr_occ.ensure(b * L.noccp)) || (rc = L.tr_y.ensure(4 * b)) || (rc = L.tr_w.ensure(4 * b)) ||
        (rc = L.tr_wsum.ensure(4)) || (rc = L.tr_Z1.ensure(4 * b * HID)) || (rc = L.tr_H1.ensure(4 * b * HID)) ||
        (rc = L.tr_Z2.ensure(4 * b * HID)) || (rc = L.tr_H2.ensure(4 * b * HID)) ||
        (rc = L.tr_g.ensure(4 * b)) || (rc = L.tr_dZ2.ensure(4 * b * HID)) ||
        (rc = L.tr_dZ1.ensure(4 * b * HID)) || (rc = L.grad.ensure(4 * (size_t)L.nparam)) ||
        (rc = L.loss.ensure(4)))
        return rc;
    L.tr_cap = B;
    return CHM_OK;
}

TrainArgs train_args(chm_sampler *S, int B) {
    chm_lp &L = S->lp;
    const Offs o = offsets(L.kp);
    TrainArgs T{};
    T.B = B;
    T.nocc = L.nocc;
    T.noccp = L.noccp;
    T.nsp = S->g.nsp;
    T.kp = L.kp;
    T.occ = L.tr_occ.as<uint8_t>();
    T.y = L.tr_y.as<float>();
    T.w = L.tr_w.as<float>();
    T.wsum = L.tr_wsum.as<int>();
    T.w1 = L.w1.as<__bf16>();
    T.w2 = L.w2.as<__bf16>();
    T.w2t = L.w2t.as<__bf16>();
    const float *p = L.params.as<float>();
    T.b1 = p + o.b1;
    T.b2 = p + o.b2;
    T.b3 = p + o.b3;
    T.w3q = L.w3q.as<float>();
    T.Z1 = L.tr_Z1.as<float>();
    T.H1 = L.tr_H1.as<float>();
    T.Z2 = L.tr_Z2.as<float>();
    T.H2 = L.tr_H2.as<float>();
    T.gval = L.tr_g.as<float>();
    T.dZ2 = L.tr_dZ2.as<float>();
    T.dZ1 = L.tr_dZ1.as<float>();
    T.grad = L.grad.as<float>();
    T.loss = L.loss.as<float>();
    T.off_b1 = o.b1;
    T.off_w2 = o.w2;
    T.off_b2 = o.b2;
    T.off_w3 = o.w3;
    T.off_b3 = o.b3;
    return T;
}

// forward + backward of the batch already in tr_occ / tr_y / tr_w / tr_wsum; gradient in L.grad
int fwd_bwd(chm_sampler *S, int B) {
    hipStream_t st = S->ctx->stream;
    TrainArgs T = train_args(S, B);
    const unsigned nb = (unsigned)((B + TR - 1) / TR);
    k_lp_fwd_train<<<nb, HID, 0, st>>>(T);
    CHM_LAUNCH_CHECK();
    k_lp_bwd_rows<<<nb, HID, 0, st>>>(T);
    CHM_LAUNCH_CHECK();
    k_lp_bwd_w<<<(unsigned)(S->lp.kp + HID + 1), HID, 0, st>>>(T);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

}  // namespace

namespace chm_smp {

int lp_prepare(chm_sampler *S) {
    chm_lp &L = S->lp;
    const Geo &g = S->g;
    L.nocc = 2 * g.ncoord + 2;
    L.noccp = (L.nocc + 15) & ~15;
    L.nfeat = g.nsp * L.nocc;
    L.kp = (L.nfeat + 31) & ~31;
    L.nparam = offsets(L.kp).n;
    if (L.kp > 1024 || g.ncoord > 256)
        return chm::set_error("learned proposal: too many occupancy features"), CHM_E_ARG;
    int rc;
    if ((rc = L.params.ensure(4 * (size_t)L.nparam)) || (rc = L.adam_m.ensure(4 * (size_t)L.nparam)) ||
        (rc = L.adam_v.ensure(4 * (size_t)L.nparam)) || (rc = L.w1t.ensure(2 * (size_t)L.kp * HID)) ||
        (rc = L.w1.ensure(2 * (size_t)L.kp * HID)) || (rc = L.w2t.ensure(2 * (size_t)HID * HID)) ||
        (rc = L.w2.ensure(2 * (size_t)HID * HID)) || (rc = L.w3q.ensure(4 * HID)) ||
        (rc = L.logit.ensure(4 * (size_t)S->nrep * MAXC)) ||
        (rc = L.gcand.ensure(4 * (size_t)S->nrep * MAXC)) ||
        (rc = L.ducand.ensure(4 * (size_t)S->nrep * MAXC)) ||
        (rc = L.pend_occ.ensure((size_t)S->nrep * L.noccp)))
        return rc;
    CHM_HIP(hipMemsetAsync(L.pend_occ.p, 0, L.pend_occ.bytes, S->ctx->stream));
    if (lds_rows(S) > 140 * 1024) return chm::set_error("learned proposal: LDS too small"), CHM_E_ARG;
    if (MR * L.nocc > IPT * NT) return chm::set_error("learned proposal: stencil too large"), CHM_E_ARG;
    hipFuncSetAttribute((const void *)k_lp_score, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_rows(S));
    hipFuncSetAttribute((const void *)k_lp_eval, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_rows(S));
    hipFuncSetAttribute((const void *)k_lp_select, hipFuncAttributeMaxDynamicSharedMemorySize,
                        (int)lds_rows(S));
    return CHM_OK;
}

int lp_launch_score(chm_sampler *S) {
    LPArgs A = lp_args(S);
    k_lp_score<<<dim3(S->lp.ncand / MR, S->nrep), NT, lds_rows(S), S->ctx->stream>>>(A);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

int lp_launch_select(chm_sampler *S, long long n) {
    LPArgs A = lp_args(S);
    A.prop_next = prow(S, n + 1);
    k_lp_select<<<dim3(1, S->nrep), NT, lds_rows(S), S->ctx->stream>>>(A);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}

}  // namespace chm_smp

// ------------------------------------------------------------------ C ABI
extern "C" int chm_lp_info(chm_sampler *S, int *nparam, int *nocc, int *nfeat, int *kp) {
    CHM_REQUIRE(S && nparam && nocc && nfeat && kp, CHM_E_ARG, "bad arguments");
    *nparam = S->lp.nparam;
    *nocc = S->lp.nocc;
    *nfeat = S->lp.nfeat;
    *kp = S->lp.kp;
    return CHM_OK;
}

extern "C" int chm_lp_set_params(chm_sampler *S, const float *flat) {
    CHM_REQUIRE(S && flat, CHM_E_ARG, "bad arguments");
    chm_lp &L = S->lp;
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    CHM_HIP(hipStreamSynchronize(st));
    CHM_HIP(hipMemcpyAsync(L.params.p, flat, 4 * (size_t)L.nparam, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemsetAsync(L.adam_m.p, 0, L.adam_m.bytes, st));
    CHM_HIP(hipMemsetAsync(L.adam_v.p, 0, L.adam_v.bytes, st));
    L.adam_t = 0;
    int rc = quantise(S);
    if (rc) return rc;
    CHM_HIP(hipStreamSynchronize(st));
    L.have_params = true;
    return CHM_OK;
}

extern "C" int chm_lp_get_params(chm_sampler *S, float *flat) {
    CHM_REQUIRE(S && flat, CHM_E_ARG, "bad arguments");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    CHM_HIP(hipMemcpy(flat, S->lp.params.p, 4 * (size_t)S->lp.nparam, hipMemcpyDeviceToHost));
    return CHM_OK;
}

extern "C" int chm_lp_configure(chm_sampler *S, int mode, int ncand, double eps, double lam, double kappa,
                                double beta_max, int lift) {
    CHM_REQUIRE(S, CHM_E_ARG, "null sampler");
    CHM_REQUIRE(mode == 0 || mode == 1, CHM_E_ARG, "mode must be 0 (uniform) or 1 (learned)");
    CHM_REQUIRE(ncand >= MR && ncand <= MAXC && ncand % MR == 0, CHM_E_ARG,
                "ncand must be a multiple of 32 in 32..256");
    CHM_REQUIRE(eps >= 0.0 && eps <= 1.0, CHM_E_ARG, "eps must be in [0, 1]");
    CHM_REQUIRE(std::isfinite(lam) && std::isfinite(kappa), CHM_E_ARG, "lam and kappa must be finite");
    CHM_REQUIRE(beta_max > 0.0, CHM_E_ARG, "beta_max must be positive");
    CHM_REQUIRE(mode == 0 || S->lp.have_params, CHM_E_STATE, "set the network parameters first");
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    chm_lp &L = S->lp;
    L.mode = mode;
    L.ncand = ncand;
    L.eps = eps;
    L.lam = lam;
    L.kappa = kappa;
    L.beta_max = beta_max;
    L.lift = (mode && lift) ? 1 : 0;
    if (L.lift) {
        // directions start alternating over the slots (the stationary law of d is +-1
        // with equal probability, independent of the configuration)
        int rc = L.dir.ensure(sizeof(int) * S->nrep);
        if (rc) return rc;
        std::vector<int> d(S->nrep);
        for (int q = 0; q < S->nrep; ++q) d[q] = (q & 1) ? -1 : 1;
        CHM_HIP(hipMemcpy(L.dir.p, d.data(), sizeof(int) * S->nrep, hipMemcpyHostToDevice));
    }
    return CHM_OK;
}

extern "C" int chm_lp_set_records(chm_sampler *S, int capacity) {
    CHM_REQUIRE(S && capacity >= 0, CHM_E_ARG, "bad arguments");
    chm_lp &L = S->lp;
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    int rc;
    if ((rc = L.rec_occ.ensure((size_t)capacity * L.noccp)) || (rc = L.rec_y.ensure(4 * (size_t)capacity)) ||
        (rc = L.rec_meta.ensure(4 * (size_t)capacity)))
        return rc;
    if (capacity) {
        CHM_HIP(hipMemset(L.rec_meta.p, 0xFF, 4 * (size_t)capacity));
        CHM_HIP(hipMemset(L.rec_occ.p, 0, (size_t)capacity * L.noccp));
        CHM_HIP(hipMemset(L.rec_y.p, 0, 4 * (size_t)capacity));
    }
    L.rec_cap = capacity;
    L.rec_from = S->hstep;
    return CHM_OK;
}

extern "C" int chm_lp_read_records(chm_sampler *S, int64_t *nvalid, uint8_t *occ, float *y, int32_t *meta,
                                   int64_t rows) {
    CHM_REQUIRE(S && nvalid, CHM_E_ARG, "bad arguments");
    chm_lp &L = S->lp;
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    const long long written = (S->hstep - L.rec_from) * (long long)S->nrep;
    *nvalid = std::min<long long>(written, L.rec_cap);
    const int64_t n = std::min<int64_t>(rows, L.rec_cap);
    if (n > 0 && occ) {
        std::vector<uint8_t> tmp((size_t)n * L.noccp);
        CHM_HIP(hipMemcpy(tmp.data(), L.rec_occ.p, tmp.size(), hipMemcpyDeviceToHost));
        for (int64_t r = 0; r < n; ++r)
            std::copy(tmp.begin() + r * L.noccp, tmp.begin() + r * L.noccp + L.nocc, occ + r * L.nocc);
    }
    if (n > 0 && y) CHM_HIP(hipMemcpy(y, L.rec_y.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    if (n > 0 && meta) CHM_HIP(hipMemcpy(meta, L.rec_meta.p, 4 * (size_t)n, hipMemcpyDeviceToHost));
    return CHM_OK;
}

extern "C" int chm_lp_eval(chm_sampler *S, int n, const uint8_t *occ, float *g) {
    CHM_REQUIRE(S && occ && g && n >= 0, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(S->lp.have_params, CHM_E_STATE, "set the network parameters first");
    chm_lp &L = S->lp;
    for (long q = 0; q < (long)n * L.nocc; ++q)
        CHM_REQUIRE(occ[q] < S->g.nsp, CHM_E_ARG, "occupancy species out of range");
    if (n == 0) return CHM_OK;
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    chm::DevBuf din, dout;
    int rc;
    if ((rc = din.ensure((size_t)n * L.nocc)) || (rc = dout.ensure(4 * (size_t)n))) return rc;
    CHM_HIP(hipMemcpyAsync(din.p, occ, (size_t)n * L.nocc, hipMemcpyHostToDevice, st));
    LPArgs A = lp_args(S);
    k_lp_eval<<<(unsigned)((n + MR - 1) / MR), NT, lds_rows(S), st>>>(A, n, din.as<uint8_t>(), dout.as<float>());
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(g, dout.p, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_lp_grad(chm_sampler *S, int n, const uint8_t *occ, const float *y, const float *w,
                           float *grad, float *loss) {
    CHM_REQUIRE(S && occ && y && w && grad && loss && n > 0, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(S->lp.have_params, CHM_E_STATE, "set the network parameters first");
    chm_lp &L = S->lp;
    int wsum = 0;
    for (int r = 0; r < n; ++r) {
        CHM_REQUIRE(w[r] == 0.f || w[r] == 1.f, CHM_E_ARG, "weights must be 0 or 1");
        wsum += w[r] != 0.f;
    }
    for (long q = 0; q < (long)n * L.nocc; ++q)
        CHM_REQUIRE(occ[q] < S->g.nsp, CHM_E_ARG, "occupancy species out of range");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    int rc = ensure_train(S, n);
    if (rc) return rc;
    std::vector<uint8_t> o((size_t)n * L.noccp, 0);
    for (int r = 0; r < n; ++r) std::copy(occ + (size_t)r * L.nocc, occ + (size_t)(r + 1) * L.nocc, o.begin() + (size_t)r * L.noccp);
    CHM_HIP(hipMemcpyAsync(L.tr_occ.p, o.data(), o.size(), hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(L.tr_y.p, y, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(L.tr_w.p, w, 4 * (size_t)n, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemcpyAsync(L.tr_wsum.p, &wsum, 4, hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemsetAsync(L.loss.p, 0, 4, st));
    if ((rc = fwd_bwd(S, n))) return rc;
    CHM_HIP(hipMemcpyAsync(grad, L.grad.p, 4 * (size_t)L.nparam, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipMemcpyAsync(loss, L.loss.p, 4, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

extern "C" int chm_lp_profile(chm_sampler *S, int iters, double *ms) {
    CHM_REQUIRE(S && S->inited && iters > 0 && ms, CHM_E_ARG, "bad arguments");
    chm_lp &L = S->lp;
    CHM_REQUIRE(L.have_params, CHM_E_STATE, "set the network parameters first");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    chm::DevBuf pr;
    int rc;
    if ((rc = pr.ensure(sizeof(int) * (size_t)S->nrep * kPropInts))) return rc;
    // both kernels on the current state; the selection writes into scratch rows, so the
    // chain is not advanced (the scores and the pending occupancy are rewritten by the
    // next real step)
    LPArgs A = lp_args(S);
    A.prop_next = pr.as<int>();
    A.eps = 0.0;   // time the learned path on every slot
    A.beta_max = 1e300;
    hipEvent_t e0, e1, e2;
    CHM_HIP(hipEventCreate(&e0));
    CHM_HIP(hipEventCreate(&e1));
    CHM_HIP(hipEventCreate(&e2));
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i)
        k_lp_score<<<dim3(L.ncand / MR, S->nrep), NT, lds_rows(S), st>>>(A);
    CHM_HIP(hipEventRecord(e1, st));
    for (int i = 0; i < iters; ++i)
        k_lp_select<<<dim3(1, S->nrep), NT, lds_rows(S), st>>>(A);
    CHM_HIP(hipEventRecord(e2, st));
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipEventSynchronize(e2));
    float t1 = 0.f, t2 = 0.f;
    CHM_HIP(hipEventElapsedTime(&t1, e0, e1));
    CHM_HIP(hipEventElapsedTime(&t2, e1, e2));
    ms[0] = t1 / iters;
    ms[1] = t2 / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipEventDestroy(e2);
    return CHM_OK;
}

extern "C" int chm_lp_train(chm_sampler *S, int iters, int batch, float lr, int max_rung, float clip,
                            uint64_t seed, float *loss) {
    CHM_REQUIRE(S && iters >= 0 && batch > 0 && lr > 0.f && clip > 0.f, CHM_E_ARG, "bad arguments");
    chm_lp &L = S->lp;
    CHM_REQUIRE(L.have_params, CHM_E_STATE, "set the network parameters first");
    const long long written = (S->hstep - L.rec_from) * (long long)S->nrep;
    const long long nvalid = std::min<long long>(written, L.rec_cap);
    CHM_REQUIRE(nvalid > 0, CHM_E_STATE, "no move records to train on");
    CHM_HIP(hipSetDevice(S->ctx->device));
    hipStream_t st = S->ctx->stream;
    int rc = ensure_train(S, batch);
    if (rc) return rc;
    const Offs o = offsets(L.kp);
    for (int it = 0; it < iters; ++it) {
        CHM_HIP(hipMemsetAsync(L.tr_wsum.p, 0, 4, st));
        k_lp_batch<<<(unsigned)batch, 64, 0, st>>>(batch, nvalid, L.noccp, seed, L.adam_t, max_rung, clip,
                                                    L.rec_occ.as<uint8_t>(), L.rec_y.as<float>(),
                                                    L.rec_meta.as<int>(), L.tr_occ.as<uint8_t>(),
                                                    L.tr_y.as<float>(), L.tr_w.as<float>(), L.tr_wsum.as<int>());
        CHM_LAUNCH_CHECK();
        if (it == iters - 1) CHM_HIP(hipMemsetAsync(L.loss.p, 0, 4, st));
        if ((rc = fwd_bwd(S, batch))) return rc;
        L.adam_t += 1;
        const float c1 = 1.f - std::pow(0.9f, (float)L.adam_t);
        const float c2 = 1.f - std::pow(0.999f, (float)L.adam_t);
        k_lp_adam<<<(unsigned)((o.n + 255) / 256), 256, 0, st>>>(o.n, L.params.as<float>(), L.grad.as<float>(),
                                                                L.adam_m.as<float>(), L.adam_v.as<float>(), lr,
                                                                c1, c2);
        CHM_LAUNCH_CHECK();
        if ((rc = quantise(S))) return rc;
    }
    if (loss) CHM_HIP(hipMemcpyAsync(loss, L.loss.p, 4, hipMemcpyDeviceToHost, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}
