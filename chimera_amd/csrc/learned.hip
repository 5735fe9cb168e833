// Learned swap proposal for the device sampler (gfx950): the per-site layer-1 sums, the
// scoring kernel of a segment's first proposal, the network evaluation for tests, the
// trainer and the C ABI.  No counterpart in the reference: its swap proposal is uniform
// (exafs_gpu/kernels.py:362-398).  The network, the per-site sums and the speculative
// scoring that runs inside the transform launches are described in learned_dev.h.
//
// Training: the sampler writes a record (occupancy, dE, rung, accepted) of every swap it
// proposes; `chm_lp_train` runs Adam on minibatches of the ring with an FP32 forward and
// backward pass (bf16 W1 and W2 with the same H1 rounding as the scorer).
#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "../../include/chimera_hip.h"
#include "common.h"
#include "learned_dev.h"
#include "sampler_dev.h"
#include "sampler_state.h"

namespace {

using namespace chm_smp;
using namespace chm_lp_dev;

__device__ __forceinline__ float silu(float z) { return z / (1.f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.f / (1.f + expf(-z));
    return s * (1.f + z * (1.f - s));
}
__device__ __forceinline__ float bf16_round(float f) { return bf16_to_float(bf16_rne(f)); }
__device__ __forceinline__ float bf(__bf16 v) { return bf16_to_float(__builtin_bit_cast(unsigned short, v)); }

// ------------------------------------------------------------------ per-site sums
// grid (natoms, nrep), HID threads: A and B rows of every site of every slot from scratch
__global__ __launch_bounds__(HID) void k_lp_cache_build(LPArgs A, int *cache) {
    __shared__ int occ[256 + 1];
    const Geo &g = A.g;
    const int c = blockIdx.x, s = blockIdx.y, h = threadIdx.x;
    const uint8_t *sp = A.species + (long)s * g.natoms_pad;
    for (int q = h; q < g.ncoord; q += blockDim.x) occ[q] = sp[A.nbr[(long)c * g.ncoord + q]];
    if (h == 0) occ[g.ncoord] = sp[c];
    __syncthreads();
    for (int role = 0; role < 2; ++role) {
        int acc = A.w1q[(long)feat_own(g, role, occ[g.ncoord]) * HID + h];
        for (int q = 0; q < g.ncoord; ++q) acc += A.w1q[(long)feat_nbr(g, role, q, occ[q]) * HID + h];
        const long o = (((long)s * 2 + role) * g.natoms + c) * HID + h;
        cache[o] = acc;
    }
}

// ------------------------------------------------------------------ a segment's first scores
// scoring workgroups only (mode 0): step n+1's candidates at the configuration after n steps
__global__ __launch_bounds__(TB) void k_lp_tiles(LPArgs A) {
    extern __shared__ __align__(16) char lds_raw[];
    lp_block(A, blockIdx.x, lds_raw);
}

// ------------------------------------------------------------------ evaluation (tests)
// grid ceil(n / MR), TB threads: g of given occupancy rows (n, nocc) through the scorer's
// network (fixed-point layer 1 summed from the occupancy, layer 2 on MFMA)
__global__ __launch_bounds__(TB) void k_lp_eval(LPArgs A, int nocc, int n, const uint8_t *__restrict__ occ_in,
                                                float *__restrict__ out) {
    extern __shared__ __align__(16) char lds_raw[];
    __shared__ float gp[TW * MP], gout[MP];
    __bf16 *H = reinterpret_cast<__bf16 *>(lds_raw);
    const int r0 = blockIdx.x * MR, nr = min(MR, n - r0), t = threadIdx.x;
    const int r = t >> 3;
    const bool on = r < nr;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        const int ch = (t & 7) + 8 * hh;
        int acc[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) acc[i] = 0;
        if (on) {
            const uint8_t *o = occ_in + (long)(r0 + r) * nocc;
            for (int q = 0; q < nocc; ++q) {
                const int sq = o[q] < A.g.nsp ? o[q] : 0;   // (rows are validated on the host)
                add_row16(acc, A.w1q + (long)(q * A.g.nsp + sq) * HID + 16 * ch, 1);
            }
        }
        h1_store(A.W, H, r, ch, acc, on);
    }
    W2Half F;
    float b2[4], w3[4];
    load_w2_half(A.W, 0, F);
    load_b2w3(A.W, b2, w3);
    __syncthreads();
    layer2(A.W, H, 2, F, b2, w3, gp, gout);
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

// bf16 W1 (the trainer's), the scorer's fixed-point W1q (bf16(W1) in units of 2^-16, half to
// even, clamped to |w| <= 2), bf16 W2 in both orders and the rounded w3
__global__ void k_lp_quant(int kp, const float *__restrict__ p, int off_w2, int off_w3, __bf16 *__restrict__ w1,
                           int *__restrict__ w1q, __bf16 *__restrict__ w2, __bf16 *__restrict__ w2t,
                           float *__restrict__ w3q) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long n1 = (long)kp * HID, n2 = (long)HID * HID;
    if (i < n1) {
        const unsigned short h = bf16_rne(p[i]);
        w1[i] = __builtin_bit_cast(__bf16, h);
        const float q = rintf(bf16_to_float(h) * 65536.f);
        w1q[i] = (int)fminf(fmaxf(q, (float)-kW1QMax), (float)kW1QMax);
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

// the parts of the scoring arguments that do not depend on the step
LPArgs lp_args(chm_sampler *S) {
    const chm_lp &L = S->lp;
    LPArgs A{};
    A.g = S->g;
    A.species = S->species.as<uint8_t>();
    A.rung_of = S->rung_of.as<int>();
    A.beta_slot = S->beta_slot.as<double>();
    A.ss_sten = S->ss_sten.as<int>();
    A.nbr = S->nbr_st.as<int>();
    A.posmap = L.posmap.as<signed char>();
    A.pm_r = L.pm_r;
    for (int q = 0; q < kMaxFs; ++q) A.fs[q] = L.fs[q];
    A.n1 = L.n1;
    A.seed = S->seed;
    A.ncand = L.ncand;
    A.nrep = S->nrep;
    A.r2max = L.r2max;
    A.eps = L.eps;
    A.lam = L.lam;
    A.kappa = L.kappa;
    A.beta_max = L.beta_max;
    A.lift = L.lift;
    for (int q = 0; q < 64; ++q) A.wpair[q] = L.wpair[q];
    const Offs o = offsets(L.kp);
    A.W.w2t = L.w2t.as<__bf16>();
    A.W.b1 = L.params.as<float>() + o.b1;
    A.W.b2 = L.params.as<float>() + o.b2;
    A.W.b3 = L.params.as<float>() + o.b3;
    A.W.w3q = L.w3q.as<float>();
    A.w1q = L.w1q.as<int>();
    A.cs = L.cs.as<int>();
    A.lq = L.lq.as<double>();
    A.dmax = L.dmax.as<float>();
    return A;
}

int quantise(chm_sampler *S) {
    chm_lp &L = S->lp;
    const Offs o = offsets(L.kp);
    const long n = (long)L.kp * HID + HID * HID + HID;
    k_lp_quant<<<(unsigned)((n + 255) / 256), 256, 0, S->ctx->stream>>>(
        L.kp, L.params.as<float>(), o.w2, o.w3, L.w1.as<__bf16>(), L.w1q.as<int>(), L.w2.as<__bf16>(),
        L.w2t.as<__bf16>(), L.w3q.as<float>());
    CHM_LAUNCH_CHECK();
    L.cache_ok = false;   // the per-site sums depend on W1
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
    for (int q = 0; q < 64; ++q) L.wpair[q] = (q / 8 != q % 8) ? 1 : 0;   // unlike bonds
    if (L.kp > 1024 || g.ncoord > 127 || g.nsp > kMaxSpW)
        return chm::set_error("learned proposal: too many occupancy features"), CHM_E_ARG;
    hipStream_t st = S->ctx->stream;
    // stencil position of every offset in the cube |d| <= pm_r (units a0/2), -1 elsewhere
    std::vector<int> ss(g.ncoord);
    CHM_HIP(hipMemcpy(ss.data(), S->ss_sten.p, sizeof(int) * g.ncoord, hipMemcpyDeviceToHost));
    auto comp = [](int v, int k) { return (int)(signed char)((v >> (8 * k)) & 0xFF); };
    int R = 0;
    for (int q = 0; q < g.ncoord; ++q)
        for (int k = 0; k < 3; ++k) R = std::max(R, std::abs(comp(ss[q], k)));
    L.pm_r = R;
    const int side = 2 * R + 1;
    if ((size_t)side * side * side > (size_t)kPosMapMax)
        return chm::set_error("learned proposal: stencil offsets too long"), CHM_E_ARG;
    // (padded to 16 bytes: the scoring workgroups read it in 16-byte pieces)
    std::vector<signed char> pm(((size_t)side * side * side + 15) & ~(size_t)15, -1);
    auto at = [&](int dx, int dy, int dz) -> signed char & {
        return pm[((size_t)(dx + R) * side + (dy + R)) * side + (dz + R)];
    };
    int n1 = 0;
    for (int q = 0; q < g.ncoord; ++q) {
        at(comp(ss[q], 0), comp(ss[q], 1), comp(ss[q], 2)) = (signed char)q;
        n1 += ((ss[q] >> 24) & 0xFF) == 0;
    }
    if (n1 > kMaxFs) return chm::set_error("learned proposal: too many first-shell neighbours"), CHM_E_ARG;
    L.n1 = 0;
    for (int q = 0; q < g.ncoord; ++q)
        if (((ss[q] >> 24) & 0xFF) == 0) L.fs[L.n1++] = (signed char)q;
    // the fold of a swap visits the stencil of each swapped site for the sites whose stencil
    // holds it: the same set only for a stencil that holds every offset's opposite (as that
    // of a Bravais lattice does)
    for (int q = 0; q < g.ncoord; ++q)
        if (at(-comp(ss[q], 0), -comp(ss[q], 1), -comp(ss[q], 2)) < 0)
            return chm::set_error("learned proposal: the stencil is not symmetric"), CHM_E_ARG;
    const size_t nr = (size_t)S->nrep, nc = MAXC;
    int rc;
    if ((rc = L.params.ensure(4 * (size_t)L.nparam)) || (rc = L.adam_m.ensure(4 * (size_t)L.nparam)) ||
        (rc = L.adam_v.ensure(4 * (size_t)L.nparam)) || (rc = L.w1q.ensure(4 * (size_t)L.kp * HID)) ||
        (rc = L.w1.ensure(2 * (size_t)L.kp * HID)) || (rc = L.w2t.ensure(2 * (size_t)HID * HID)) ||
        (rc = L.w2.ensure(2 * (size_t)HID * HID)) || (rc = L.w3q.ensure(4 * HID)) ||
        (rc = L.posmap.ensure(pm.size())) ||
        (rc = L.sg.ensure(4 * 2 * 2 * nr * nc)) || (rc = L.sdu.ensure(4 * 2 * 2 * nr * nc)) ||
        (rc = L.st8.ensure(2 * nr * nc)) || (rc = L.cs.ensure(4 * nr)) ||
        (rc = L.lq.ensure(8 * nr * MAXT * kLqDoubles)) || (rc = L.dmax.ensure(4 * nr * MAXT * 2)) ||
        (rc = L.dir.ensure(4 * 2 * nr)) || (rc = L.lpn.ensure(8 * nr)) || (rc = L.pend_occ.ensure(nr * L.noccp)))
        return rc;
    CHM_HIP(hipMemcpyAsync(L.posmap.p, pm.data(), pm.size(), hipMemcpyHostToDevice, st));
    CHM_HIP(hipMemsetAsync(L.pend_occ.p, 0, L.pend_occ.bytes, st));
    CHM_HIP(hipMemsetAsync(L.cs.p, 0, L.cs.bytes, st));
    CHM_HIP(hipMemsetAsync(L.st8.p, 0, L.st8.bytes, st));
    CHM_HIP(hipStreamSynchronize(st));
    return CHM_OK;
}

// The per-site layer-1 sums of every slot from its species (after a change of W1, of the
// species from the host, or of the proposal mode)
int lp_ensure_cache(chm_sampler *S) {
    chm_lp &L = S->lp;
    if (L.cache_ok) return CHM_OK;
    int rc = L.cache.ensure(4 * (size_t)S->nrep * 2 * S->g.natoms * HID);
    if (rc) return rc;
    LPArgs A = lp_args(S);
    k_lp_cache_build<<<dim3(S->g.natoms, S->nrep), HID, 0, S->ctx->stream>>>(A, L.cache.as<int>());
    CHM_LAUNCH_CHECK();
    L.cache_ok = true;
    return CHM_OK;
}

size_t lp_tile_lds(const chm_sampler *S) { return tile_lds(S->g.natoms_pad); }

// The scoring arguments of the transform launch of step n (mode 1; 2 for the last step of a
// segment) or of a segment's first scoring launch after n steps (mode 0)
LPArgs lp_step_args(chm_sampler *S, int mode, long long n) {
    chm_lp &L = S->lp;
    LPArgs A = lp_args(S);
    A.mode = mode;
    A.stepc = S->stepc.as<uint64_t>();
    A.lpn = L.lpn.as<uint64_t>();
    A.cache = L.cache.as<int>();
    const size_t np = (size_t)S->nrep * 2 * MAXC, nt = (size_t)S->nrep * MAXC;
    const long long nx = n + 1;
    int *dir = L.dir.as<int>();
    if (mode == 0) {
        A.dir = dir + (size_t)(nx & 1) * S->nrep;
    } else {
        A.prop_n = prow(S, n);
        A.dir = dir + (size_t)(n & 1) * S->nrep;
        A.sg_in = L.sg.as<float>() + (size_t)(n & 1) * np;
        A.sdu_in = L.sdu.as<int>() + (size_t)(n & 1) * np;
        A.st_in = L.st8.as<int8_t>() + (size_t)(n & 1) * nt;
    }
    A.sg_out = L.sg.as<float>() + (size_t)(nx & 1) * np;
    A.sdu_out = L.sdu.as<int>() + (size_t)(nx & 1) * np;
    A.st_out = L.st8.as<int8_t>() + (size_t)(nx & 1) * nt;
    const int slots8 = (S->nrep + 7) / 8 * 8;
    A.ntile_blocks = (L.ncand / MR) * slots8;
    return A;
}

// a segment's first scoring launch, after h steps
int lp_launch_tiles0(chm_sampler *S, long long h) {
    LPArgs A = lp_step_args(S, 0, h);
    k_lp_tiles<<<(unsigned)A.ntile_blocks, TB, lp_tile_lds(S), S->ctx->stream>>>(A);
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
    L.cache_ok = false;   // the sums were not kept while the proposal was off
    if (L.lift) {
        // directions start alternating over the slots (the stationary law of d is +-1
        // with equal probability, independent of the configuration); two copies, the
        // direction before and after a step, alternate by step
        std::vector<int> d(2 * (size_t)S->nrep);
        for (int q = 0; q < 2 * S->nrep; ++q) d[q] = ((q % S->nrep) & 1) ? -1 : 1;
        CHM_HIP(hipMemcpy(L.dir.p, d.data(), sizeof(int) * d.size(), hipMemcpyHostToDevice));
    }
    return CHM_OK;
}

extern "C" int chm_lp_set_tilt_weights(chm_sampler *S, const int32_t *w, int nsp) {
    CHM_REQUIRE(S && w, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(nsp == S->g.nsp, CHM_E_ARG, "weights must be nsp x nsp");
    for (int i = 0; i < nsp; ++i)
        for (int j = 0; j < nsp; ++j) {
            CHM_REQUIRE(w[i * nsp + j] == w[j * nsp + i], CHM_E_ARG, "weights must be symmetric");
            CHM_REQUIRE(w[i * nsp + j] >= -63 && w[i * nsp + j] <= 63, CHM_E_ARG, "weights must be in [-63, 63]");
        }
    CHM_HIP(hipSetDevice(S->ctx->device));
    CHM_HIP(hipStreamSynchronize(S->ctx->stream));
    drop_graph(S);
    for (int q = 0; q < 64; ++q) S->lp.wpair[q] = 0;
    for (int i = 0; i < nsp; ++i)
        for (int j = 0; j < nsp; ++j) S->lp.wpair[8 * i + j] = (signed char)w[i * nsp + j];
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
    k_lp_eval<<<(unsigned)((n + MR - 1) / MR), TB, (size_t)MR * LDH * 2, st>>>(A, L.nocc, n, din.as<uint8_t>(),
                                                                                dout.as<float>());
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
    int rc = lp_ensure_cache(S);
    if (rc) return rc;
    // the scoring workgroups of a step on the current state (every slot learned, the last
    // step's proposal as the swap; they rewrite only buffers the next segment's first
    // launches rewrite again).  ms[1] = 0: an accepted swap is folded into the per-site
    // sums inside the per-replica kernel, with no launch of its own.
    const double eps = L.eps, bmax = L.beta_max;
    L.eps = 0.0;
    L.beta_max = 1e300;
    const LPArgs A = lp_step_args(S, 1, S->hstep);
    L.eps = eps;
    L.beta_max = bmax;
    hipEvent_t e0, e1;
    CHM_HIP(hipEventCreate(&e0));
    CHM_HIP(hipEventCreate(&e1));
    CHM_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters; ++i) k_lp_tiles<<<(unsigned)A.ntile_blocks, TB, lp_tile_lds(S), st>>>(A);
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipEventRecord(e1, st));
    CHM_HIP(hipEventSynchronize(e1));
    float t1 = 0.f;
    CHM_HIP(hipEventElapsedTime(&t1, e0, e1));
    ms[0] = t1 / iters;
    ms[1] = 0.0;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
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
