// Windowed k -> R transform as a DFT-matrix contraction (gfx950).
//
// The reference resamples chi(k) onto a uniform grid, multiplies by W(k) k^w and
// takes an FFT per channel (exafs_gpu/fourier.py:69-80).  Every step of that is linear
// in chi, so it is folded once into a dense matrix D_e (nk_model x nkw*nR*2) per edge,
// and the transform of any batch of rows (replicas, basis rows) is one GEMM against
// it.  The R grid is arbitrary: the FFT grid j*pi/(nfft dk) or the fit's own grid.
#include <cmath>

#include "common.h"

namespace {

// D[e][n][(kwl*nR + j)*2 + {0,1}] = sum_u coef_u(n) * weight[c][u] * norm * e^{-2i k_u R_j}
// One thread per (channel, j); it owns column j of its channel, so no two threads
// write the same element.
__global__ void k_build_dft(const int *__restrict__ idx, const double *__restrict__ frac,
                            const double *__restrict__ inside, const double *__restrict__ weight,
                            const double *__restrict__ kuni, const double *__restrict__ R,
                            const int *__restrict__ chan_edge, const int *__restrict__ chan_kwl,
                            double *__restrict__ D, int nchan, int nk, int nfft, int nR,
                            int ncol_e, double norm) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)nchan * nR) return;
    int c = (int)(tid / nR), j = (int)(tid % nR);
    int e = chan_edge[c], kwl = chan_kwl[c];
    double *De = D + (long)e * nk * ncol_e;
    int col = (kwl * nR + j) * 2;
    for (int n = 0; n < nk; ++n) {
        De[(long)n * ncol_e + col] = 0.0;
        De[(long)n * ncol_e + col + 1] = 0.0;
    }
    const double Rj = R[j];
    for (int u = 0; u < nfft; ++u) {
        double w = weight[(long)c * nfft + u];
        double ins = inside[(long)c * nfft + u];
        if (w == 0.0 || ins == 0.0) continue;
        double s, co;
        sincos(2.0 * kuni[u] * Rj, &s, &co);
        double zr = w * norm * co, zi = -w * norm * s;
        int n0 = idx[(long)c * nfft + u];
        double fr = frac[(long)c * nfft + u];
        double a0 = (1.0 - fr) * ins, a1 = fr * ins;
        De[(long)n0 * ncol_e + col] += a0 * zr;
        De[(long)n0 * ncol_e + col + 1] += a0 * zi;
        De[(long)(n0 + 1) * ncol_e + col] += a1 * zr;
        De[(long)(n0 + 1) * ncol_e + col + 1] += a1 * zi;
    }
}

// C = A @ B, FP64, 256 threads, 32 x 64 output tile, K staged by 16.
constexpr int BM = 32, BN = 64, BK = 16;

__global__ __launch_bounds__(256) void k_gemm_f64(int M, int N, int K,
                                                   const double *__restrict__ A, int lda, long sA,
                                                   const double *__restrict__ B, int ldb, long sB,
                                                   double *__restrict__ C, int ldc, long sC) {
    __shared__ double As[BK][BM + 1];
    __shared__ double Bs[BK][BN];
    const int b = blockIdx.z;
    A += (long)b * sA;
    B += (long)b * sB;
    C += (long)b * sC;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int tm = threadIdx.x / 16, tn = threadIdx.x % 16;
    double acc[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    for (int k0 = 0; k0 < K; k0 += BK) {
        // A tile: 32 x 16 = 512 elements, 2 per thread
        for (int q = threadIdx.x; q < BM * BK; q += 256) {
            int mm = q / BK, kk = q % BK;
            int gm = m0 + mm, gk = k0 + kk;
            As[kk][mm] = (gm < M && gk < K) ? A[(long)gm * lda + gk] : 0.0;
        }
        // B tile: 16 x 64 = 1024 elements, 4 per thread, coalesced along n
        for (int q = threadIdx.x; q < BK * BN; q += 256) {
            int kk = q / BN, nn = q % BN;
            int gk = k0 + kk, gn = n0 + nn;
            Bs[kk][nn] = (gk < K && gn < N) ? B[(long)gk * ldb + gn] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; ++kk) {
            double a0 = As[kk][tm], a1 = As[kk][tm + 16];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double bv = Bs[kk][tn + 16 * v];
                acc[0][v] = fma(a0, bv, acc[0][v]);
                acc[1][v] = fma(a1, bv, acc[1][v]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        int gm = m0 + tm + 16 * h;
        if (gm >= M) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int gn = n0 + tn + 16 * v;
            if (gn < N) C[(long)gm * ldc + gn] = acc[h][v];
        }
    }
}

// out[r][u] = (a + (b - a) frac) inside with a, b = chi[r][idx], chi[r][idx + 1]: the
// reference's linear resample onto the uniform grid (exafs_gpu/fourier.py:69-73), each
// operation rounded on its own (no fused multiply-add) so the result is NumPy's bit for bit
__global__ void k_resample(const double *__restrict__ chi, const int *__restrict__ idx,
                           const double *__restrict__ frac, const double *__restrict__ inside,
                           double *__restrict__ out, int nrows, int nk, int nfft) {
#pragma clang fp contract(off)
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)nrows * nfft) return;
    const int r = (int)(tid / nfft), u = (int)(tid % nfft);
    const int i = idx[u];
    const double a = chi[(long)r * nk + i], b = chi[(long)r * nk + i + 1];
    out[tid] = (a + (b - a) * frac[u]) * inside[u];
}

__global__ void k_add_extra(double *__restrict__ x, const double *__restrict__ extra, long n,
                            int per) {
    long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid < n) x[tid] += extra[tid % per];
}

}  // namespace

namespace chm {
int gemm_f64(hipStream_t st, int nbatch, int M, int N, int K, const double *A, int lda, long sA,
             const double *B, int ldb, long sB, double *C, int ldc, long sC) {
    if (M <= 0 || N <= 0 || nbatch <= 0) return CHM_OK;
    dim3 grid((N + BN - 1) / BN, (M + BM - 1) / BM, nbatch);
    k_gemm_f64<<<grid, 256, 0, st>>>(M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}
}  // namespace chm

extern "C" int chm_ft_create(chm_ctx *ctx, int nchan, int nk_model, int nfft, int nR,
                             const int32_t *chan_edge, const int32_t *idx, const double *frac,
                             const double *inside, const double *weight, const double *kuni,
                             const double *R, double norm, chm_ft **out) {
    CHM_REQUIRE(ctx && out, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(nchan > 0 && nk_model >= 2 && nfft > 0 && nR > 0, CHM_E_ARG, "bad sizes");
    CHM_REQUIRE(chan_edge && idx && frac && inside && weight && kuni && R, CHM_E_ARG,
                "null input");
    // channels edge-major, the same number per edge (method_a_rmc.py:127-135)
    int nsp = 0;
    for (int c = 0; c < nchan; ++c) nsp = std::max(nsp, chan_edge[c] + 1);
    CHM_REQUIRE(nchan % nsp == 0, CHM_E_ASSERT, "channels do not divide into edges");
    int nkw = nchan / nsp;
    std::vector<int> kwl(nchan);
    for (int c = 0; c < nchan; ++c) {
        CHM_REQUIRE(chan_edge[c] == c / nkw, CHM_E_ASSERT, "channels must be ordered edge-major");
        kwl[c] = c % nkw;
    }
    for (long q = 0; q < (long)nchan * nfft; ++q)
        CHM_REQUIRE(idx[q] >= 0 && idx[q] + 1 < nk_model, CHM_E_ARG, "resample index out of range");
    CHM_HIP(hipSetDevice(ctx->device));
    auto *ft = new chm_ft();
    ft->ctx = ctx;
    ft->nchan = nchan;
    ft->nk = nk_model;
    ft->nR = nR;
    ft->nsp = nsp;
    ft->chan_edge.assign(chan_edge, chan_edge + nchan);
    for (int e = 0; e < nsp; ++e) {
        ft->edge_first.push_back(e * nkw);
        ft->edge_nkw.push_back(nkw);
    }
    const int ncol_e = nkw * nR * 2;
    chm::DevBuf dw, dk, dR, dce, dkwl;
    chm::DevBuf &didx = ft->ridx, &dfrac = ft->rfrac, &dins = ft->rins;   // kept for resample
    ft->nfft = nfft;
    auto up = [&](chm::DevBuf &b, const void *h, size_t n) -> int {
        int rc = b.ensure(n);
        if (rc) return rc;
        CHM_HIP(hipMemcpyAsync(b.p, h, n, hipMemcpyHostToDevice, ctx->stream));
        return CHM_OK;
    };
    int rc = 0;
    size_t cf = (size_t)nchan * nfft;
    if ((rc = up(didx, idx, sizeof(int32_t) * cf)) || (rc = up(dfrac, frac, sizeof(double) * cf)) ||
        (rc = up(dins, inside, sizeof(double) * cf)) || (rc = up(dw, weight, sizeof(double) * cf)) ||
        (rc = up(dk, kuni, sizeof(double) * nfft)) || (rc = up(dR, R, sizeof(double) * nR)) ||
        (rc = up(dce, chan_edge, sizeof(int32_t) * nchan)) ||
        (rc = up(dkwl, kwl.data(), sizeof(int32_t) * nchan)) ||
        (rc = ft->D.ensure(sizeof(double) * (size_t)nsp * nk_model * ncol_e))) {
        delete ft;
        return rc;
    }
    long nt = (long)nchan * nR;
    k_build_dft<<<dim3((unsigned)((nt + 127) / 128)), 128, 0, ctx->stream>>>(
        didx.as<int>(), dfrac.as<double>(), dins.as<double>(), dw.as<double>(), dk.as<double>(),
        dR.as<double>(), dce.as<int>(), dkwl.as<int>(), ft->D.as<double>(), nchan, nk_model, nfft,
        nR, ncol_e, norm);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        delete ft;
        chm::set_error(std::string("build_dft: ") + hipGetErrorString(e));
        return CHM_E_HIP;
    }
    chm::ctx_retain(ctx);
    *out = ft;
    return CHM_OK;
}

extern "C" int chm_ft_destroy(chm_ft *ft) {
    if (ft) {
        chm_ctx *ctx = ft->ctx;
        hipSetDevice(ctx->device);
        hipStreamSynchronize(ctx->stream);
        delete ft;
        chm::ctx_release(ctx);
    }
    return CHM_OK;
}

extern "C" int chm_ft_apply(chm_ft *ft, int nrows, const double *chi, const double *chi_extra,
                            double *out) {
    CHM_REQUIRE(ft && nrows > 0 && chi && out, CHM_E_ARG, "bad arguments");
    chm_ctx *ctx = ft->ctx;
    CHM_HIP(hipSetDevice(ctx->device));
    const int nk = ft->nk, nsp = ft->nsp, nR = ft->nR;
    const int nkw = ft->nchan / nsp, ncol_e = nkw * nR * 2;
    size_t nin = (size_t)nrows * nsp * nk, nout = (size_t)nrows * ft->nchan * nR * 2;
    int rc;
    if ((rc = ft->in.ensure(sizeof(double) * nin)) || (rc = ft->out.ensure(sizeof(double) * nout)))
        return rc;
    CHM_HIP(hipMemcpyAsync(ft->in.p, chi, sizeof(double) * nin, hipMemcpyHostToDevice, ctx->stream));
    if (chi_extra) {
        if ((rc = ft->extra.ensure(sizeof(double) * nsp * nk))) return rc;
        CHM_HIP(hipMemcpyAsync(ft->extra.p, chi_extra, sizeof(double) * nsp * nk,
                               hipMemcpyHostToDevice, ctx->stream));
        k_add_extra<<<dim3((unsigned)((nin + 255) / 256)), 256, 0, ctx->stream>>>(
            ft->in.as<double>(), ft->extra.as<double>(), (long)nin, nsp * nk);
        CHM_LAUNCH_CHECK();
    }
    // batch over edges: A_e = chi[:, e, :], B_e = D_e, C_e = out[:, e*nkw .. ]
    rc = chm::gemm_f64(ctx->stream, nsp, nrows, ncol_e, nk, ft->in.as<double>(), nsp * nk, nk,
                       ft->D.as<double>(), ncol_e, (long)nk * ncol_e, ft->out.as<double>(),
                       ft->nchan * nR * 2, ncol_e);
    if (rc) return rc;
    CHM_HIP(hipMemcpyAsync(out, ft->out.p, sizeof(double) * nout, hipMemcpyDeviceToHost,
                           ctx->stream));
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

extern "C" int chm_ft_apply_edge(chm_ft *ft, int edge, int nrows, const double *rows,
                                 double *out) {
    CHM_REQUIRE(ft && nrows > 0 && rows && out, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(edge >= 0 && edge < ft->nsp, CHM_E_ARG, "no such edge");
    chm_ctx *ctx = ft->ctx;
    CHM_HIP(hipSetDevice(ctx->device));
    const int nk = ft->nk, nR = ft->nR;
    const int nkw = ft->nchan / ft->nsp, ncol_e = nkw * nR * 2;
    size_t nin = (size_t)nrows * nk, nout = (size_t)nrows * ncol_e;
    int rc;
    if ((rc = ft->in.ensure(sizeof(double) * nin)) || (rc = ft->out.ensure(sizeof(double) * nout)))
        return rc;
    CHM_HIP(hipMemcpyAsync(ft->in.p, rows, sizeof(double) * nin, hipMemcpyHostToDevice,
                           ctx->stream));
    rc = chm::gemm_f64(ctx->stream, 1, nrows, ncol_e, nk, ft->in.as<double>(), nk, 0,
                       ft->D.as<double>() + (size_t)edge * nk * ncol_e, ncol_e, 0,
                       ft->out.as<double>(), ncol_e, 0);
    if (rc) return rc;
    CHM_HIP(hipMemcpyAsync(out, ft->out.p, sizeof(double) * nout, hipMemcpyDeviceToHost,
                           ctx->stream));
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}

extern "C" int chm_ft_resample(chm_ft *ft, int channel, int nrows, const double *chi, double *out) {
    CHM_REQUIRE(ft && nrows > 0 && chi && out, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(channel >= 0 && channel < ft->nchan, CHM_E_ARG, "no such channel");
    chm_ctx *ctx = ft->ctx;
    CHM_HIP(hipSetDevice(ctx->device));
    const size_t nin = (size_t)nrows * ft->nk, nout = (size_t)nrows * ft->nfft;
    int rc;
    if ((rc = ft->in.ensure(sizeof(double) * nin)) || (rc = ft->out.ensure(sizeof(double) * nout)))
        return rc;
    CHM_HIP(hipMemcpyAsync(ft->in.p, chi, sizeof(double) * nin, hipMemcpyHostToDevice, ctx->stream));
    const size_t off = (size_t)channel * ft->nfft;
    k_resample<<<dim3((unsigned)((nout + 255) / 256)), 256, 0, ctx->stream>>>(
        ft->in.as<double>(), ft->ridx.as<int>() + off, ft->rfrac.as<double>() + off,
        ft->rins.as<double>() + off, ft->out.as<double>(), nrows, ft->nk, ft->nfft);
    CHM_LAUNCH_CHECK();
    CHM_HIP(hipMemcpyAsync(out, ft->out.p, sizeof(double) * nout, hipMemcpyDeviceToHost, ctx->stream));
    CHM_HIP(hipStreamSynchronize(ctx->stream));
    return CHM_OK;
}
