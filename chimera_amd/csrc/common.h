// Shared host/device declarations of libchimera_hip (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/chimera_hip.h"

namespace chm {

void set_error(const std::string &msg);

#define CHM_HIP(expr)                                                                \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            chm::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));        \
            return CHM_E_HIP;                                                        \
        }                                                                            \
    } while (0)

#define CHM_REQUIRE(cond, code, msg)                                                 \
    do {                                                                             \
        if (!(cond)) {                                                               \
            chm::set_error(msg);                                                     \
            return code;                                                             \
        }                                                                            \
    } while (0)

#define CHM_LAUNCH_CHECK()                                                           \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess) {                                                      \
            chm::set_error(std::string("kernel launch: ") + hipGetErrorString(e_));   \
            return CHM_E_HIP;                                                        \
        }                                                                            \
    } while (0)

constexpr int kMaxNsp = 6;

// A device buffer owned by the library.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    ~DevBuf() { release(); }
    void release() {
        if (p) hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t n) {
        if (n <= bytes && p) return CHM_OK;
        release();
        if (n == 0) return CHM_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
            return CHM_E_HIP;
        }
        bytes = n;
        return CHM_OK;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

}  // namespace chm

struct chm_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int refs = 1;   // the owner's reference plus one per sampler / transform made on it
    // tables
    int nsp = 0, nk = 0;
    double s02 = 1.0;
    chm::DevBuf k, f, delta, lam;
    chm::DevBuf f32, drev32;   // FP32 f and delta / 2 pi for the FP32 path sum
    std::vector<double> h_k;
    // topology
    int natoms = 0, npairs = 0, nshell = 1;
    chm::DevBuf i_idx, j_idx, shift, shell, sig2;
    chm::DevBuf drtab, c3tab, s2tab;
    bool c3_zero = true;   // every shell's C3 is 0 (the adjoint may use its k recurrence)
    // pairs of each atom as absorber / as scatterer (CSR, pair order), for the adjoints
    chm::DevBuf csr_ioff, csr_i, csr_joff, csr_j;
    // scratch for the forward path
    chm::DevBuf pos, species, r, evec, ptype, partial, chi, rows_in, rows_out;
    // scratch for the adjoints and the relaxed-occupancy path sum
    chm::DevBuf w, gchi, gr, gwi, gwj, gpos, gw, soft_fc, soft_fs;
    // nconf of the last call of each profiled kind (chm_fm_profile): path sum FP64,
    // path sum FP32, hard adjoint, soft path sum, soft adjoint; 0 = none
    int prof_nconf[5] = {0, 0, 0, 0, 0};
};

struct chm_ft {
    chm_ctx *ctx = nullptr;
    int nchan = 0, nk = 0, nR = 0, nsp = 0;
    std::vector<int> chan_edge;
    std::vector<int> edge_first, edge_nkw;  // channels of each edge (contiguous)
    chm::DevBuf D;       // (nsp, nk, nkw_e * nR * 2): per edge, channels edge-major
    chm::DevBuf in, out, extra;
    int nfft = 0;
    chm::DevBuf ridx, rfrac, rins;   // (nchan, nfft) resample index, fraction, inside flag
};

namespace chm {
// A context lives until its owner and every object made on it have been destroyed, in
// any order (chm_ctx_destroy drops the owner's reference).
inline void ctx_retain(chm_ctx *c) { ++c->refs; }
inline void ctx_release(chm_ctx *c) {
    if (--c->refs > 0) return;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    hipStreamDestroy(c->stream);
    delete c;
}

// Launch shape of the pair-space path sum (forward.hip): first k index above 0, lanes per
// workgroup, workgroups per configuration, pairs per workgroup, k chunks and their width.
struct ChiShape {
    int k0, nthr, nwg, ppw;
    int nkc, kspan;   // chunks of the k > 0 points (grid z) and columns per chunk
};
int chi_shape(const chm_ctx *ctx, int nconf, ChiShape &sh);
// Path sum of nconf configurations from device distances r and pair types pt (nconf,
// npairs) into chi (nconf, nsp, nk), partial (nconf, nwg, nsp, nk) holding the per-
// workgroup sums; on stream st, nothing allocated.
int chi_into(chm_ctx *ctx, hipStream_t st, int nconf, int precision, const ChiShape &sh,
             const double *r, const int *pt, double *partial, double *chi);
// noalloc.hip: the device address of the allocation-free path sum's static scratch array,
// looked up once per device (called by chm_ctx_create with the device current).
int scratch_init(int device);
// Pair distances of nconf configurations (device positions) into r, on stream st.
int distances_into(chm_ctx *ctx, hipStream_t st, int nconf, const double *pos, double *r);

// Generic FP64 GEMM on the context stream: C[b] (M x N) = A[b] (M x K) @ B[b] (K x N),
// row-major with leading dimensions and per-batch element strides.
int gemm_f64(hipStream_t st, int nbatch, int M, int N, int K, const double *A, int lda,
             long sA, const double *B, int ldb, long sB, double *C, int ldc, long sC);
}  // namespace chm
