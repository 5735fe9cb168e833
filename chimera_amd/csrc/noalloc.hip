// chm_fm_chi_noalloc: the pair-space path sum of chm_fm_chi on device-resident inputs and
// output, allocating nothing.  Its intermediate arrays (pair distances, pair types, the
// per-workgroup partial sums and a per-configuration input flag) are carved out of one
// static device array, `scratch`, instead of buffers the library allocates on demand, so a
// caller that must not allocate device memory (a stream being captured into a graph, a
// serving loop) can call it.  The kernels are the ones chm_fm_chi launches.
//
// Host code never passes `scratch` itself (on the host the name is only a stand-in): the
// device address of each device's copy is looked up once per device by scratch_init,
// called from chm_ctx_create with that device current, and kept in g_scratch.
#include <mutex>
#include <string>

#include "common.h"

namespace {

constexpr size_t kScratchBytes = (size_t)32 << 20;   // per device, held until the process ends
constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
unsigned char *g_scratch[kMaxDevices] = {};   // device address of `scratch` on each device

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Pair types with the species range checked: an out-of-range species gives type 0 (no
// table is read out of bounds) and marks its configuration in bad[].
__global__ void k_ptypes_checked(const int *__restrict__ species, const int *__restrict__ iidx,
                                 const int *__restrict__ jidx, int *__restrict__ pt,
                                 int *__restrict__ bad, int npairs, int natoms, int nconf, int nsp) {
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (long)npairs * nconf) return;
    const int ic = (int)(tid / npairs), p = (int)(tid % npairs);
    const int a = species[(long)ic * natoms + iidx[p]], b = species[(long)ic * natoms + jidx[p]];
    const bool ok = (unsigned)a < (unsigned)nsp && (unsigned)b < (unsigned)nsp;
    pt[tid] = ok ? a * nsp + b : 0;
    if (!ok) atomicOr(&bad[ic], 1);
}

// a fill kernel rather than hipMemsetAsync, which must not be captured into a graph here
__global__ void k_zero_i32(int *__restrict__ p, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0;
}

// the rows of a configuration marked in bad[] become NaN (its species input was invalid)
__global__ void k_poison_rows(double *__restrict__ chi, const int *__restrict__ bad, int n_per_conf,
                              int nconf) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_per_conf * nconf) return;
    if (bad[t / n_per_conf]) chi[t] = __builtin_nan("");
}

}  // namespace

namespace chm {
// Once per device, with that device current: the device address of `scratch`.
int scratch_init(int device) {
    CHM_REQUIRE(device >= 0 && device < kMaxDevices, CHM_E_ARG, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return CHM_OK;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        return CHM_E_HIP;
    }
    g_scratch[device] = static_cast<unsigned char *>(p);
    return CHM_OK;
}
}  // namespace chm

extern "C" int chm_fm_chi_noalloc(chm_ctx *ctx, int nconf, const double *d_pos,
                                  const int32_t *d_species, double *d_chi, int precision,
                                  void *stream) {
    CHM_REQUIRE(ctx && ctx->npairs > 0, CHM_E_STATE, "set tables and topology first");
    CHM_REQUIRE(nconf > 0 && d_pos && d_species && d_chi, CHM_E_ARG, "bad arguments");
    CHM_REQUIRE(precision == CHM_PREC_F64 || precision == CHM_PREC_F32, CHM_E_ARG,
                "precision must be CHM_PREC_F64 or CHM_PREC_F32");
    CHM_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices, CHM_E_ARG, "device index too large");
    unsigned char *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch[ctx->device];
    }
    CHM_REQUIRE(base, CHM_E_STATE, "scratch address not set up for this device");
    chm::ChiShape sh;
    int rc = chm::chi_shape(ctx, nconf, sh);
    if (rc) return rc;
    const size_t tot = (size_t)nconf * ctx->npairs;
    const size_t n_per_conf = (size_t)ctx->nsp * ctx->nk;
    const size_t b_r = align256(sizeof(double) * tot), b_pt = align256(sizeof(int32_t) * tot);
    const size_t b_part = align256(sizeof(double) * (size_t)nconf * sh.nwg * n_per_conf);
    const size_t b_bad = align256(sizeof(int) * (size_t)nconf);
    const size_t need = b_r + b_pt + b_part + b_bad;
    if (need > kScratchBytes) {
        chm::set_error("chm_fm_chi_noalloc: " + std::to_string(nconf) + " configurations need " +
                       std::to_string(need) + " bytes of scratch, the maximum is " +
                       std::to_string(kScratchBytes) + ": pass fewer configurations per call");
        return CHM_E_ARG;
    }
    double *r = reinterpret_cast<double *>(base);
    int *pt = reinterpret_cast<int *>(base + b_r);
    double *partial = reinterpret_cast<double *>(base + b_r + b_pt);
    int *bad = reinterpret_cast<int *>(base + b_r + b_pt + b_part);
    CHM_HIP(hipSetDevice(ctx->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    k_zero_i32<<<dim3((unsigned)((nconf + 255) / 256)), 256, 0, st>>>(bad, nconf);
    CHM_LAUNCH_CHECK();
    if ((rc = chm::distances_into(ctx, st, nconf, d_pos, r))) return rc;
    k_ptypes_checked<<<dim3((unsigned)((tot + 255) / 256)), 256, 0, st>>>(
        d_species, ctx->i_idx.as<int>(), ctx->j_idx.as<int>(), pt, bad, ctx->npairs, ctx->natoms,
        nconf, ctx->nsp);
    CHM_LAUNCH_CHECK();
    if ((rc = chm::chi_into(ctx, st, nconf, precision, sh, r, pt, partial, d_chi))) return rc;
    const long nout = (long)nconf * n_per_conf;
    k_poison_rows<<<dim3((unsigned)((nout + 255) / 256)), 256, 0, st>>>(d_chi, bad, (int)n_per_conf,
                                                                       nconf);
    CHM_LAUNCH_CHECK();
    return CHM_OK;
}
