// plb_noalloc.hip -- the macro read-back for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops, harnesses that
// forbid hipMalloc).  plb_read_macros stages its V-order gather through a
// device buffer it allocates and frees on every call; this variant stages
// through `scratch`, a fixed array every GPU the library runs on holds until
// the process ends.
//
// `scratch` is a device symbol: its name in host code is not a device
// address.  Each device's own copy is looked up once with
// hipGetSymbolAddress, with that device current, when the first context on
// the device is created; host code only ever passes the looked-up pointer.
// Every call on one GPU shares that GPU's copy: overlapping calls on one GPU
// race, and a call leaves the array holding the macros it read.
#include <stdio.h>

#include <mutex>

#include "plb_internal.hpp"

namespace plb {

constexpr int64_t NOALLOC_MAX = PLB_NOALLOC_MAX_NODES;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[4 * NOALLOC_MAX];

static double* scratch_ptr[MAX_DEVICES];
static std::once_flag scratch_once[MAX_DEVICES];
static hipError_t scratch_err[MAX_DEVICES];

// Called by context creation with `device` current.
int noalloc_setup(int device)
{
    if (device < 0 || device >= MAX_DEVICES) return PLB_ERR_INVALID;
    std::call_once(scratch_once[device], [device] {
        void* p = nullptr;
        scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        scratch_ptr[device] = scratch_err[device] == hipSuccess ? (double*)p : nullptr;
    });
    if (scratch_err[device] != hipSuccess) {
        plb_set_hip_error(scratch_err[device], "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
        return PLB_ERR_HIP;
    }
    return PLB_OK;
}

// out[c][r] = macros[c][internal id of V-order node first + r], c = 0..3.
__global__ void k_gather_macros4(const double* __restrict__ macros, int64_t Npad,
                                 const uint32_t* __restrict__ inv_perm, int64_t first, int64_t n,
                                 double* __restrict__ out)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int64_t g = first + r;
    const int64_t k = inv_perm ? (int64_t)inv_perm[g] : g;
#pragma unroll
    for (int comp = 0; comp < 4; comp++) out[comp * n + r] = macros[comp * Npad + k];
}

}  // namespace plb

using namespace plb;

extern "C" int plb_read_macros_noalloc(plb_ctx* c, int64_t first, int64_t n, double* rho, double* ux,
                                       double* uy, double* uz)
{
    if (!c || first < 0 || n < 0 || first + n > c->N) return PLB_ERR_INVALID;
    if (n > NOALLOC_MAX) {
        fprintf(stderr, "libplb: plb_read_macros_noalloc reads at most %lld nodes per call (asked %lld)\n",
                (long long)NOALLOC_MAX, (long long)n);
        return PLB_ERR_RANGE;
    }
    if (n == 0) return PLB_OK;
    if (c->device < 0 || c->device >= MAX_DEVICES || !scratch_ptr[c->device]) return PLB_ERR_HIP;
    PLB_HIP(hipSetDevice(c->device));
    double* d = scratch_ptr[c->device];
    hipLaunchKernelGGL(k_gather_macros4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->macros,
                       c->Npad, c->inv_perm, first, n, d);
    PLB_HIP(hipGetLastError());
    double* outs[4] = {rho, ux, uy, uz};
    for (int comp = 0; comp < 4; comp++)
        if (outs[comp])
            PLB_HIP(hipMemcpyAsync(outs[comp], d + comp * n, sizeof(double) * n, hipMemcpyDeviceToHost,
                                   c->stream));
    PLB_HIP(hipStreamSynchronize(c->stream));
    return PLB_OK;
}
