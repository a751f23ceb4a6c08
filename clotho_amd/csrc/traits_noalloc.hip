// traits_noalloc.hip — the exact trait contraction for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  No counterpart in the reference.
//
// The effect digit planes the int8 MFMA path multiplies by (and the two
// width words its kernels read from device memory) live in `scratch`, a
// file-scope device array of fixed size.  Each GPU holds its own copy until
// the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name on
// the host is a stand-in, not a device address).  Every call on one GPU
// shares that GPU's array: calls that overlap on one GPU race, and a call
// overwrites everything of the previous call that it reads (the width words
// and every plane its traits kernel multiplies by), so nothing needs to be
// cleared between calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/clotho_amd.h"
#include "traits.hpp"

namespace clotho {
namespace {

constexpr size_t kDimsBytes = 256;  // {words, sites}, padded to the planes' alignment
constexpr size_t kPlanesBytes =
    (size_t)((CLOTHO_NOALLOC_MAX_TRAITS + 3) / 4) * (CLOTHO_NOALLOC_MAX_SITES / 64 / 16) * 16 * 64 *
    sizeof(uint4);
constexpr size_t kScratchBytes = kDimsBytes + kPlanesBytes;
static_assert(kScratchBytes == CLOTHO_NOALLOC_SCRATCH_BYTES, "scratch size as declared");

__device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
uint8_t* g_scratch[kMaxDevices] = {};

// the device address of `scratch` on `device` (current), looked up once
int scratch_of(int device, uint8_t** out) {
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return CLOTHO_ERR_HIP;
        g_scratch[device] = static_cast<uint8_t*>(p);
    }
    *out = g_scratch[device];
    return CLOTHO_OK;
}

__global__ void k_set_dims(uint32_t* dims, uint32_t words, uint32_t sites) {
    if (threadIdx.x == 0) {
        dims[0] = words;
        dims[1] = sites;
    }
}

}  // namespace
}  // namespace clotho

using namespace clotho;

extern "C" int clotho_trait_contract_exact_noalloc(const uint64_t* rows, uint32_t haplotypes,
                                                   uint32_t words, const int32_t* effects,
                                                   uint32_t sites, uint32_t traits, int64_t* sums,
                                                   void* stream) {
    if (traits == 0 || words == 0 || words % 8 || (size_t)words * 64 < sites)
        return CLOTHO_ERR_ARG;
    if (traits > CLOTHO_NOALLOC_MAX_TRAITS || (size_t)words * 64 > CLOTHO_NOALLOC_MAX_SITES)
        return CLOTHO_ERR_CAPACITY;
    if (haplotypes == 0) return CLOTHO_OK;
    if (!rows || !sums || (sites && !effects)) return CLOTHO_ERR_ARG;
    int device = 0;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices)
        return CLOTHO_ERR_HIP;
    uint8_t* base = nullptr;
    const int rc = scratch_of(device, &base);
    if (rc != CLOTHO_OK) return rc;
    uint32_t* dims = reinterpret_cast<uint32_t*>(base);
    uint4* planes = reinterpret_cast<uint4*>(base + kDimsBytes);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_set_dims, dim3(1), dim3(64), 0, s, dims, words, sites);
    if (hipGetLastError() != hipSuccess) return CLOTHO_ERR_HIP;
    if (launch_pack_i8(effects, traits, dims, dims + 1, words, planes, s) != hipSuccess)
        return CLOTHO_ERR_HIP;
    if (launch_traits_i8(rows, haplotypes, dims, traits, planes, sums, s) != hipSuccess)
        return CLOTHO_ERR_HIP;
    return CLOTHO_OK;
}
