// traits.hpp — device workspace of the standalone trait contraction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/clotho_amd.h"

namespace clotho {

hipError_t launch_traits_dense_fp32(const uint64_t* rows, uint32_t h, uint32_t w,
                                    const float* eff, uint32_t sites, uint32_t traits, float* z,
                                    hipStream_t s);
hipError_t launch_traits_walk_fp32(const uint64_t* rows, uint32_t h, uint32_t w,
                                   const float* eff, uint32_t sites, uint32_t traits, float* z,
                                   hipStream_t s);
hipError_t launch_traits_walk_fp64(const uint64_t* rows, uint32_t h, uint32_t w,
                                   const double* eff, uint32_t sites, uint32_t traits, double* z,
                                   hipStream_t s);
// Exact traits of the pipeline (traits_mfma.hip): int8 digit planes of the
// quantised effects, int32 MFMA accumulation, int64 recombination.
uint32_t i8_ntiles(uint32_t traits);  // N tiles of the kernel for `traits` (1, 2, 4, 8)
size_t i8_packed_elems(uint32_t traits, uint32_t w_cap);  // uint4 elements
hipError_t launch_pack_i8(const int32_t* eff, uint32_t traits, const uint32_t* p_words,
                          const uint32_t* p_alleles, uint32_t w_cap, uint4* packed,
                          hipStream_t s);
hipError_t launch_traits_i8(const uint64_t* rows, uint32_t h, const uint32_t* p_words,
                            uint32_t traits, const uint4* packed, int64_t* out, hipStream_t s);
// MFMA contraction (traits_mfma.hip): effects pre-split into `terms` BF16
// matrices laid out for the kernel; FP64 partials per K chunk reduced in FP64.
size_t mfma_effect_elems(uint32_t words, int terms);
void mfma_pack_effects(const double* effects, uint32_t sites, uint32_t traits, uint32_t words,
                       int terms, uint16_t* packed);
size_t mfma_partial_elems(uint32_t h, uint32_t w, uint32_t traits);
hipError_t launch_traits_mfma(const uint64_t* rows, uint32_t h, uint32_t w,
                              const uint16_t* packed, uint32_t sites, uint32_t traits, int terms,
                              double* partial, double* z, hipStream_t s);

// Quantum exponent of the exact int8 contraction: the largest q with every
// |effect| 2^q <= 2^30 - 1.  Effects must be finite, and their largest
// magnitude must give q in [-60, 60] (roughly 2^-31 <= max |effect| <= 2^89);
// otherwise the call is rejected with CLOTHO_ERR_ARG before anything runs.
inline int i8_quantum_exponent(const double* eff, size_t n, int* q_out) {
    double mx = 0.0;
    for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(eff[i])) return CLOTHO_ERR_ARG;
        mx = std::max(mx, std::fabs(eff[i]));
    }
    int q = 0;
    if (mx > 0.0) {
        q = 29 - std::ilogb(mx);
        while (std::ldexp(mx, q) > 1073741823.0) --q;
        if (q < -60 || q > 60) return CLOTHO_ERR_ARG;
    }
    *q_out = q;
    return CLOTHO_OK;
}

struct TraitWorkspace {
    int mode = 0, terms = 0;
    uint32_t h = 0, w = 0, sites = 0, traits = 0;
    uint64_t* d_rows = nullptr;
    float* d_eff32 = nullptr;
    double* d_eff64 = nullptr;
    uint16_t* d_packed = nullptr;
    float* d_z32 = nullptr;
    double* d_z64 = nullptr;
    double* d_partial = nullptr;
    int32_t* d_qeff = nullptr;   // EXACT_I8: effects as integers k (effect ~ k 2^-q)
    uint32_t* d_dims = nullptr;  // EXACT_I8: {words, sites} as the pack kernel reads them
    uint4* d_packed8 = nullptr;  // EXACT_I8: int8 digit planes
    int64_t* d_sum = nullptr;    // EXACT_I8: exact sums of k
    int q = 0;

    int init(const uint64_t* rows, uint32_t h_, uint32_t w_, const double* eff, uint32_t s_,
             uint32_t t_, int mode_, int terms_, hipStream_t st) {
        mode = mode_;
        terms = terms_;
        h = h_;
        w = w_;
        sites = s_;
        traits = t_;
        if (hipMalloc((void**)&d_rows, sizeof(uint64_t) * h * w) != hipSuccess)
            return CLOTHO_ERR_HIP;
        if (rows && hipMemcpyAsync(d_rows, rows, sizeof(uint64_t) * h * w,
                                   hipMemcpyHostToDevice, st) != hipSuccess)
            return CLOTHO_ERR_HIP;
        if (mode == CLOTHO_TRAITS_PLAIN_FP32 || mode == CLOTHO_TRAITS_WALK_FP32) {
            std::vector<float> e((size_t)sites * traits);
            for (size_t i = 0; i < e.size(); ++i) e[i] = (float)eff[i];
            if (hipMalloc((void**)&d_eff32, 4 * e.size()) != hipSuccess ||
                hipMalloc((void**)&d_z32, 4 * (size_t)h * traits) != hipSuccess)
                return CLOTHO_ERR_HIP;
            if (hipMemcpy(d_eff32, e.data(), 4 * e.size(), hipMemcpyHostToDevice) != hipSuccess)
                return CLOTHO_ERR_HIP;
        } else if (mode == CLOTHO_TRAITS_EXACT_I8) {
            // largest q with every |effect| 2^q <= 2^30 - 1, then round to nearest
            if (i8_quantum_exponent(eff, (size_t)sites * traits, &q) != CLOTHO_OK)
                return CLOTHO_ERR_ARG;
            std::vector<int32_t> k((size_t)sites * traits);
            for (size_t i = 0; i < k.size(); ++i) k[i] = (int32_t)std::llround(std::ldexp(eff[i], q));
            const uint32_t dims[2] = {w, sites};
            if (hipMalloc((void**)&d_qeff, 4 * k.size()) != hipSuccess ||
                hipMalloc((void**)&d_dims, sizeof(dims)) != hipSuccess ||
                hipMalloc((void**)&d_packed8, sizeof(uint4) * i8_packed_elems(traits, w)) !=
                    hipSuccess ||
                hipMalloc((void**)&d_sum, 8 * (size_t)h * traits) != hipSuccess)
                return CLOTHO_ERR_HIP;
            if (hipMemcpy(d_qeff, k.data(), 4 * k.size(), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemcpy(d_dims, dims, sizeof(dims), hipMemcpyHostToDevice) != hipSuccess)
                return CLOTHO_ERR_HIP;
            // the digit planes are built once per effect set, as in a generation
            if (launch_pack_i8(d_qeff, traits, d_dims, d_dims + 1, w, d_packed8, st) != hipSuccess)
                return CLOTHO_ERR_HIP;
        } else if (mode == CLOTHO_TRAITS_EXACT_FP64) {
            if (hipMalloc((void**)&d_eff64, 8 * (size_t)sites * traits) != hipSuccess ||
                hipMalloc((void**)&d_z64, 8 * (size_t)h * traits) != hipSuccess)
                return CLOTHO_ERR_HIP;
            if (hipMemcpy(d_eff64, eff, 8 * (size_t)sites * traits, hipMemcpyHostToDevice) !=
                hipSuccess)
                return CLOTHO_ERR_HIP;
        } else {
            const size_t ne = mfma_effect_elems(w, terms);
            std::vector<uint16_t> packed(ne);
            mfma_pack_effects(eff, sites, traits, w, terms, packed.data());
            if (hipMalloc((void**)&d_packed, 2 * ne) != hipSuccess ||
                hipMalloc((void**)&d_partial, 8 * mfma_partial_elems(h, w, traits)) !=
                    hipSuccess ||
                hipMalloc((void**)&d_z64, 8 * (size_t)h * traits) != hipSuccess)
                return CLOTHO_ERR_HIP;
            if (hipMemcpy(d_packed, packed.data(), 2 * ne, hipMemcpyHostToDevice) != hipSuccess)
                return CLOTHO_ERR_HIP;
        }
        return hipStreamSynchronize(st) == hipSuccess ? CLOTHO_OK : CLOTHO_ERR_HIP;
    }

    hipError_t run(hipStream_t st) {
        if (mode == CLOTHO_TRAITS_PLAIN_FP32)
            return launch_traits_dense_fp32(d_rows, h, w, d_eff32, sites, traits, d_z32, st);
        if (mode == CLOTHO_TRAITS_WALK_FP32)
            return launch_traits_walk_fp32(d_rows, h, w, d_eff32, sites, traits, d_z32, st);
        if (mode == CLOTHO_TRAITS_EXACT_FP64)
            return launch_traits_walk_fp64(d_rows, h, w, d_eff64, sites, traits, d_z64, st);
        if (mode == CLOTHO_TRAITS_EXACT_I8)
            return launch_traits_i8(d_rows, h, d_dims, traits, d_packed8, d_sum, st);
        return launch_traits_mfma(d_rows, h, w, d_packed, sites, traits, terms, d_partial, d_z64,
                                  st);
    }

    int download(double* z, hipStream_t st) {
        const size_t n = (size_t)h * traits;
        if (hipStreamSynchronize(st) != hipSuccess) return CLOTHO_ERR_HIP;
        if (mode == CLOTHO_TRAITS_PLAIN_FP32 || mode == CLOTHO_TRAITS_WALK_FP32) {
            std::vector<float> tmp(n);
            if (hipMemcpy(tmp.data(), d_z32, 4 * n, hipMemcpyDeviceToHost) != hipSuccess)
                return CLOTHO_ERR_HIP;
            for (size_t i = 0; i < n; ++i) z[i] = tmp[i];
        } else if (mode == CLOTHO_TRAITS_EXACT_I8) {
            std::vector<int64_t> tmp(n);
            if (hipMemcpy(tmp.data(), d_sum, 8 * n, hipMemcpyDeviceToHost) != hipSuccess)
                return CLOTHO_ERR_HIP;
            // |sum| < 2^53 (sites < 2^23), so the conversion and the scaling are exact
            for (size_t i = 0; i < n; ++i) z[i] = std::ldexp((double)tmp[i], -q);
        } else {
            if (hipMemcpy(z, d_z64, 8 * n, hipMemcpyDeviceToHost) != hipSuccess)
                return CLOTHO_ERR_HIP;
        }
        return CLOTHO_OK;
    }

    void release() {
        void* ptrs[] = {d_rows, d_eff32, d_eff64, d_packed, d_z32, d_z64, d_partial,
                        d_qeff, d_dims, d_packed8, d_sum};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    }
};

}  // namespace clotho
