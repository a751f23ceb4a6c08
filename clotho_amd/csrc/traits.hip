// traits.hip — genotype x effect contraction over a resident population:
// z[h][t] = sum_j bit(h, j) * effect[j][t], the population-wide form of
// common_phenotype_accumulator::operator()
// (include/clotho/data_spaces/phenotype_evaluator/common_phenotype_accumulator_row_block_alignment.hpp:47-66).
//
// Modes: PLAIN_FP32 walks the set bits and sums FP32 effects (the plain
// popcount-weighted baseline); EXACT_FP64 does the same walk in FP64;
// MFMA_BF16 expands bits to BF16 and runs the contraction on matrix cores
// with effects split into BF16 terms (traits_mfma.hip); EXACT_I8 is the
// pipeline's exact path (int8 digit planes of quantised effects).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/clotho_amd.h"
#include "device_common.hpp"
#include "traits.hpp"

namespace clotho {
namespace {

constexpr int kBlock = 256;

template <typename Acc, int TMAX>
__global__ void __launch_bounds__(kBlock) k_traits_walk(const uint64_t* __restrict__ rows,
                                                        uint32_t h, uint32_t w,
                                                        const Acc* __restrict__ eff,
                                                        uint32_t sites, uint32_t traits,
                                                        Acc* __restrict__ z) {
    const uint32_t row = (blockIdx.x * kBlock + threadIdx.x) >> 6;
    const uint32_t lane = threadIdx.x & 63;
    if (row >= h) return;
    Acc acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = 0;
    const uint64_t* r = rows + (size_t)row * w;
    for (uint32_t k = lane; k < w; k += 64) {
        uint64_t x = r[k];
        while (x) {
            const uint32_t j = k * 64 + (uint32_t)__builtin_ctzll(x);
            x &= x - 1;
            if (j >= sites) break;
            const Acc* e = eff + (size_t)j * traits;
#pragma unroll
            for (int t = 0; t < TMAX; ++t)
                if (t < (int)traits) acc[t] += e[t];
        }
    }
#pragma unroll
    for (int t = 0; t < TMAX; ++t) {
        if (t >= (int)traits) break;
        Acc s = acc[t];
        for (int d = 32; d > 0; d >>= 1) s += __shfl_xor(s, d, 64);
        if (lane == 0) z[(size_t)row * traits + t] = s;
    }
}

// Dense popcount-weighted FP32 contraction: lane = haplotype row, effects of
// a 64-site block staged in LDS and broadcast, z_t = fma(bit, w_t, z_t) for
// every site (the reference's per-sequence bit x effect accumulation,
// CU/phenotype/translation_kernels.hpp:123-181, generalised to T traits).
// The library builds with -ffp-contract=off for bit-exact parity elsewhere,
// so the fused multiply-add is written out: one v_fma_f32 per site and trait.
template <int TMAX>
__global__ void __launch_bounds__(kBlock) k_traits_dense(const uint64_t* __restrict__ rows,
                                                         uint32_t h, uint32_t w,
                                                         const float* __restrict__ eff,
                                                         uint32_t sites, uint32_t traits,
                                                         float* __restrict__ z) {
    __shared__ float s_eff[64 * TMAX];
    const uint32_t row = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = row < h;
    const uint64_t* r = rows + (size_t)(valid ? row : 0) * w;
    float acc[TMAX];
#pragma unroll
    for (int t = 0; t < TMAX; ++t) acc[t] = 0.f;
    for (uint32_t k = 0; k < w; ++k) {
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < 64 * TMAX; i += kBlock) {
            const uint32_t j = k * 64 + i / TMAX, t = i % TMAX;
            s_eff[i] = (j < sites && t < traits) ? eff[(size_t)j * traits + t] : 0.f;
        }
        __syncthreads();
        const uint64_t x = valid ? r[k] : 0;
        for (uint32_t j = 0; j < 64; ++j) {
            const float bit = (float)((uint32_t)(x >> j) & 1u);
#pragma unroll
            for (int t = 0; t < TMAX; ++t) acc[t] = __builtin_fmaf(bit, s_eff[j * TMAX + t], acc[t]);
        }
    }
    if (valid)
        for (uint32_t t = 0; t < traits && t < (uint32_t)TMAX; ++t)
            z[(size_t)row * traits + t] = acc[t];
}

}  // namespace

hipError_t launch_traits_dense_fp32(const uint64_t* rows, uint32_t h, uint32_t w,
                                    const float* eff, uint32_t sites, uint32_t traits, float* z,
                                    hipStream_t s) {
    const uint32_t blocks = (h + kBlock - 1) / kBlock;
    if (traits <= 1)
        hipLaunchKernelGGL((k_traits_dense<1>), dim3(blocks), dim3(kBlock), 0, s, rows, h, w, eff,
                           sites, traits, z);
    else if (traits <= 4)
        hipLaunchKernelGGL((k_traits_dense<4>), dim3(blocks), dim3(kBlock), 0, s, rows, h, w, eff,
                           sites, traits, z);
    else
        hipLaunchKernelGGL((k_traits_dense<16>), dim3(blocks), dim3(kBlock), 0, s, rows, h, w, eff,
                           sites, traits, z);
    return hipGetLastError();
}

hipError_t launch_traits_walk_fp32(const uint64_t* rows, uint32_t h, uint32_t w,
                                   const float* eff, uint32_t sites, uint32_t traits, float* z,
                                   hipStream_t s) {
    const uint32_t blocks = (h + 3) / 4;
    if (traits <= 1)
        hipLaunchKernelGGL((k_traits_walk<float, 1>), dim3(blocks), dim3(kBlock), 0, s, rows, h, w,
                           eff, sites, traits, z);
    else if (traits <= 4)
        hipLaunchKernelGGL((k_traits_walk<float, 4>), dim3(blocks), dim3(kBlock), 0, s, rows, h, w,
                           eff, sites, traits, z);
    else
        hipLaunchKernelGGL((k_traits_walk<float, 16>), dim3(blocks), dim3(kBlock), 0, s, rows, h,
                           w, eff, sites, traits, z);
    return hipGetLastError();
}

hipError_t launch_traits_walk_fp64(const uint64_t* rows, uint32_t h, uint32_t w,
                                   const double* eff, uint32_t sites, uint32_t traits, double* z,
                                   hipStream_t s) {
    const uint32_t blocks = (h + 3) / 4;
    if (traits <= 1)
        hipLaunchKernelGGL((k_traits_walk<double, 1>), dim3(blocks), dim3(kBlock), 0, s, rows, h,
                           w, eff, sites, traits, z);
    else
        hipLaunchKernelGGL((k_traits_walk<double, 16>), dim3(blocks), dim3(kBlock), 0, s, rows, h,
                           w, eff, sites, traits, z);
    return hipGetLastError();
}

// Genotype bits drawn on the device: bit j of row r set when Philox word
// (seed, generation 0, stream 7, replicate 0, index r, sub j/4)[j%4] < thr[j].
__global__ void __launch_bounds__(kBlock) k_synth_genotypes(uint64_t seed, const uint32_t* thr,
                                                            uint32_t sites, uint32_t h,
                                                            uint32_t words, uint64_t* rows) {
    const size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= (size_t)h * words) return;
    const uint32_t r = (uint32_t)(id / words), w = (uint32_t)(id % words);
    uint64_t x = 0;
    for (uint32_t q = 0; q < 16; ++q) {
        const uint32_t j0 = w * 64 + q * 4;
        if (j0 >= sites) break;
        const Draw4 d = philox_draw(seed, 0, 7, 0, r, j0 / 4);
        for (uint32_t k = 0; k < 4; ++k)
            if (j0 + k < sites && d.x[k] < thr[j0 + k]) x |= (uint64_t)1 << (q * 4 + k);
    }
    rows[id] = x;
}

}  // namespace clotho

using namespace clotho;

extern "C" int clotho_trait_contract(const uint64_t* rows, uint32_t h, uint32_t w,
                                     const double* effects, uint32_t sites, uint32_t traits,
                                     int mode, int split_terms, int repeats, double* z,
                                     double* avg_ms, int device) {
    // rows are loaded four words (two 16-byte loads) at a time: whole
    // 512-bit row blocks, as the reference's row_block_alignment pads them
    if (!rows || !effects || !z || h == 0 || w == 0 || w % 8 || traits == 0 || traits > 16 ||
        (size_t)w * 64 < sites || repeats < 1)
        return CLOTHO_ERR_ARG;
    if (mode != CLOTHO_TRAITS_PLAIN_FP32 && mode != CLOTHO_TRAITS_EXACT_FP64 &&
        mode != CLOTHO_TRAITS_MFMA_BF16 && mode != CLOTHO_TRAITS_WALK_FP32 &&
        mode != CLOTHO_TRAITS_EXACT_I8)
        return CLOTHO_ERR_ARG;
    if (mode == CLOTHO_TRAITS_EXACT_I8 && sites >= (1u << 23)) return CLOTHO_ERR_ARG;
    if (mode == CLOTHO_TRAITS_MFMA_BF16 && (split_terms < 1 || split_terms > 3))
        return CLOTHO_ERR_ARG;
    if (mode == CLOTHO_TRAITS_EXACT_I8) {
        int q = 0;
        if (i8_quantum_exponent(effects, (size_t)sites * traits, &q) != CLOTHO_OK)
            return CLOTHO_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) return CLOTHO_ERR_HIP;
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return CLOTHO_ERR_HIP;
    TraitWorkspace ws;
    int rc = ws.init(rows, h, w, effects, sites, traits, mode, split_terms, s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double ms_total = 0.0;
    if (rc == CLOTHO_OK) {
        // one untimed launch, then `repeats` timed launches
        if (ws.run(s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) rc = CLOTHO_ERR_HIP;
    }
    if (rc == CLOTHO_OK) {
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
            rc = CLOTHO_ERR_HIP;
    }
    if (rc == CLOTHO_OK) {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < repeats && rc == CLOTHO_OK; ++r)
            if (ws.run(s) != hipSuccess) rc = CLOTHO_ERR_HIP;
        (void)hipEventRecord(e1, s);
        if (rc == CLOTHO_OK && hipEventSynchronize(e1) != hipSuccess) rc = CLOTHO_ERR_HIP;
        float ms = 0.f;
        if (rc == CLOTHO_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) rc = CLOTHO_ERR_HIP;
        ms_total = ms;
    }
    if (rc == CLOTHO_OK) rc = ws.download(z, s);
    if (avg_ms) *avg_ms = ms_total / repeats;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    ws.release();
    (void)hipStreamDestroy(s);
    return rc;
}

extern "C" int clotho_trait_contract_synthetic(uint32_t h, uint32_t sites, const double* freqs,
                                               uint64_t seed, const double* effects,
                                               uint32_t traits, int mode, int split_terms,
                                               int repeats, double* z, double* avg_ms,
                                               int device) {
    if (!freqs || !effects || !z || h == 0 || sites == 0 || traits == 0 || traits > 16 ||
        repeats < 1)
        return CLOTHO_ERR_ARG;
    if (mode != CLOTHO_TRAITS_PLAIN_FP32 && mode != CLOTHO_TRAITS_EXACT_FP64 &&
        mode != CLOTHO_TRAITS_MFMA_BF16 && mode != CLOTHO_TRAITS_WALK_FP32 &&
        mode != CLOTHO_TRAITS_EXACT_I8)
        return CLOTHO_ERR_ARG;
    if (mode == CLOTHO_TRAITS_EXACT_I8 && sites >= (1u << 23)) return CLOTHO_ERR_ARG;
    if (mode == CLOTHO_TRAITS_MFMA_BF16 && (split_terms < 1 || split_terms > 3))
        return CLOTHO_ERR_ARG;
    if (mode == CLOTHO_TRAITS_EXACT_I8) {
        int q = 0;
        if (i8_quantum_exponent(effects, (size_t)sites * traits, &q) != CLOTHO_OK)
            return CLOTHO_ERR_ARG;
    }
    if (hipSetDevice(device) != hipSuccess) return CLOTHO_ERR_HIP;
    const uint32_t words = ((sites + 511) / 512) * 8;
    std::vector<uint32_t> thr(sites);
    for (uint32_t j = 0; j < sites; ++j) {
        const double f = std::min(std::max(freqs[j], 0.0), 1.0);
        thr[j] = f >= 1.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(f * 4294967296.0);
    }
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return CLOTHO_ERR_HIP;
    uint32_t* d_thr = nullptr;
    int rc = CLOTHO_OK;
    TraitWorkspace ws;
    if (hipMalloc((void**)&d_thr, 4 * (size_t)sites) != hipSuccess ||
        hipMemcpy(d_thr, thr.data(), 4 * (size_t)sites, hipMemcpyHostToDevice) != hipSuccess)
        rc = CLOTHO_ERR_HIP;
    if (rc == CLOTHO_OK) rc = ws.init(nullptr, h, words, effects, sites, traits, mode, split_terms, s);
    if (rc == CLOTHO_OK) {
        const size_t n = (size_t)h * words;
        hipLaunchKernelGGL(k_synth_genotypes, dim3((uint32_t)((n + kBlock - 1) / kBlock)),
                           dim3(kBlock), 0, s, seed, d_thr, sites, h, words, ws.d_rows);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
            rc = CLOTHO_ERR_HIP;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == CLOTHO_OK && (ws.run(s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess))
        rc = CLOTHO_ERR_HIP;
    if (rc == CLOTHO_OK &&
        (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = CLOTHO_ERR_HIP;
    if (rc == CLOTHO_OK) {
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < repeats && rc == CLOTHO_OK; ++r)
            if (ws.run(s) != hipSuccess) rc = CLOTHO_ERR_HIP;
        (void)hipEventRecord(e1, s);
        float ms = 0.f;
        if (rc == CLOTHO_OK && (hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = CLOTHO_ERR_HIP;
        if (avg_ms) *avg_ms = ms / repeats;
    }
    if (rc == CLOTHO_OK) rc = ws.download(z, s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (d_thr) (void)hipFree(d_thr);
    ws.release();
    (void)hipStreamDestroy(s);
    return rc;
}
