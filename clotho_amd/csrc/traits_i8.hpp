// traits_i8.hpp — exact genome traits on v_mfma_i32_16x16x64_i8 (device
// code shared by the pipeline's single-population and batched launches and
// the standalone contraction).
//
// Quantised int32 effects are written as four balanced base-256 digits in
// [-128, 127]; genotype bits expand to int8 0/1; each 16x16x64 product
// accumulates exactly in int32 and the digit planes recombine in int64, so
// the sums equal the sequential FP64 sums of the reference bit for bit
// (common_phenotype_accumulator,
// include/clotho/data_spaces/phenotype_evaluator/common_phenotype_accumulator_row_block_alignment.hpp:47-66).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace clotho {
namespace i8 {

typedef int i8x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));  // 16-byte loads and stores
typedef __attribute__((address_space(1))) u32x4 g_u32x4;

// Callers may hold these buffers as generic pointers read from device memory;
// viewing them in the global address space keeps every access a global (not
// FLAT) instruction.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* as_global(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// N tiles (four traits each) the traits kernels multiply by: 1, 2, 4 or 8.
// The digit planes are packed (zeros past the last trait) and allocated for
// this many tiles, so every plane a kernel reads was written.
__host__ __device__ inline uint32_t kernel_tiles(uint32_t traits) {
    return traits <= 4 ? 1u : traits <= 8 ? 2u : traits <= 16 ? 4u : 8u;
}

// four genotype bits -> four int8 0/1 bytes
__device__ inline uint32_t spread4(uint32_t nib) { return (nib * 0x00204081u) & 0x01010101u; }

__device__ inline int32_t balanced_digit(int32_t q, int p) {
    int32_t d = 0;
    for (int i = 0; i <= p; ++i) {
        d = ((q + 128) & 255) - 128;
        q = (q - d) >> 8;
    }
    return d;
}

// K-block order.  Rows are read 16 words (one 128-byte line) at a time per
// row: lane group g = lane / 16 loads words 4g..4g+3 of the group with two
// 16-byte loads, and K block q (of 16 per group) takes 16-bit slice q % 4 of
// the lane's word q / 4.  The site order inside a K block is free as long as
// the packed effects follow it, so no lane ever needs another lane's bits.
constexpr uint32_t kGroupWords = 16;

__device__ inline uint32_t i8_site(uint32_t G, uint32_t q, uint32_t g, uint32_t i) {
    return 64 * (kGroupWords * G + 4 * g + q / 4) + 16 * (q & 3) + i;
}

// B fragments of the exact path: packed[((nt * NG + G) * 16 + q) * 64 + lane]
// holds the 16 int8 digits for column n = lane % 16 (trait 4 nt + n / 4,
// digit plane n % 4) and sites i8_site(G, q, lane / 16, i).
__device__ __forceinline__ void pack_i8_block(const int32_t* __restrict__ eff_in, uint32_t T,
                                              uint32_t ntiles, const uint32_t* p_words,
                                              const uint32_t* p_alleles, uint32_t ng_cap,
                                              uint4* __restrict__ packed_in, uint32_t block) {
    const auto eff = as_global(eff_in);
    const auto packed = as_global(packed_in);
    const uint32_t W = (uint32_t)__builtin_amdgcn_readfirstlane((int)*as_global(p_words));
    const uint32_t A = (uint32_t)__builtin_amdgcn_readfirstlane((int)*as_global(p_alleles));
    const uint32_t NG = (W + kGroupWords - 1) / kGroupWords;
    const size_t id = (size_t)block * 256 + threadIdx.x;
    const uint32_t lane = (uint32_t)(id & 63);
    const uint32_t q = (uint32_t)((id >> 6) & 15);
    const size_t gn = id >> 10;
    const uint32_t G = (uint32_t)(gn % ng_cap), nt = (uint32_t)(gn / ng_cap);
    if (nt >= ntiles || G >= NG) return;
    const uint32_t n = lane & 15, g = lane >> 4;
    const uint32_t t = 4 * nt + n / 4, p = n & 3;
    uint32_t d[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < 16; ++i) {
        const uint32_t s = i8_site(G, q, g, i);
        const int32_t v = (s < A && t < T) ? eff[(size_t)s * T + t] : 0;
        d[i >> 2] |= ((uint32_t)balanced_digit(v, (int)p) & 0xFFu) << (8 * (i & 3));
    }
    *(g_u32x4*)(packed + (((size_t)nt * NG + G) * 16 + q) * 64 + lane) =
        u32x4{d[0], d[1], d[2], d[3]};
}

// Row tiles of one wave: 16-row tiles t0 .. t0 + cnt - 1 (cnt <= RT).
struct WaveTiles {
    uint32_t t0, cnt;
};

// RT consecutive tiles per wave, waves in block order.
__device__ inline WaveTiles wave_tiles_fixed(uint32_t block, uint32_t wave, uint32_t RT) {
    return WaveTiles{(block * 4 + wave) * RT, RT};
}

// The ceil(h / 16) tiles split over the 4 * blocks waves as evenly as
// possible (counts differ by at most one), so that a grid that is a multiple
// of the compute-unit count gives every unit the same number of rows.  The
// caller launches blocks >= ceil(h / (64 RT)), so that no wave gets more than
// RT tiles.
__device__ inline WaveTiles wave_tiles_even(uint32_t block, uint32_t wave, uint32_t h,
                                            uint32_t blocks) {
    const uint64_t tiles = (h + 15u) / 16u, waves = 4ull * blocks, i = 4ull * block + wave;
    const uint32_t t0 = (uint32_t)(i * tiles / waves);
    return WaveTiles{t0, (uint32_t)((i + 1) * tiles / waves) - t0};
}

// End of a stage: wait until at most N vector-memory loads are in flight,
// then until the wave's LDS accesses are done, then meet the block.  The
// builtin wait tells the compiler which loads have landed (so that it adds
// no wait of its own before the next stage's LDS reads); the asm statement
// repeats it with the barrier, and its memory clobber keeps the compiler
// from moving a memory access across.  __syncthreads instead waits for every
// load in flight (its fences cover global memory, and a fence restricted to
// LDS counts direct-to-LDS loads as pending LDS writes), so loads of the next
// group could not stay in flight across it.  The caller pins its loads with
// compiler barriers right after issuing them, so that none is issued
// between the wait and the barrier.
template <int N>
__device__ __forceinline__ void stage_barrier() {
    static_assert(N >= 0 && N < 16, "vmcnt immediate");
    __builtin_amdgcn_s_waitcnt(0x0F70 | N);
    asm volatile("s_waitcnt vmcnt(%0)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

// Exact traits: one wave = up to RT row tiles of 16 rows over all K; NT N tiles.
// The block stages the B fragments in LDS, shared by its four waves, in
// stages of 16 KB: stage s of group G holds q-blocks s QS .. s QS + QS - 1
// (QS = 16 / NT) of all NT tiles, double buffered and moved by direct
// global-to-LDS loads while the previous stage is multiplied.  A group's rows
// are loaded at its first stage for the next group.  Bits expand to int8 0/1
// through a 256-entry LDS table (byte -> 8 bytes).
//
// traits_i8_wave is one wave's share: rows base .. base + 16 cnt - 1 (cnt <=
// RT, base and cnt wave-uniform); s_b0, s_b1 (the two stage buffers) and
// s_lut are the block's tables.  The stage buffers are two LDS objects, not
// halves of one: a read of one buffer is then known not to touch the other,
// which direct-to-LDS loads are filling, and needs no wait for those loads
// (with an even tile count the buffer of every stage is a constant).
template <int RT, int NT>
__device__ __forceinline__ void traits_i8_wave(const uint64_t* __restrict__ rows_in, uint32_t h,
                                               uint32_t W, uint32_t T,
                                               const uint4* __restrict__ packed_in,
                                               int64_t* __restrict__ traits_in, uint32_t base,
                                               uint32_t cnt, u32x4* s_b0, u32x4* s_b1,
                                               uint2* s_lut) {
    static_assert(NT == 1 || NT == 2 || NT == 4 || NT == 8, "N tiles: 1, 2, 4 or 8");
    constexpr int QS = 16 / NT;  // q-blocks per stage
    const auto rows = as_global(rows_in);
    const auto packed = as_global(packed_in);
    const auto traits = as_global(traits_in);
    const uint32_t NG = (W + kGroupWords - 1) / kGroupWords;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t m = lane & 15, g = lane >> 4;
    {
        const uint32_t v = threadIdx.x;
        s_lut[v] = make_uint2(spread4(v & 15), spread4(v >> 4));
    }
    i8x4 acc[RT][NT];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[r][nt] = i8x4{0, 0, 0, 0};
    const __attribute__((address_space(1))) uint64_t* rp[RT];
    bool rv[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const uint32_t row = base + 16 * r + m;
        rv[r] = (uint32_t)r < cnt && row < h;
        rp[r] = rows + (size_t)(rv[r] ? row : 0) * W;
    }
    auto load_rows = [&](uint32_t G, u32x4 (&x)[RT][2]) {
        const uint32_t w = kGroupWords * G + 4 * g;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            if (rv[r] && w < W) {
                const g_u32x4* src = (const g_u32x4*)(rp[r] + w);
                x[r][0] = src[0];
                x[r][1] = src[1];
            } else {
                x[r][0] = u32x4{0, 0, 0, 0};
                x[r][1] = u32x4{0, 0, 0, 0};
            }
        }
    };
    // With several stages per group (NT >= 2) the next group's rows are
    // loaded after the first stage's B loads (a compiler barrier keeps that
    // order), all 2 RT loads issued in every group (the group and word
    // clamped in bounds: W is a multiple of 8) and the invalid ones zeroed
    // when the group is taken over.  The wait that ends the first stage then
    // leaves exactly these loads in flight, so they have two stages of
    // products to land in instead of one (loads complete in issue order: a
    // wait for a later B load is a wait for every earlier row load).
    // (With eight stages per group the rows would gain one stage in eight,
    // and the 32-trait kernel would lose a wave per SIMD to the registers.)
    constexpr bool kLateRows = NT == 2 || NT == 4;
    auto load_rows_all = [&](uint32_t G, u32x4 (&x)[RT][2]) {
        const uint32_t w = min(kGroupWords * G + 4 * g, W - 4);
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            const g_u32x4* src = (const g_u32x4*)(rp[r] + w);
            x[r][0] = src[0];
            x[r][1] = src[1];
        }
    };
    // stage k = (group k / NT, part k mod NT) into buffer k & 1: piece i =
    // (tile i / QS, q-block part QS + i mod QS), 1 KB each, four per wave
    const uint32_t n_stages = NG * NT;
    // four pieces per wave in a loop of known trip count (the block has four
    // waves) over a wave-uniform index: a loop the compiler takes as
    // divergent, of unknown length, leaves it unable to count the loads in
    // flight, and every later wait for a load becomes a wait for all of them
    const uint32_t wave_s = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave);
    auto stage_b_at = [&](uint32_t k) {  // k < n_stages
        const uint32_t G = k / NT, q0 = (k % NT) * QS;
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t i = wave_s + 4 * j;
            const uint32_t nt = i / QS, q = q0 + i % QS;
            const g_u32x4* src =
                (const g_u32x4*)(packed + (((size_t)nt * NG + G) * 16 + q) * 64) + lane;
            __builtin_amdgcn_global_load_lds(
                (const __attribute__((address_space(1))) void*)src,
                (__attribute__((address_space(3))) void*)(((k & 1) ? s_b1 : s_b0) + i * 64), 16,
                0, 0);
        }
    };
    auto stage_b = [&](uint32_t k) {
        if (k < n_stages) stage_b_at(k);
    };
    // vmcnt(0) only (expcnt and lgkmcnt at their maxima): the direct-to-LDS
    // loads and the next group's rows have landed
    auto wait_loads = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };
    u32x4 x[RT][2];
    if (NG) load_rows(0, x);
    stage_b(0);
    wait_loads();
    __syncthreads();
    for (uint32_t G = 0; G < NG; ++G) {
        u32x4 xn[RT][2];
        if (!kLateRows && G + 1 < NG) load_rows(G + 1, xn);
#pragma unroll
        for (int s = 0; s < NT; ++s) {
            const uint32_t k = G * NT + s;
            if (kLateRows && s == 0) {
                stage_b_at(k + 1);  // G NT + 1 < NG NT
                asm volatile("" ::: "memory");
                load_rows_all(min(G + 1, NG - 1), xn);
                asm volatile("" ::: "memory");
            } else {
                stage_b(k + 1);
                if constexpr (kLateRows) asm volatile("" ::: "memory");
            }
#pragma unroll
            for (int qq = 0; qq < QS; ++qq) {
                const int q = s * QS + qq;
                i8x4 a[RT];
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    const u32x4 v = x[r][q >> 3];
                    const uint32_t dw = ((q >> 1) & 3) == 0 ? v.x
                                      : ((q >> 1) & 3) == 1 ? v.y
                                      : ((q >> 1) & 3) == 2 ? v.z : v.w;
                    const uint32_t b16 = (q & 1) ? (dw >> 16) : (dw & 0xFFFFu);
                    const uint2 lo = s_lut[b16 & 0xFFu], hi = s_lut[b16 >> 8];
                    a[r] = i8x4{(int)lo.x, (int)lo.y, (int)hi.x, (int)hi.y};
                }
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) {
                    const u32x4 bq = ((k & 1) ? s_b1 : s_b0)[(nt * QS + qq) * 64 + lane];
                    const i8x4 bv = i8x4{(int)bq.x, (int)bq.y, (int)bq.z, (int)bq.w};
#pragma unroll
                    for (int r = 0; r < RT; ++r)
                        acc[r][nt] =
                            __builtin_amdgcn_mfma_i32_16x16x64_i8(a[r], bv, acc[r][nt], 0, 0, 0);
                }
            }
            // stage k + 1 complete; buffer k & 1 free
            if constexpr (kLateRows) {
                if (s == 0)
                    stage_barrier<2 * RT>();  // the next group's rows stay in flight
                else
                    stage_barrier<0>();
            } else {
                wait_loads();
                __syncthreads();
            }
        }
        if constexpr (kLateRows) {
            const bool wok = kGroupWords * (G + 1) + 4 * g < W;
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                const bool ok = rv[r] && wok;
                x[r][0] = ok ? xn[r][0] : u32x4{0, 0, 0, 0};
                x[r][1] = ok ? xn[r][1] : u32x4{0, 0, 0, 0};
            }
        } else {
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                x[r][0] = xn[r][0];
                x[r][1] = xn[r][1];
            }
        }
    }
    // recombine the digit planes: column n = 4 t_local + p held by lane m = n
#pragma unroll
    for (int r = 0; r < RT; ++r) {
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int64_t c0 = acc[r][nt][i];
                const int64_t c1 = __shfl_down((int)acc[r][nt][i], 1, 64);
                const int64_t c2 = __shfl_down((int)acc[r][nt][i], 2, 64);
                const int64_t c3 = __shfl_down((int)acc[r][nt][i], 3, 64);
                const uint32_t row = base + 16 * r + 4 * g + i;
                const uint32_t t = 4 * nt + m / 4;
                if ((m & 3) == 0 && (uint32_t)r < cnt && row < h && t < T)
                    traits[(size_t)row * T + t] = c0 + (c1 << 8) + (c2 << 16) + (c3 << 24);
            }
        }
    }
}

// Exact traits of the rows of WaveTiles wt, one wave each.  kSplit: a wave
// of fewer than RT tiles runs the RT - 1 body, so that it issues no lookups
// or products for an empty tile (both bodies pass the same barriers).
template <int RT, int NT, bool kSplit = false>
__device__ __forceinline__ void traits_i8_block(const uint64_t* __restrict__ rows, uint32_t h,
                                                const uint32_t* p_words, uint32_t T,
                                                const uint4* __restrict__ packed,
                                                int64_t* __restrict__ traits, WaveTiles wt) {
    __shared__ u32x4 s_b0[16 * 64];
    __shared__ u32x4 s_b1[16 * 64];
    __shared__ uint2 s_lut[256];
    const uint32_t W = (uint32_t)__builtin_amdgcn_readfirstlane((int)*as_global(p_words));
    const uint32_t cnt = (uint32_t)__builtin_amdgcn_readfirstlane((int)min(wt.cnt, (uint32_t)RT));
    const uint32_t base = 16u * (uint32_t)__builtin_amdgcn_readfirstlane((int)wt.t0);
    if constexpr (kSplit && RT > 1) {
        if (cnt < (uint32_t)RT) {
            traits_i8_wave<RT - 1, NT>(rows, h, W, T, packed, traits, base, cnt, s_b0, s_b1,
                                       s_lut);
            return;
        }
    }
    traits_i8_wave<RT, NT>(rows, h, W, T, packed, traits, base, cnt, s_b0, s_b1, s_lut);
}

}  // namespace i8
}  // namespace clotho
