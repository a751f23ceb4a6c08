// traits_mfma.hip — genotype x effect contraction on CDNA4 matrix cores.
//
// z[h][t] = sum_j bit(h, j) * effect[j][t] over bit-packed rows (the
// population-wide common_phenotype_accumulator,
// include/clotho/data_spaces/phenotype_evaluator/common_phenotype_accumulator_row_block_alignment.hpp:47-66).
//
// Exact path (v_mfma_i32_16x16x64_i8): quantised int32 effects are written as
// four balanced base-256 digits d_p in [-128, 127]; genotype bits expand to
// int8 0/1; each 16x16x64 product accumulates exactly in int32 and the digit
// planes recombine in int64, so the sums equal the FP64 sequential sums of
// the reference bit for bit (effects are multiples of 2^q).  One N tile holds
// 4 traits x 4 digit planes.
//
// Split-BF16 path (v_mfma_f32_16x16x32_bf16): bits expand to BF16 0/1 (exact),
// effects split into 1-3 BF16 terms hi + mid + lo, FP32 accumulation per
// 64-word K chunk, chunks reduced in FP64.
//
// Operand layouts (16x16xK, wave64): lane l supplies A row (l % 16) and B
// column (l % 16) for K elements [K/4 * (l / 16), K/4 * (l / 16 + 1)); the
// accumulator lane l holds column (l % 16), rows 4 * (l / 16) + i.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "pipeline.hpp"
#include "traits.hpp"
#include "traits_i8.hpp"

namespace clotho {
namespace {

using i8::kGroupWords;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));


// two genotype bits -> two BF16 0/1 halves (1.0 = 0x3F80)
__device__ inline uint32_t bits2_bf16(uint32_t two) {
    return ((two * 0x8001u) & 0x10001u) * 0x3F80u;
}


// Exact-path kernels: bodies in traits_i8.hpp.
__global__ void __launch_bounds__(256) k_pack_i8(const int32_t* __restrict__ eff, uint32_t T,
                                                 uint32_t ntiles, const uint32_t* p_words,
                                                 const uint32_t* p_alleles, uint32_t ng_cap,
                                                 uint4* __restrict__ packed) {
    i8::pack_i8_block(eff, T, ntiles, p_words, p_alleles, ng_cap, packed, blockIdx.x);
}

template <int RT, int NT>
__global__ void __launch_bounds__(256) k_traits_i8(const uint64_t* __restrict__ rows, uint32_t h,
                                                   const uint32_t* p_words, uint32_t T,
                                                   const uint4* __restrict__ packed,
                                                   int64_t* __restrict__ traits) {
    i8::traits_i8_block<RT, NT>(rows, h, p_words, T, packed, traits,
                                i8::wave_tiles_fixed(blockIdx.x, threadIdx.x >> 6, RT));
}


// Split-BF16 contraction, K chunked: block = 4 waves x RT row tiles, one
// chunk of kChunkGroups 16-word groups.  K block q (of 32 per group) takes
// byte q % 8 of the lane's word q / 8, i.e. sites bf_site(G, q, lane / 16, i);
// bits expand to BF16 0/1 through a 256-entry LDS table (byte -> 8 BF16).
// The B fragments (TERMS per K block, 1 KB each) are staged in LDS for the
// whole block, kStageBlocks K blocks at a time, double buffered and moved by
// direct global-to-LDS loads while the previous stage is multiplied: every
// wave reads its B operands from LDS instead of each wave fetching them from
// L2.  The FP32 accumulators are flushed into FP64 after every stage (256
// sites), so FP32 rounding acts only on short partial sums; FP64
// partial[chunk][row][t].
constexpr uint32_t kChunkGroups = 16;
constexpr uint32_t kStageBlocks = 8;

__host__ __device__ inline uint32_t bf_site(uint32_t G, uint32_t q, uint32_t g, uint32_t i) {
    return 64 * (16 * G + 4 * g + q / 8) + 8 * (q & 7) + i;
}

template <int RT, int TERMS>
__global__ void __launch_bounds__(256) k_traits_bf16(const uint64_t* __restrict__ rows_in,
                                                     uint32_t h, uint32_t W, uint32_t T,
                                                     const uint4* __restrict__ packed_in,
                                                     double* __restrict__ partial) {
    constexpr uint32_t QS = kStageBlocks;
    constexpr uint32_t FR = QS * TERMS;       // 1 KB fragments per stage
    constexpr uint32_t SPG = 32 / QS;         // stages per group
    static_assert(FR % 4 == 0 && SPG % 2 == 0, "stage pieces per wave, stage parity per group");
    // the two stage buffers are two LDS objects, not halves of one: a read of
    // one is then known not to touch the other, which direct-to-LDS loads are
    // filling, and needs no wait for those loads (SPG is even, so the buffer
    // of every stage is a constant)
    __shared__ uint4 s_b0[FR * 64];
    __shared__ uint4 s_b1[FR * 64];
    __shared__ uint4 s_lut[256];
    const auto rows = i8::as_global(rows_in);
    const auto packed = i8::as_global(packed_in);
    const uint32_t NG = (W + 15) / 16;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t m = lane & 15, g = lane >> 4;
    const uint32_t base = (blockIdx.x * 4 + wave) * (16 * RT);
    const uint32_t G0 = blockIdx.y * kChunkGroups;
    const uint32_t G1 = min(NG, G0 + kChunkGroups);
    {
        const uint32_t b8 = threadIdx.x;
        s_lut[b8] = make_uint4(bits2_bf16(b8 & 3), bits2_bf16((b8 >> 2) & 3),
                               bits2_bf16((b8 >> 4) & 3), bits2_bf16(b8 >> 6));
    }
    double acc64[RT][4];
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) acc64[r][i] = 0.0;
    const __attribute__((address_space(1))) uint64_t* rp[RT];
    bool rv[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        const uint32_t row = base + 16 * r + m;
        rv[r] = row < h;
        rp[r] = rows + (size_t)(rv[r] ? row : 0) * W;
    }
    auto load_rows = [&](uint32_t G, i8::u32x4 (&x)[RT][2]) {
        const uint32_t w = 16 * G + 4 * g;
#pragma unroll
        for (int r = 0; r < RT; ++r) {
            if (rv[r] && w < W) {
                const i8::g_u32x4* src = (const i8::g_u32x4*)(rp[r] + w);
                x[r][0] = src[0];
                x[r][1] = src[1];
            } else {
                x[r][0] = i8::u32x4{0, 0, 0, 0};
                x[r][1] = i8::u32x4{0, 0, 0, 0};
            }
        }
    };
    // stage k of the chunk: group G0 + k / SPG, K blocks (k % SPG) QS ..
    // + QS - 1, every term: FR contiguous 1 KB fragments, wave v moving
    // fragments v, v + 4, ...
    const uint32_t n_stages = (G1 > G0 ? G1 - G0 : 0) * SPG;
    // FR / 4 pieces per wave in a loop of known trip count over a
    // wave-uniform index (a loop of unknown length leaves the compiler unable
    // to count the loads in flight)
    const uint32_t wave_s = (uint32_t)__builtin_amdgcn_readfirstlane((int)wave);
    auto stage_b = [&](uint32_t k) {
        if (k < n_stages) {
            const uint32_t G = G0 + k / SPG, qb = (k % SPG) * QS;
            const auto src = packed + (((size_t)G * 32 + qb) * TERMS) * 64 + lane;
#pragma unroll
            for (uint32_t j = 0; j < FR / 4; ++j) {
                const uint32_t i = wave_s + 4 * j;
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(src + (size_t)i * 64),
                    (__attribute__((address_space(3))) void*)(((k & 1) ? s_b1 : s_b0) + i * 64),
                    16, 0, 0);
            }
        }
    };
    // vmcnt(0) only: the direct-to-LDS loads and the next group's rows have landed
    auto wait_loads = [] { __builtin_amdgcn_s_waitcnt(0x0F70); };
    i8::u32x4 x[RT][2];
    if (n_stages) load_rows(G0, x);
    stage_b(0);
    wait_loads();
    __syncthreads();
    for (uint32_t G = G0; G < G1; ++G) {
        i8::u32x4 xn[RT][2];
        if (G + 1 < G1) load_rows(G + 1, xn);
#pragma unroll
        for (uint32_t st = 0; st < SPG; ++st) {
            const uint32_t k = (G - G0) * SPG + st;
            stage_b(k + 1);
            v4f acc[RT];
#pragma unroll
            for (int r = 0; r < RT; ++r) acc[r] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (uint32_t qq = 0; qq < QS; ++qq) {
                const uint32_t q = st * QS + qq;
                v8bf a[RT];
#pragma unroll
                for (int r = 0; r < RT; ++r) {
                    const i8::u32x4 v = x[r][q >> 4];
                    const uint32_t sel = (q >> 2) & 3;
                    const uint32_t dw = sel == 0 ? v.x : sel == 1 ? v.y : sel == 2 ? v.z : v.w;
                    const uint32_t b8 = (dw >> (8 * (q & 3))) & 0xFFu;
                    const uint4 u = s_lut[b8];
                    __builtin_memcpy(&a[r], &u, 16);
                }
#pragma unroll
                for (int tm = 0; tm < TERMS; ++tm) {
                    const uint4 bq = ((k & 1) ? s_b1 : s_b0)[(qq * TERMS + tm) * 64 + lane];
                    v8bf bv;
                    __builtin_memcpy(&bv, &bq, 16);
#pragma unroll
                    for (int r = 0; r < RT; ++r)
                        acc[r] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[r], bv, acc[r], 0, 0, 0);
                }
            }
#pragma unroll
            for (int r = 0; r < RT; ++r)
#pragma unroll
                for (int i = 0; i < 4; ++i) acc64[r][i] += (double)acc[r][i];
            wait_loads();
            __syncthreads();  // stage k + 1 complete; buffer k & 1 free
        }
        if (G + 1 < G1) {
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                x[r][0] = xn[r][0];
                x[r][1] = xn[r][1];
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RT; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const uint32_t row = base + 16 * r + 4 * g + i;
            if (row < h && m < T)
                partial[((size_t)blockIdx.y * h + row) * T + m] = acc64[r][i];
        }
}

__global__ void __launch_bounds__(256) k_reduce_chunks(const double* __restrict__ partial,
                                                       uint32_t chunks, size_t n,
                                                       double* __restrict__ z) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) s += partial[(size_t)c * n + i];
    z[i] = s;
}

uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    const uint32_t r = u + 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(r >> 16);
}
float bf16_to_f(uint16_t b) {
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

}  // namespace

// ---- exact int8 path (used by the pipeline) -------------------------------
uint32_t i8_ntiles(uint32_t traits) { return i8::kernel_tiles(traits); }

size_t i8_packed_elems(uint32_t traits, uint32_t w_cap) {
    const size_t ng = (w_cap + kGroupWords - 1) / kGroupWords;
    return (size_t)i8_ntiles(traits) * ng * 16 * 64;
}

hipError_t launch_pack_i8(const int32_t* eff, uint32_t traits, const uint32_t* p_words,
                          const uint32_t* p_alleles, uint32_t w_cap, uint4* packed,
                          hipStream_t s) {
    const uint32_t nt = i8_ntiles(traits);
    const uint32_t ng_cap = (w_cap + kGroupWords - 1) / kGroupWords;
    const size_t n = (size_t)nt * ng_cap * 16 * 64;
    hipLaunchKernelGGL(k_pack_i8, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, eff, traits,
                       nt, p_words, p_alleles, ng_cap, packed);
    return hipGetLastError();
}

hipError_t launch_traits_i8(const uint64_t* rows, uint32_t h, const uint32_t* p_words,
                            uint32_t traits, const uint4* packed, int64_t* out, hipStream_t s) {
    const uint32_t nt = i8_ntiles(traits);
    if (nt <= 1) {
        const uint32_t blocks = (h + 4 * 64 - 1) / (4 * 64);
        hipLaunchKernelGGL((k_traits_i8<4, 1>), dim3(blocks), dim3(256), 0, s, rows, h, p_words,
                           traits, packed, out);
    } else if (nt <= 2) {
        const uint32_t blocks = (h + 4 * 32 - 1) / (4 * 32);
        hipLaunchKernelGGL((k_traits_i8<2, 2>), dim3(blocks), dim3(256), 0, s, rows, h, p_words,
                           traits, packed, out);
    } else if (nt <= 4) {  // four row tiles per wave, as the pipeline's 16-trait kernel
        const uint32_t blocks = (h + 4 * 64 - 1) / (4 * 64);
        hipLaunchKernelGGL((k_traits_i8<4, 4>), dim3(blocks), dim3(256), 0, s, rows, h, p_words,
                           traits, packed, out);
    } else {
        const uint32_t blocks = (h + 4 * 16 - 1) / (4 * 16);
        hipLaunchKernelGGL((k_traits_i8<1, 8>), dim3(blocks), dim3(256), 0, s, rows, h, p_words,
                           traits, packed, out);
    }
    return hipGetLastError();
}

// ---- split-BF16 path (standalone contraction) -----------------------------
// The packed effects cover whole 16-word groups of the padded row width
// (sites >= `sites` are 0): packed[((G * 32 + q) * terms + term) * 64 + lane]
// = 8 BF16 of column lane % 16 (trait) at sites bf_site(G, q, lane / 16, i).
size_t mfma_effect_elems(uint32_t words, int terms) {
    const size_t ng = (words + 15) / 16;
    return ng * 32 * (size_t)terms * 64 * 8;  // bf16 elements
}

void mfma_pack_effects(const double* effects, uint32_t sites, uint32_t traits, uint32_t words,
                       int terms, uint16_t* packed) {
    const uint32_t ng = (words + 15) / 16;
    for (uint32_t G = 0; G < ng; ++G)
        for (uint32_t q = 0; q < 32; ++q)
            for (uint32_t lane = 0; lane < 64; ++lane) {
                const uint32_t n = lane & 15, g = lane >> 4;
                for (uint32_t i = 0; i < 8; ++i) {
                    const size_t site = bf_site(G, q, g, i);
                    double x = (site < sites && n < traits) ? effects[site * traits + n] : 0.0;
                    for (int tm = 0; tm < terms; ++tm) {
                        const uint16_t hb = bf16_rne((float)x);
                        packed[((((size_t)G * 32 + q) * terms + tm) * 64 + lane) * 8 + i] = hb;
                        x -= (double)bf16_to_f(hb);
                    }
                }
            }
}

size_t mfma_partial_elems(uint32_t h, uint32_t w, uint32_t traits) {
    const size_t chunks = ((w + 15) / 16 + kChunkGroups - 1) / kChunkGroups;
    return chunks * h * traits;
}

hipError_t launch_traits_mfma(const uint64_t* rows, uint32_t h, uint32_t w,
                              const uint16_t* packed, uint32_t sites, uint32_t traits, int terms,
                              double* partial, double* z, hipStream_t s) {
    (void)sites;
    const uint32_t chunks = ((w + 15) / 16 + kChunkGroups - 1) / kChunkGroups;
    const dim3 grid((h + 4 * 64 - 1) / (4 * 64), chunks);
    const uint4* pk = reinterpret_cast<const uint4*>(packed);
    if (terms == 1)
        hipLaunchKernelGGL((k_traits_bf16<4, 1>), grid, dim3(256), 0, s, rows, h, w, traits, pk,
                           partial);
    else if (terms == 2)
        hipLaunchKernelGGL((k_traits_bf16<4, 2>), grid, dim3(256), 0, s, rows, h, w, traits, pk,
                           partial);
    else
        hipLaunchKernelGGL((k_traits_bf16<4, 3>), grid, dim3(256), 0, s, rows, h, w, traits, pk,
                           partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t n = (size_t)h * traits;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s,
                       partial, chunks, n, z);
    return hipGetLastError();
}

}  // namespace clotho
