// pipeline.hip — gfx950 kernels of one generation of the offspring pipeline.
//
// Generation = draw -> recombine+mutate (+fused free-space partials) ->
// exact traits -> free-space reduce/compact -> fitness -> commit, all on the
// context stream, all shapes read from GenState so a captured hipGraph
// replays generation after generation.
//
// Reference path restated (DS/ = include/clotho/data_spaces/,
// QS/ = examples/QTLSim/):
//   QS/engine_parallel_pipeline.hpp:172-239 (simulate), :327-367 (mutations),
//   :393-447 (child_max_alleles, removeFixedAlleles);
//   DS/offspring_generator/offspring_generator_population_space_row_alignment.hpp:84-260;
//   DS/crossover/block_crossover.hpp:46-63; DS/free_space/free_space_details.hpp:359-408;
//   DS/fitness/general_fitness.hpp:148-177.
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "pipeline.hpp"
#include "traits.hpp"
#include "traits_i8.hpp"

namespace clotho {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// Buffers reach the batched kernels through device memory, so their pointers
// carry no address space and the compiler emits FLAT accesses (which also
// count against the LDS wait counter and stall every LDS wait behind the
// parent-row prefetches).  Every kernel body therefore views its buffers
// as global-address-space pointers.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gbl(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

// A block-uniform word of device memory, kept in a scalar register.  In the
// persistent batched kernels the compiler cannot prove that GenState is not
// clobbered by earlier stores of the same launch and reads it with vector
// loads; making the value explicitly uniform keeps it out of VGPRs.
template <class Ptr>
__device__ __forceinline__ uint32_t uniform_read(Ptr p) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)*p);
}

// Scan blocks over the parent population's fitness (selection set-up).
__host__ __device__ inline uint32_t parent_blocks(const Params& P) {
    return (P.n_par + kBlock - 1) / kBlock;
}


template <class Ptr>
__device__ inline uint32_t poisson_draw(Ptr cdf, int n, double u) {
    uint32_t k = 0;
    while ((int)k + 1 < n && !(u < cdf[k])) ++k;
    return k;
}

// Inclusive scan over a 256-thread block (four 64-lane waves).
template <typename T>
__device__ inline T block_inclusive_scan(T v, T* lds, T& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    T x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const T y = __shfl_up(x, d, 64);
        if (lane >= d) x = x + y;
    }
    if (lane == 63) lds[wave] = x;
    __syncthreads();
    T off = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
        if (q < wave) off = off + lds[q];
        tot = tot + lds[q];
    }
    __syncthreads();
    total = tot;
    return off + x;
}

// Scan of nb <= kTotalsLds block totals into LDS by one 256-thread block
// (eight totals per thread): dst[i] = src[0] + ... + src[i] (kInclusive) or
// src[0] + ... + src[i - 1]; returns the sum of all.  Every value is an
// integer below 2^53, so the order of the additions does not matter.
constexpr uint32_t kTotalsLds = 2048;
template <typename T, bool kInclusive, class Src>
__device__ inline T block_scan_totals(Src src, uint32_t nb, T* dst, T* lds) {
    const uint32_t t0 = threadIdx.x * 8u;
    T v[8], sum = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        v[j] = t0 + j < nb ? src[t0 + j] : T(0);
        sum = sum + v[j];
    }
    T total;
    T run = block_inclusive_scan<T>(sum, lds, total) - sum;
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        if (kInclusive) run = run + v[j];
        if (t0 + j < nb) dst[t0 + j] = run;
        if (!kInclusive) run = run + v[j];
    }
    __syncthreads();
    return total;
}

// ---------------------------------------------------------------------------
// Draw phase
// ---------------------------------------------------------------------------

// Per-gamete mutation counts, Poisson(mu) by inversion (the reference's
// Poisson(mu * 2N) total with uniform rows has the same law), and their
// block-local exclusive scan.
struct MutCounts {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.mut_blocks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto cdf_mu = gbl(B.cdf_mu);
        const auto count = gbl(B.mut_count);
        const auto local = gbl(B.mut_local);
        const auto bsum = gbl(B.mut_bsum);
        __shared__ uint32_t lds[kWaves];
        const uint32_t o = blk * kBlock + threadIdx.x;
        const uint32_t gen = st->gen;
        uint32_t c = 0;
        if (o < P.h) {
            const Draw4 d = philox_draw(P.seed, gen, S_MUTCOUNT, P.replicate, o, 0);
            c = poisson_draw(cdf_mu, P.n_cdf_mu, u53(d.x[0], d.x[1]));
            count[o] = c;
        }
        uint32_t total;
        const uint32_t inc = block_inclusive_scan<uint32_t>(c, lds, total);
        if (o < P.h) local[o] = inc - c;
        if (threadIdx.x == 0) bsum[blk] = total;
    }
};

// Per-block max and all-equal flag of the parent fitness (selection set-up);
// also clears the per-parent child counts of the parent order.
struct FitnessStats {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return parent_blocks(P); }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto fit = gbl(B.fitness[sel]);
        const auto stats = gbl(B.fit_stats);
        __shared__ double smax[kBlock];
        __shared__ int seq[kBlock];
        const uint32_t i = blk * kBlock + threadIdx.x;
        const double f0 = fit[0];
        double m = 0.0;
        int eq = 1;
        if (i < P.n_par) {
            const double f = fit[i];
            m = f;
            eq = (f == f0) ? 1 : 0;
            if (B.by_parent) gbl(B.parent_cnt)[i] = 0;  // counted by the gamete draws
        }
        smax[threadIdx.x] = m;
        seq[threadIdx.x] = eq;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
                seq[threadIdx.x] = seq[threadIdx.x] & seq[threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            stats[2 * blk] = smax[0];
            stats[2 * blk + 1] = seq[0] ? 1.0 : 0.0;
        }
    }
};

// Single block: mutation offsets, child allele count / width (child_max_alleles),
// capacity check, selection quantisation exponent.
struct Prepare {
    static constexpr int kThreads = 1024;
    __host__ __device__ static uint32_t blocks(const Params& P) { return 1u; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto mut_bsum = gbl(B.mut_bsum);
        const auto stats = gbl(B.fit_stats);
        (void)blk;
        __shared__ uint32_t s_carry;
        __shared__ uint32_t s_wave[16];
        __shared__ double s_max[1024];
        __shared__ int s_eq[1024];
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        if (tid == 0) s_carry = 0;
        __syncthreads();
        // exclusive scan of the per-block mutation totals, in place
        for (uint32_t base = 0; base < P.mut_blocks; base += 1024) {
            const uint32_t idx = base + tid;
            const uint32_t v = idx < P.mut_blocks ? mut_bsum[idx] : 0;
            uint32_t x = v;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if (lane >= d) x += y;
            }
            if (lane == 63) s_wave[wave] = x;
            __syncthreads();
            uint32_t off = s_carry, tot = 0;
            for (int q = 0; q < 16; ++q) {
                if (q < wave) off += s_wave[q];
                tot += s_wave[q];
            }
            if (idx < P.mut_blocks) mut_bsum[idx] = off + x - v;
            __syncthreads();
            if (tid == 0) s_carry += tot;
            __syncthreads();
        }
        // selection: constant fitness -> uniform; else quantisation exponent
        double m = 0.0;
        int eq = 1;
        for (uint32_t b = tid; b < parent_blocks(P); b += 1024) {
            m = fmax(m, stats[2 * b]);
            eq &= stats[2 * b + 1] != 0.0 ? 1 : 0;
        }
        s_max[tid] = m;
        s_eq[tid] = eq;
        __syncthreads();
        for (int s = 512; s > 0; s >>= 1) {
            if (tid < s) {
                s_max[tid] = fmax(s_max[tid], s_max[tid + s]);
                s_eq[tid] &= s_eq[tid + s];
            }
            __syncthreads();
        }
        if (tid == 0) {
            const uint32_t M = s_carry;
            const uint32_t F = st->free_count;
            const uint32_t all_size = st->alleles + (M > F ? M - F : 0);
            st->m = M;
            uint32_t as = all_size;
            if (all_size > P.s_cap) {
                if (!st->error) {
                    st->error = 1;
                    st->error_gen = st->gen;
                }
                as = P.s_cap;
            }
            st->all_size = as;
            st->wc = ((as + 511) / 512) * 8;
            const double mx = s_max[0];
            int constant = s_eq[0] || !(mx > 0.0);
            st->fit_constant = constant ? 1u : 0u;
            if (!constant) {
                int L = 0;
                while (((uint64_t)1 << L) < P.n_par) ++L;
                st->sel_shift = 52 - L - (det_ilogb(mx) + 1);
            }
        }
    }
};

// New-allele attributes for every mutation of gamete o: slot = free[k] (fixed
// ascending, then lost ascending) else appended (generate_child_mutations);
// position uniform (AlleleGenerator), age = generation, per-trait effect
// neutral ? 0 : quantised Normal(mean, sigma) (TraitSpaceGenerator2).
struct MutAttrs {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.mut_blocks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto count = gbl(B.mut_count);
        const auto local = gbl(B.mut_local);
        const auto boff = gbl(B.mut_bsum);
        const auto free_list = gbl(B.free_list);
        const auto keys = gbl(B.keys);
        const auto effects = gbl(B.effects);
        const auto ages = gbl(B.ages);
        const uint32_t o = blk * kBlock + threadIdx.x;
        if (o >= P.h) return;
        const uint32_t c = count[o];
        if (!c) return;
        const uint32_t gen = st->gen, F = st->free_count, A = st->alleles, fb = st->free_base;
        const uint32_t k0 = local[o] + boff[o / kBlock];
        for (uint32_t k = k0; k < k0 + c; ++k) {
            const uint32_t slot = k < F ? free_list[fb + k] : A + (k - F);
            if (slot >= P.s_cap) continue;
            const Draw4 d = philox_draw(P.seed, gen, S_MUTATTR, P.replicate, k, 0);
            keys[slot] = d.x[0];
            ages[slot] = gen;
            for (uint32_t t = 0; t < P.traits; ++t) {
                const Draw4 e = philox_draw(P.seed, gen, S_MUTATTR, P.replicate, k, 1 + t);
                int32_t q = 0;
                if (!(u53(e.x[0], e.x[1]) < P.neutral_p)) {
                    const double w = P.trait_mean + P.trait_sigma * normal_ppf(uopen(e.x[2], e.x[3]));
                    double n = rint(w * P.eff_scale);
                    if (n > kEffectLimit) n = kEffectLimit;
                    if (n < -kEffectLimit) n = -kEffectLimit;
                    q = (int32_t)n;
                }
                effects[(size_t)slot * P.traits + t] = q;
            }
        }
    }
};

// Quantised fitness q_i = floor(f_i 2^shift) and its block-local inclusive
// scan; every partial sum is an integer below 2^52, exact in FP64.
struct SelScan {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return parent_blocks(P); }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto fit = gbl(B.fitness[sel]);
        const auto cum_local = gbl(B.cum_local);
        const auto cum_block = gbl(B.cum_block);
        __shared__ double lds[kWaves];
        if (st->fit_constant) return;
        const uint32_t i = blk * kBlock + threadIdx.x;
        const int sh = st->sel_shift;
        double q = 0.0;
        if (i < P.n_par) {
            const double f = fit[i];
            q = floor(det_ldexp(f > 0.0 ? f : 0.0, sh));
        }
        double total;
        const double inc = block_inclusive_scan<double>(q, lds, total);
        if (i < P.n_par) cum_local[i] = inc;
        if (threadIdx.x == 0) cum_block[blk] = total;
    }
};

struct SelBlockscan {
    static constexpr int kThreads = 1024;
    __host__ __device__ static uint32_t blocks(const Params& P) { return 1u; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto cum_block = gbl(B.cum_block);
        (void)blk;
        __shared__ double s_wave[16];
        __shared__ double s_carry;
        if (st->fit_constant) return;
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const uint32_t nb = parent_blocks(P);
        if (tid == 0) s_carry = 0.0;
        __syncthreads();
        for (uint32_t base = 0; base < nb; base += 1024) {
            const uint32_t idx = base + tid;
            double x = idx < nb ? cum_block[idx] : 0.0;
            for (int d = 1; d < 64; d <<= 1) {
                const double y = __shfl_up(x, d, 64);
                if (lane >= d) x = x + y;
            }
            if (lane == 63) s_wave[wave] = x;
            __syncthreads();
            double off = s_carry, tot = 0.0;
            for (int q = 0; q < 16; ++q) {
                if (q < wave) off = off + s_wave[q];
                tot = tot + s_wave[q];
            }
            if (idx < nb) cum_block[idx] = off + x;
            __syncthreads();
            if (tid == 0) s_carry = s_carry + tot;
            __syncthreads();
        }
        if (tid == 0) st->sel_total = (uint64_t)s_carry;
    }
};

// Parent pick for one uniform: uniform over [0, N) when fitness is constant,
// else min{i : target < C[i]} with target = floor(k * C[N-1] / 2^53).
template <class LPtr, class BPtr>
__device__ inline uint32_t pick_parent(const Params& P, bool constant, uint64_t total, uint64_t k,
                                       LPtr cum_local, BPtr cum_block) {
    if (constant) return (uint32_t)__umul64hi(k << 11, (uint64_t)P.n_par);
    const double target = (double)__umul64hi(k << 11, total);
    // first block whose inclusive total exceeds target
    uint32_t lo = 0, hi = parent_blocks(P);
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (cum_block[mid] > target)
            hi = mid;
        else
            lo = mid + 1;
    }
    const uint32_t b = lo;
    const double base = b ? cum_block[b - 1] : 0.0;
    const double rel = target - base;
    uint32_t a = b * kBlock, e = min((b + 1) * kBlock, P.n_par);
    while (a < e) {
        const uint32_t mid = (a + e) >> 1;
        if (cum_local[mid] > rel)
            e = mid;
        else
            a = mid + 1;
    }
    return a;
}

// Per offspring gamete o = 2i + s: parent (id0, id1 of offspring i), event
// count E ~ Poisson(rho), strand swap ~ Bernoulli(bias), E event keys
// (common_crossover::operator(), fill_events), and the top/bottom parent rows.
// Selection block totals (kTotals), held in LDS so that the block search is a
// chain of LDS reads, not of round trips: kTotalsScan has each block scan the
// raw totals itself (up to kTotalsLds of them), in place of a separate
// SelBlockscan launch; kTotalsCopy copies totals a SelBlockscan launch
// already scanned (the batched launches, up to 1024 of them).  Beyond those
// counts the search reads the scanned totals in global memory.  The
// event-count table is read from LDS too: the inversion walks it with one
// dependent read per step (up to ~10 steps at rho = 5).
constexpr int kTotalsScan = 1, kTotalsCopy = 2;
template <int kTotals>
struct GametesOp {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.mut_blocks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto cdf_rho = gbl(B.cdf_rho);
        const auto cum_local = gbl(B.cum_local);
        const auto cum_block = gbl(B.cum_block);
        const auto top_out = gbl(B.top);
        const auto bot_out = gbl(B.bottom);
        const auto nev_out = gbl(B.n_events);
        const auto events = gbl(B.events);
        // copied totals up to 1024 (8 KB: eight 256-thread blocks per CU
        // still fit), scanned ones up to kTotalsLds; beyond, the global search
        constexpr uint32_t kCap = kTotals == kTotalsScan ? kTotalsLds : 1024u;
        __shared__ double s_cb[kCap];
        __shared__ double s_lds[kWaves];
        __shared__ double s_cdf[kMaxCdf];
        const bool constant = uniform_read(&st->fit_constant) != 0;
        const uint32_t nb = parent_blocks(P);
        const bool staged = nb <= kCap && !constant;
        for (uint32_t i = threadIdx.x; i < P.n_cdf_rho; i += kThreads) s_cdf[i] = cdf_rho[i];
        if (kTotals == kTotalsCopy && staged)
            for (uint32_t i = threadIdx.x; i < nb; i += kThreads) s_cb[i] = cum_block[i];
        __syncthreads();
        double scanned_total = 0.0;
        if constexpr (kTotals == kTotalsScan) {
            if (staged) scanned_total = block_scan_totals<double, true>(cum_block, nb, s_cb, s_lds);
        }
        const uint32_t o = blk * kBlock + threadIdx.x;
        if (o >= P.h) return;
        const uint32_t gen = st->gen;
        const uint32_t i = o >> 1;
        const Draw4 pd = philox_draw(P.seed, gen, S_PARENTS, P.replicate, i, 0);
        const uint64_t k = (o & 1) ? k53(pd.x[2], pd.x[3]) : k53(pd.x[0], pd.x[1]);
        // given parents (one generation): Crossover::update's mate pairs
        const bool given = uniform_read(&st->given) != 0;
        const uint64_t total = kTotals == kTotalsScan ? (uint64_t)scanned_total : st->sel_total;
        const uint32_t p =
            given ? min(gbl(B.given_parents)[o], P.n_par - 1)
            : staged ? pick_parent(P, false, total, k, cum_local, (const double*)s_cb)
                     : pick_parent(P, constant, st->sel_total, k, cum_local, cum_block);
        const Draw4 g = philox_draw(P.seed, gen, S_GAMETE, P.replicate, o, 0);
        const uint32_t ne = poisson_draw((const double*)s_cdf, P.n_cdf_rho, u53(g.x[0], g.x[1]));
        const bool swap = u53(g.x[2], g.x[3]) < P.bias;
        top_out[o] = swap ? 2 * p + 1 : 2 * p;
        bot_out[o] = swap ? 2 * p : 2 * p + 1;
        nev_out[o] = ne;
        // the gamete's rank among its parent's picks (atomic order) places its
        // record in the parent order without a second atomic
        uint32_t rank = 0;
        const bool ranked = B.by_parent && p < P.n_par;
        if (ranked)
            rank = __hip_atomic_fetch_add(gbl(B.parent_cnt) + p, 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
        // one 16-byte store per draw of four keys: the keys past ne in the
        // last group are never read (every reader stops at the event count),
        // and e + 3 < ev_cap since e < ne <= ev_cap and both e and ev_cap
        // are multiples of 4 (so is the row offset: the stores are aligned)
        const auto ev = events + (size_t)o * P.ev_cap;
        for (uint32_t e = 0; e < ne; e += 4) {
            const Draw4 d = philox_draw(P.seed, gen, S_GAMETE, P.replicate, o, 1 + e / 4);
            *(__attribute__((address_space(1))) i8::u32x4*)(ev + e) =
                i8::u32x4{d.x[0], d.x[1], d.x[2], d.x[3]};
        }
        if (ranked) gbl(B.prank)[o] = rank;
    }
};

using Gametes = GametesOp<kTotalsCopy>;  // the batched launches (SelBlockscan scans)

// Parent order of the child rows: a counting sort of the gametes by parent
// individual (counts from the gamete draws, block-local exclusive scan, scan
// of the block totals, scatter).  Only the processing order of the
// recombination depends on it; which of two siblings comes first is decided by
// the scatter's atomics and changes nothing in the output.
struct ParentScan {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return parent_blocks(P); }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto cnt = gbl(B.parent_cnt);
        __shared__ uint32_t lds[kWaves];
        const uint32_t i = blk * kBlock + threadIdx.x;
        const uint32_t c = i < P.n_par ? cnt[i] : 0u;
        uint32_t total;
        const uint32_t inc = block_inclusive_scan<uint32_t>(c, lds, total);
        if (i < P.n_par) cnt[i] = inc - c;
        if (threadIdx.x == 0) gbl(B.parent_bsum)[blk] = total;
    }
};

struct ParentBlockscan {
    static constexpr int kThreads = 1024;
    __host__ __device__ static uint32_t blocks(const Params& P) { return 1u; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        (void)blk;
        const auto bsum = gbl(B.parent_bsum);
        __shared__ uint32_t s_wave[16];
        __shared__ uint32_t s_carry;
        const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const uint32_t nb = parent_blocks(P);
        if (tid == 0) s_carry = 0;
        __syncthreads();
        for (uint32_t base = 0; base < nb; base += 1024) {
            const uint32_t idx = base + tid;
            const uint32_t v = idx < nb ? bsum[idx] : 0u;
            uint32_t x = v;
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if ((int)lane >= d) x += y;
            }
            if (lane == 63) s_wave[wave] = x;
            __syncthreads();
            uint32_t off = s_carry, tot = 0;
            for (uint32_t q = 0; q < 16; ++q) {
                if (q < wave) off += s_wave[q];
                tot += s_wave[q];
            }
            if (idx < nb) bsum[idx] = off + x - v;
            __syncthreads();
            if (tid == 0) s_carry += tot;
            __syncthreads();
        }
    }
};

// kScanTotals: each block scans the per-block parent counts into LDS itself
// (up to kTotalsLds of them), in place of a separate ParentBlockscan launch.
template <bool kScanTotals>
struct ParentScatterOp {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.mut_blocks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        __shared__ uint32_t s_off[kScanTotals ? kTotalsLds : 1];
        __shared__ uint32_t s_lds[kWaves];
        bool staged = false;
        if constexpr (kScanTotals) {
            const uint32_t nb = parent_blocks(P);
            staged = nb <= kTotalsLds;
            if (staged) (void)block_scan_totals<uint32_t, false>(gbl(B.parent_bsum), nb, s_off, s_lds);
        }
        const uint32_t o = blk * kBlock + threadIdx.x;
        if (o >= P.h) return;
        const uint32_t top = gbl(B.top)[o], p = top >> 1;
        if (p >= P.n_par) return;
        const uint32_t pos = (staged ? s_off[p / kBlock] : gbl(B.parent_bsum)[p / kBlock]) +
                             gbl(B.parent_cnt)[p] + gbl(B.prank)[o];
        const uint32_t k0 = gbl(B.mut_local)[o] + gbl(B.mut_bsum)[o / kBlock];
        const uint32_t nx = gbl(B.n_events)[o] | (gbl(B.mut_count)[o] << 8);
        if (pos < P.h)
            *(__attribute__((address_space(1))) i8::u32x4*)(B.by_parent + pos) =
                i8::u32x4{o, top, nx, k0};
    }
};

// ---------------------------------------------------------------------------
// Recombination + mutation (+ free-space partials)
// ---------------------------------------------------------------------------
// Per-word order statistics, once per generation: the 64 position keys of
// word w sorted ascending (sorted[w][k]), prefix masks pm[w][c] = bits of the
// c smallest keys, and a 256-bucket table bucket[w][b] = #{k : key_k < b<<24}
// (four buckets per dword), each stored tile-major (table_at, prefix_at).  For a gamete event e, c(e) = #{k : key_k < e}
// is bucket[e>>24] plus the few keys of that bucket below e, and
// {j in w : key_j >= e} = ~pm[w][c(e)]; the crossover mask of the word,
// {j : odd #{events e <= key_j}} (block_crossover + region_classifier's
// upper_bound), is the XOR over events of ~pm[w][c(e)].
struct WordSort {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return (P.w_cap + kWaves - 1) / kWaves; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto keys = gbl(B.keys);
        const auto sorted = gbl(B.sorted);
        const auto pm = gbl(B.pm);
        const auto bucket = gbl(B.bucket);
        __shared__ uint32_t s_key[kWaves][64];
        __shared__ uint32_t s_perm[kWaves][64];
        const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const uint32_t w = blk * kWaves + wave;
        const uint32_t wc = st->wc, A = st->alleles;
        if (w >= wc) return;  // uniform per wave; no block barrier follows
        const uint32_t slot = w * 64 + lane;
        const uint32_t key = slot < A ? keys[slot] : 0u;
        s_key[wave][lane] = key;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        uint32_t rank = 0;
        for (uint32_t k = 0; k < 64; ++k) {
            const uint32_t v = s_key[wave][k];
            rank += (v < key || (v == key && k < lane)) ? 1u : 0u;
        }
        s_perm[wave][rank] = lane;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        const uint32_t j = s_perm[wave][lane];  // slot bit of the lane-th smallest key
        sorted[table_at(w, lane)] = s_key[wave][j];
        uint64_t x = (uint64_t)1 << j;
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t y = __shfl_up(x, d, 64);
            if ((int)lane >= d) x |= y;
        }
        pm[prefix_at(w, lane + 1)] = x;
        if (lane == 0) pm[prefix_at(w, 0)] = 0;
        // buckets 4 lane .. 4 lane + 3: count of keys below b << 24
        uint32_t packed = 0;
        for (uint32_t q = 0; q < 4; ++q) {
            const uint32_t bound = (4 * lane + q) << 24;
            uint32_t cnt = 0;
            for (uint32_t k = 0; k < 64; ++k) cnt += s_key[wave][k] < bound ? 1u : 0u;
            packed |= cnt << (8 * q);  // cnt <= 64 fits a byte
        }
        bucket[table_at(w, lane)] = packed;
    }
};

// Grid: x = 64-word column tiles, y = row chunks; kRecWaves waves per block,
// one 64-bit word per lane.  The tile's bucket table ([b/4][lane]), sorted
// keys ([k][lane], 4 sentinel rows) and prefix masks ([c][lane]) sit in LDS,
// each lane reading its own bank column.  For each child row:
//   t, b  = top/bottom parent words with fixed columns cleared
//   mask  = XOR_e ~pm[c(e)]
//   child = (t & ~mask) | (b & mask), then XOR of this row's new mutations
// OR/AND of the child words accumulate in registers (free_space_accumulator).
//
// Latency structure: with parent-ordered gamete records a wave owns a
// contiguous segment of positions (without them, positions c0 + wave +
// kRecWaves j).  The records of 64 positions (child row, top row, event and
// mutation counts, mutation offset) are loaded with one vector load, lane q
// holding position j0 + q, and read back with v_readlane, so the row loop
// issues no scalar loads.  Runs of positions with one parent individual are
// cut into items of at most kSibs children: an item loads the parent pair
// once and the event keys of each child (lane e holds key e).  Items
// alternate between two register sets: the loads of item k + 2 are issued as
// soon as item k is done and stay in flight while item k + 1 is classified
// (no register rotation, hence no wait on loads still in flight; three or
// four sets measured no faster on the same device).  Events are classified
// kEventGroup at a time so that their LDS lookups overlap; a group is padded
// with rank-64 entries (pm[64] is all ones), and since #real + #padding =
// kEventGroup is even, the XOR of the group's prefix masks equals the XOR of
// the complements of the real events.
constexpr int kRecWaves = 16;
constexpr int kRecBlock = 64 * kRecWaves;
// events classified together; any even size keeps the padding parity
// argument (#real + #padding even), and 2 pads Poisson(5) counts least
constexpr int kEventGroup = 2;
// children of one parent individual produced from one load of its row pair
// (parent-ordered gametes; the batched launches keep child order and take 1)
constexpr int kSibsOrdered = 3;
// dword 3 of a raw (stride 0) buffer resource on gfx9-family parts: 32-bit
// data format, offsets range-checked against num_records in bytes
constexpr int kBufferWord3 = 0x00020000;
// cache-policy operand of the buffer intrinsics: NT (non-temporal, streaming)
constexpr int kCachePolicyNT = 2;

__device__ __forceinline__ uint32_t lane_read(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// A block's LDS copy of column tile bx's order-statistic tables (the
// tile-major images of table_at / prefix_at): every 16-byte load is issued
// before the first LDS write, so the fill costs one round trip, not one per
// table row.  Columns at or past wc read as no keys: sorted all ones, bucket
// and prefix masks zero.  fw < 64 (narrow rows, tile 0; a power of two of at
// least 4): column l holds word l mod fw, the tile's first fw words repeated
// across the 64 columns.
__device__ __forceinline__ void load_word_tables(const Buffers& B, uint32_t bx, uint32_t wc,
                                                 uint32_t fw, uint32_t* s_bucket,
                                                 uint32_t* s_sorted, uint64_t* s_pm) {
    static_assert(kRecBlock == 1024, "one 16-byte table piece per thread");
    using i8::u32x4;
    using i8::g_u32x4;
    uint32_t t = threadIdx.x;
    // opaque here, so the per-lane table offsets are computed per call: hoisted
    // out of a persistent block's loop they outlive it and spill to scratch
    asm volatile("" : "+v"(t));
    const uint32_t w0 = bx * 64;
    // sorted and bucket: 64 rows x 64 columns of dwords, four columns per thread
    const uint32_t k = t >> 4, c = (t & 15u) * 4u, sc = c & (fw - 1);
    const u32x4 vs = *(const g_u32x4*)(gbl(B.sorted) + table_at(w0 + sc, k));
    const u32x4 vb = *(const g_u32x4*)(gbl(B.bucket) + table_at(w0 + sc, k));
    // prefix masks: 65 rows x 64 columns of qwords, two columns per piece;
    // pieces t, t + 1024 and (first 32 threads) t + 2048 of 2080
    u32x4 vp[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const uint32_t x = min(t + 1024u * i, 65u * 32u - 1u);  // in bounds; the store is guarded
        vp[i] = *(const g_u32x4*)(gbl(B.pm) + prefix_at(w0 + (((x & 31u) * 2u) & (fw - 1)), x >> 5));
    }
    u32x4 ms, mb;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const bool ok = w0 + sc + e < wc;
        ms[e] = ok ? vs[e] : 0xFFFFFFFFu;
        mb[e] = ok ? vb[e] : 0u;
    }
    *(u32x4*)(s_sorted + k * 64 + c) = ms;
    *(u32x4*)(s_bucket + k * 64 + c) = mb;
    if (t < 64) *(u32x4*)(s_sorted + 64 * 64 + t * 4) = u32x4{~0u, ~0u, ~0u, ~0u};  // rows 64..67
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        if (i == 2 && t >= 32) break;
        const uint32_t x = t + 1024u * i, col = (x & 31u) * 2u;
        const uint32_t sw = w0 + (col & (fw - 1));
        u32x4 v = vp[i];
        if (sw >= wc) v.x = v.y = 0u;
        if (sw + 1 >= wc) v.z = v.w = 0u;
        *(u32x4*)(s_pm + (x >> 5) * 64 + col) = v;
    }
}

// Memory operations of the row loop are issued unconditionally (indices
// clamped in bounds, the child store through a range-checked buffer resource
// that drops the lanes past the row), so the number of vector memory
// operations in flight is the same on every path and the waits before a row's
// compute cover only that row's loads.
// Narrow rows (W <= 32 words, child order, at most 64 events per gamete):
// a wave produces 64 / FW rows per step, lane = (row f, word lane mod FW), so
// a replicate of 8-word rows keeps every lane busy instead of one in eight.
// The LDS tables hold the tile's FW words repeated across the 64 columns, so
// lane l still reads its own bank column.  For events j FW .. j FW + FW - 1
// each lane holds key j FW + (l mod FW) of its row (the first 2 FW keys are
// loaded with the parent words, later ones only when some row of the step
// has that many events) and takes key x of its row from lane
// f FW + x mod FW; rows of a step share the event loop up to the largest
// event count, lookups past a row's own count take rank 64 (pm[64] = all
// ones), and since every lane does the same even number of lookups the XOR
// of the prefix masks is still the XOR of the complements of the real events.
// One code path serves FW = 8, 16 and 32 (lg = log2 FW): a persistent
// batched block holds every path of the kernel in one register allocation of
// at most 64 registers, and one narrow path keeps it free of spills.
__device__ __forceinline__ void recombine_folded(const Params& P, const Buffers& B, uint32_t sel,
                                                 uint32_t by, uint32_t lg, uint32_t* s_bucket,
                                                 uint32_t* s_sorted, uint64_t* s_pm) {
    const uint32_t FW = 1u << lg;     // words per row slot
    const uint32_t NF = 64u >> lg;    // rows per wave step
    // event-key registers loaded with the parent words: keys 0 .. 2 FW - 1
    // (more than 8 crossover events per gamete is common at rho = 5)
    constexpr uint32_t NK = 2;
    const auto st = gbl(B.state);
    const auto parent = gbl(B.rows[sel]);
    const auto child = gbl(B.rows[1 - sel]);
    const auto free_list = gbl(B.free_list);
    const auto fixed_bits = gbl(B.fixed_bits);
    const auto top_in = gbl(B.top);
    const auto nev_in = gbl(B.n_events);
    const auto events = gbl(B.events);
    const auto count = gbl(B.mut_count);
    const auto local = gbl(B.mut_local);
    const auto boff = gbl(B.mut_bsum);
    const uint32_t wc = uniform_read(&st->wc), wp = uniform_read(&st->wp);
    const uint32_t A = uniform_read(&st->alleles), F = uniform_read(&st->free_count),
                   S = uniform_read(&st->all_size), fb = uniform_read(&st->free_base);
    const uint32_t ev_cap = P.ev_cap;  // <= 64 on this path
    // opaque thread index: in a persistent block's loop the lane-derived
    // values of each row width would otherwise be hoisted and held throughout
    uint32_t tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const uint32_t lane = tid & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t f = lane >> lg, w = lane & (FW - 1);
    const bool valid = w < wc;
    const bool inpar = valid && w < wp;
    const uint32_t col = wp ? min(w, wp - 1) : 0;

    load_word_tables(B, 0, wc, FW, s_bucket, s_sorted, s_pm);
    const uint64_t keep = inpar ? ~fixed_bits[w] : 0;
    __syncthreads();

    const uint32_t rows_per_chunk = (P.h + P.chunks - 1) / P.chunks;
    const uint32_t c0 = min(P.h, by * rows_per_chunk);
    const uint32_t c1 = min(P.h, c0 + rows_per_chunk);
    uint64_t acc_or = 0, acc_and = ~(uint64_t)0;

    // Each wave owns a contiguous segment of the chunk's rows (a multiple of
    // NF rows).  The gamete records of 64 rows (event and mutation counts,
    // top row) are loaded with one load per field, lane q holding row q of
    // the batch; a step of NF rows reads its rows' records from those registers
    // with shuffles, so only the parent words and event keys of a step are
    // loads of their own.  Steps alternate between two register sets: the
    // loads of step k + 2 are issued as soon as step k is done and stay in
    // flight while step k + 1 is classified.
    const uint32_t seg = ((c1 - c0 + kRecWaves - 1) / kRecWaves + NF - 1) & ~(NF - 1);
    const uint32_t s0 = min(c1, c0 + wave * seg);
    const uint32_t nrow = min(c1, s0 + seg) - s0;  // rows of this wave
    struct Par {
        uint64_t t, b;
        uint32_t kv[NK];  // event keys x FW + w of the lane's row
    };

    for (uint32_t j0 = 0; j0 < nrow; j0 += 64) {
        const uint32_t nb = min(64u, nrow - j0);  // rows of this record batch
        const uint32_t nst = (nb + NF - 1) >> (6 - lg);  // its steps
        const uint32_t rr = s0 + j0 + min(lane, nb - 1);  // in bounds
        const uint32_t g_nev = nev_in[rr], g_cnt = count[rr], g_top = top_in[rr];
        // record lane of the lane's row in step k (clamped into the batch)
        auto rec_lane = [&](uint32_t k) { return min(k * NF + f, nb - 1); };
        auto load_step = [&](uint32_t k, Par& pr) {
            const uint32_t q = rec_lane(min(k, nst - 1));
            const uint32_t top = (uint32_t)__shfl((int)g_top, (int)q, 64);
            // opaque column indices: the addresses are formed here, per call,
            // rather than hoisted and held as 64-bit pointers across the loop
            uint32_t cl = col, wl = w;
            asm volatile("" : "+v"(cl), "+v"(wl));
            pr.t = parent[(size_t)top * wp + cl];
            pr.b = parent[(size_t)(top ^ 1u) * wp + cl];
            const auto ev = events + (size_t)(s0 + j0 + q) * ev_cap;
#pragma unroll
            for (uint32_t x = 0; x < NK; ++x) pr.kv[x] = ev[min(x * FW + wl, ev_cap - 1)];
        };
        auto process = [&](uint32_t k, const Par& pr) {
            const uint32_t q = rec_lane(k);
            const bool live = k * NF + f < nb;
            // shuffles with every lane active: a lane reading a record must
            // not depend on the source lane being live itself
            const uint32_t nev_q = (uint32_t)__shfl((int)g_nev, (int)q, 64);
            const uint32_t cnt_q = (uint32_t)__shfl((int)g_cnt, (int)q, 64);
            const uint32_t ne = live ? nev_q : 0u;
            const uint32_t nm = live ? cnt_q : 0u;
            const uint32_t row = s0 + j0 + q;
            // loop bounds: while some row of the step has events (mutations)
            // left.  kv holds events j FW .. j FW + FW - 1 of the lane's row.
            auto classify = [&](uint32_t kv, uint32_t j, uint64_t& m) {
                for (uint32_t x = 0; x < FW && __ballot(j * FW + x < ne); x += kEventGroup) {
                    uint32_t key[kEventGroup], p[kEventGroup];
                    bool more[kEventGroup];
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i)
                        key[i] = (uint32_t)__shfl((int)kv, (int)(f * FW + ((x + i) & (FW - 1))), 64);
                    const uint8_t* sb = reinterpret_cast<const uint8_t*>(s_bucket);
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i)
                        p[i] = sb[lane * 4u + (((key[i] >> 26) << 8) | ((key[i] >> 24) & 3u))];
                    uint32_t sk[kEventGroup][3];
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i)
#pragma unroll
                        for (int u = 0; u < 3; ++u) sk[i][u] = s_sorted[(p[i] + u) * 64 + lane];
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i) {
                        const bool real = j * FW + x + i < ne;
                        const uint32_t c = (sk[i][0] < key[i] ? 1u : 0u) + (sk[i][1] < key[i] ? 1u : 0u);
                        more[i] = real && sk[i][2] < key[i];
                        p[i] = real ? p[i] + c + (more[i] ? 1u : 0u) : 64u;
                    }
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i)
                        if (more[i])
                            while (p[i] < 64 && s_sorted[p[i] * 64 + lane] < key[i]) ++p[i];
                    uint64_t xm = 0;
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i) xm ^= s_pm[p[i] * 64 + lane];
                    m ^= xm;
                }
            };
            uint64_t m = 0;
            // the first NK FW events: keys already in registers, no load here
            // (a key that might come from a load issued in this loop would make
            // every use wait for all loads in flight, the prefetched ones included)
#pragma unroll
            for (uint32_t j = 0; j < NK; ++j)
                if (__ballot(j * FW < ne)) classify(pr.kv[j], j, m);
            // further events (rare): keys loaded here
            for (uint32_t j = NK; __ballot(j * FW < ne); ++j)
                classify(events[(size_t)row * ev_cap + min(j * FW + w, ev_cap - 1)], j, m);
            uint64_t c = ((pr.t & ~m) | (pr.b & m)) & keep;
            // new mutations (rare at these rates: their offsets load on demand)
            const bool any_nm = __ballot(nm != 0) != 0;
            const uint32_t k0 = any_nm ? local[row] + boff[row / kBlock] : 0u;
            for (uint32_t k = 0; __ballot(k < nm); ++k) {
                if (k < nm) {
                    const uint32_t kk = k0 + k;
                    const uint32_t slot = kk < F ? free_list[fb + kk] : A + (kk - F);
                    if (slot < S && (slot >> 6) == w) c ^= (uint64_t)1 << (slot & 63);
                }
            }
            // the step's rows through a range-checked buffer resource: lanes of
            // rows past the batch or of words past the row store out of range
            // and are dropped, so the store is issued on every path
            const uint32_t rbase = s0 + j0 + k * NF;  // first row of the step
            const uint32_t nrows = min(NF, nb - k * NF);
            const __amdgpu_buffer_rsrc_t out = __builtin_amdgcn_make_buffer_rsrc(
                child + (size_t)rbase * wc, 0, (int)(nrows * wc * 8u), kBufferWord3);
            typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
            const u32x2 v = {(uint32_t)c, (uint32_t)(c >> 32)};
            __builtin_amdgcn_raw_buffer_store_b64(v, out, (live && valid) ? (f * wc + w) * 8u : 0x80000000u,
                                                  0, kCachePolicyNT);
            acc_or |= live ? c : 0;
            acc_and &= live ? c : ~(uint64_t)0;
        };

        Par pa, pb;
        load_step(0, pa);
        load_step(1, pb);
        for (uint32_t k = 0; k < nst; k += 2) {
            process(k, pa);
            load_step(k + 2, pa);
            if (k + 1 >= nst) break;
            process(k + 1, pb);
            load_step(k + 3, pb);
        }
    }
    // rows of one word live in lanes w, w + FW, ...: fold them, then the waves
    for (uint32_t d = FW; d < 64; d <<= 1) {
        acc_or |= (uint64_t)__shfl_xor((long long)acc_or, (int)d, 64);
        acc_and &= (uint64_t)__shfl_xor((long long)acc_and, (int)d, 64);
    }
    __syncthreads();  // every wave is past its last prefix-mask read
    uint64_t* s_or = s_pm;
    uint64_t* s_and = s_pm + kRecWaves * 64;
    s_or[wave * 64 + lane] = acc_or;
    s_and[wave * 64 + lane] = acc_and;
    __syncthreads();
    if (wave == 0 && lane < FW && valid) {
        uint64_t vo = s_or[lane], va = s_and[lane];
        for (int q = 1; q < kRecWaves; ++q) {
            vo |= s_or[q * 64 + lane];
            va &= s_and[q * 64 + lane];
        }
        gbl(B.var_part)[(size_t)by * P.w_cap + w] = vo;
        gbl(B.fix_part)[(size_t)by * P.w_cap + w] = va;
    }
}

template <bool kWideEvents, int kSibs>
struct Recombine {
    static constexpr int kThreads = kRecBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.tiles * P.chunks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto parent = gbl(B.rows[sel]);
        const auto child = gbl(B.rows[1 - sel]);
        const auto free_list = gbl(B.free_list);
        const auto fixed_bits = gbl(B.fixed_bits);
        const auto top_in = gbl(B.top);
        const auto nev_in = gbl(B.n_events);
        const auto events = gbl(B.events);
        const auto count = gbl(B.mut_count);
        const auto local = gbl(B.mut_local);
        const auto boff = gbl(B.mut_bsum);
        const auto var_part = gbl(B.var_part);
        const auto fix_part = gbl(B.fix_part);
        const auto by_parent = (const __attribute__((address_space(1))) i8::u32x4*)B.by_parent;
        const bool ordered = B.by_parent != nullptr;
        const uint32_t bx = blk % P.tiles, by = blk / P.tiles;
        __shared__ __align__(16) uint32_t s_bucket[64 * 64];
        __shared__ __align__(16) uint32_t s_sorted[68 * 64];
        __shared__ __align__(16) uint64_t s_pm[65 * 64];  // reused for the OR/AND reduction at the end

        const uint32_t wc = uniform_read(&st->wc);
        const uint32_t w0 = bx * 64;
        if (w0 >= wc) return;  // uniform for the block
        if constexpr (kSibs == 1 && !kWideEvents) {
            // narrow rows: several rows per wave (tile 0 is the only live tile)
            if (wc <= 32)
                return recombine_folded(P, B, sel, by, wc <= 8 ? 3u : wc <= 16 ? 4u : 5u, s_bucket,
                                        s_sorted, s_pm);
        }
        const uint32_t wp = uniform_read(&st->wp);
        const uint32_t lane = threadIdx.x & 63;
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const uint32_t w = w0 + lane;
        const bool valid = w < wc;
        const bool inpar = valid && w < wp;
        const uint32_t col = wp ? min(w, wp - 1) : 0;  // in-bounds parent column
        const uint32_t A = uniform_read(&st->alleles), F = uniform_read(&st->free_count),
                       S = uniform_read(&st->all_size), fb = uniform_read(&st->free_base);
        const uint32_t ev_cap = P.ev_cap;  // >= 4, multiple of 4, <= 128

        load_word_tables(B, bx, wc, 64u, s_bucket, s_sorted, s_pm);
        const uint64_t keep = inpar ? ~fixed_bits[w] : 0;  // also clears columns >= wp
        __syncthreads();

        // positions [c0, c1) of this row chunk.  With sibling items a wave
        // takes the contiguous segment [s0, s1), so the children of one parent
        // meet in one wave; without, the waves interleave (wave w takes
        // c0 + w + 16 k), so a block writes consecutive child rows together.
        const uint32_t rows_per_chunk = (P.h + P.chunks - 1) / P.chunks;
        const uint32_t c0 = min(P.h, by * rows_per_chunk);
        const uint32_t c1 = min(P.h, c0 + rows_per_chunk);
        uint32_t first, stride, npos;
        if (kSibs > 1) {
            const uint32_t seg = (c1 - c0 + kRecWaves - 1) / kRecWaves;
            first = min(c1, c0 + wave * seg);
            stride = 1;
            npos = min(c1, first + seg) - first;
        } else {
            first = c0 + wave;
            stride = kRecWaves;
            npos = c1 > first ? (c1 - first + kRecWaves - 1) / kRecWaves : 0;
        }
        uint64_t acc_or = 0, acc_and = ~(uint64_t)0;

        for (uint32_t j0 = 0; j0 < npos; j0 += 64) {
            const uint32_t nb = min(64u, npos - j0);
            // gamete record of position j0 + lane (the last one repeated past
            // the batch): child row, top row (bottom = top ^ 1), events |
            // mutations << 8, first mutation
            const uint32_t pos = first + stride * (j0 + min(lane, nb - 1));
            uint32_t g_row, g_top, g_nx, g_k0;
            if (ordered) {
                const i8::u32x4 r = by_parent[pos];
                g_row = r.x;
                g_top = r.y;
                g_nx = r.z;
                g_k0 = r.w;
            } else {
                g_row = pos;
                g_top = top_in[pos];
                g_nx = nev_in[pos] | (count[pos] << 8);
                g_k0 = local[pos] + boff[pos / kBlock];
            }
            // items: runs of at most kSibs consecutive positions whose gametes
            // share the parent individual; one item loads the parent pair once
            uint64_t items = 0;
            if (kSibs > 1) {
                const uint32_t parent_id = g_top >> 1;
                const uint32_t prev_id = __shfl_up(parent_id, 1, 64);
                const uint64_t starts = __ballot(lane == 0 || lane >= nb || parent_id != prev_id);
                const uint64_t upto = lane == 63 ? ~(uint64_t)0 : (((uint64_t)2 << lane) - 1);
                const uint32_t run0 = 63u - (uint32_t)__builtin_clzll(starts & upto);
                items = __ballot(lane < nb && (lane - run0) % kSibs == 0);
            }
            uint32_t next_q = 0;  // kSibs == 1: every position is an item
            struct Item {
                uint32_t q, len;  // first position (lane) and number of children
                uint64_t t, b;    // parent rows top-of-first-child and its partner
                uint32_t lo[kSibs], hi[kSibs];
            };
            auto next_item = [&](Item& it) {
                if (kSibs == 1) {
                    it.q = next_q < nb ? next_q++ : 64u;
                    it.len = 1;
                } else if (items) {
                    it.q = (uint32_t)__builtin_ctzll(items);
                    items &= items - 1;
                    it.len = (items ? (uint32_t)__builtin_ctzll(items) : nb) - it.q;
                } else {
                    it.q = 64;
                    it.len = 0;
                }
            };
            // loads of an item (issued for every item slot, indices clamped in bounds)
            auto load_item = [&](Item& it) {
                const uint32_t qq = min(it.q, nb - 1);
                const uint32_t top = lane_read(g_top, qq);
                it.t = parent[(size_t)top * wp + col];
                it.b = parent[(size_t)(top ^ 1u) * wp + col];
#pragma unroll
                for (int k = 0; k < kSibs; ++k) {
                    const auto ev = events + (size_t)lane_read(g_row, min(qq + k, nb - 1)) * ev_cap;
                    it.lo[k] = ev[min(lane, ev_cap - 1)];
                    if (kWideEvents) it.hi[k] = ev[min(lane + 64, ev_cap - 1)];
                }
            };

            // A child whose top row is the item's bottom row takes the
            // complemented mask: (b & ~m) | (t & m) = (t & ~~m) | (b & ~m).
            auto child_row = [&](uint32_t q, uint64_t t, uint64_t b, uint32_t lo, uint32_t hi,
                                 bool flip) {
                const uint32_t nx = lane_read(g_nx, q);
                const uint32_t ne = nx & 0xFFu;
                uint64_t m = flip ? ~(uint64_t)0 : 0;
                for (uint32_t e = 0; e < ne; e += kEventGroup) {
                    uint32_t key[kEventGroup], p[kEventGroup];
                    bool more[kEventGroup];
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i) {
                        const uint32_t x = e + i;
                        const uint32_t a = lane_read(lo, x & 63);
                        key[i] = (kWideEvents && x >= 64) ? lane_read(hi, x & 63) : a;
                    }
                    // byte (key >> 24) & 3 of dword [key >> 26][lane]: one byte read
                    const uint8_t* sb = reinterpret_cast<const uint8_t*>(s_bucket);
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i)
                        p[i] = sb[lane * 4u + (((key[i] >> 26) << 8) | ((key[i] >> 24) & 3u))];
                    uint32_t sk[kEventGroup][3];
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) sk[i][j] = s_sorted[(p[i] + j) * 64 + lane];
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i) {
                        const bool live = e + i < ne;
                        const uint32_t c = (sk[i][0] < key[i] ? 1u : 0u) + (sk[i][1] < key[i] ? 1u : 0u);
                        more[i] = live && sk[i][2] < key[i];
                        p[i] = live ? p[i] + c + (more[i] ? 1u : 0u) : 64u;
                    }
                    bool any_more = false;
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i) any_more = any_more || more[i];
                    if (any_more) {  // three or more keys of a bucket below an event: rare
#pragma unroll
                        for (int i = 0; i < kEventGroup; ++i)
                            if (more[i])
                                while (p[i] < 64 && s_sorted[p[i] * 64 + lane] < key[i]) ++p[i];
                    }
                    uint64_t xm = 0;
#pragma unroll
                    for (int i = 0; i < kEventGroup; ++i) xm ^= s_pm[p[i] * 64 + lane];
                    m ^= xm;
                }
                // keep clears fixed columns and columns past the parent width in both
                uint64_t c = ((t & ~m) | (b & m)) & keep;
                // new mutations of this row (free columns: zero in every parent row)
                const uint32_t nm = nx >> 8;
                if (nm) {
                    const uint32_t k0 = lane_read(g_k0, q);
                    for (uint32_t k = k0; k < k0 + nm; ++k) {
                        const uint32_t slot = k < F ? free_list[fb + k] : A + (k - F);
                        if (slot < S && (slot >> 6) == w) c ^= (uint64_t)1 << (slot & 63);
                    }
                }
                {
                    // non-temporal store (cache policy NT): the traits pass that
                    // reads the child rows next runs ~12% faster than after
                    // ordinary write-back stores
                    const __amdgpu_buffer_rsrc_t row = __builtin_amdgcn_make_buffer_rsrc(
                        child + (size_t)lane_read(g_row, q) * wc, 0, (int)(wc * 8u), kBufferWord3);
                    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
                    const u32x2 v = {(uint32_t)c, (uint32_t)(c >> 32)};
                    __builtin_amdgcn_raw_buffer_store_b64(v, row, valid ? w * 8u : 0x80000000u, 0,
                                                          kCachePolicyNT);
                }
                acc_or |= c;
                acc_and &= c;
            };

            // each child takes the loaded pair in its own strand order
            auto process_item = [&](const Item& it) {
                const uint32_t top0 = lane_read(g_top, it.q);
#pragma unroll
                for (int k = 0; k < kSibs; ++k) {
                    if ((uint32_t)k < it.len) {
                        const uint32_t q = it.q + k;
                        const bool flip = lane_read(g_top, q) != top0;
                        child_row(q, it.t, it.b, it.lo[k], kWideEvents ? it.hi[k] : 0u, flip);
                    }
                }
            };

            Item iA, iB;
            next_item(iA);
            load_item(iA);
            next_item(iB);
            load_item(iB);
            while (iA.q < 64) {
                process_item(iA);
                next_item(iA);
                load_item(iA);
                if (iB.q >= 64) break;
                process_item(iB);
                next_item(iB);
                load_item(iB);
            }
        }
        __syncthreads();  // every wave is past its last prefix-mask read
        uint64_t* s_or = s_pm;
        uint64_t* s_and = s_pm + kRecWaves * 64;
        s_or[wave * 64 + lane] = acc_or;
        s_and[wave * 64 + lane] = acc_and;
        __syncthreads();
        if (wave == 0 && valid) {
            uint64_t vo = s_or[lane], va = s_and[lane];
            for (int q = 1; q < kRecWaves; ++q) {
                vo |= s_or[q * 64 + lane];
                va &= s_and[q * 64 + lane];
            }
            var_part[(size_t)by * P.w_cap + w] = vo;
            fix_part[(size_t)by * P.w_cap + w] = va;
        }
    }
};

// ---------------------------------------------------------------------------
// Free space: reduce chunk partials per word, classify, ordered compaction
// ---------------------------------------------------------------------------
// 64 words per block; the four waves fold interleaved chunk partials, then
// wave 0 classifies its word and scans the per-word counts.
constexpr uint32_t kFreeWords = 64;

struct FreeReduce {
    static constexpr int kThreads = 1024;  // 16 waves share the chunk partials of 64 words
    static constexpr int kRWaves = kThreads / 64;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.word_blocks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto var_part = gbl(B.var_part);
        const auto fix_part = gbl(B.fix_part);
        const uint32_t chunks = P.chunks;
        const auto fixed_bits = gbl(B.fixed_bits);
        const auto lost_bits = gbl(B.var_part) + (size_t)P.max_chunks * P.w_cap;
        const auto word_cnt = gbl(B.word_cnt);
        const auto word_bsum = gbl(B.word_bsum);
        const uint32_t word_blocks = P.word_blocks;
        __shared__ uint64_t s_or[kRWaves][64];
        __shared__ uint64_t s_and[kRWaves][64];
        const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const uint32_t w = blk * kFreeWords + lane;
        const uint32_t wc = st->wc, S = st->all_size;
        uint64_t vo = 0, va = ~(uint64_t)0;
        if (w < wc) {
            for (uint32_t c = wave; c < chunks; c += kRWaves) {
                vo |= var_part[(size_t)c * P.w_cap + w];
                va &= fix_part[(size_t)c * P.w_cap + w];
            }
        }
        s_or[wave][lane] = vo;
        s_and[wave][lane] = va;
        __syncthreads();
        if (wave) return;  // no barrier below
        for (int q = 1; q < kRWaves; ++q) {
            vo |= s_or[q][lane];
            va &= s_and[q][lane];
        }
        uint32_t cf = 0, cl = 0;
        if (w < wc) {
            const uint32_t lo = w * 64;
            uint64_t validm = 0;
            if (S > lo) validm = (S - lo >= 64) ? ~(uint64_t)0 : (((uint64_t)1 << (S - lo)) - 1);
            const uint64_t fx = va & validm, ls = ~vo & validm;
            cf = (uint32_t)__popcll(fx);
            cl = (uint32_t)__popcll(ls);
            // after an intermediate generation nothing is cleared from the
            // next parents (their fixed columns were handed out as free slots)
            fixed_bits[w] = st->inter == P.inter_k ? fx : 0;
            lost_bits[w] = ls;
        }
        // wave-level exclusive scans of the fixed and lost counts
        uint32_t xf = cf, xl = cl;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t yf = __shfl_up(xf, d, 64), yl = __shfl_up(xl, d, 64);
            if ((int)lane >= d) {
                xf += yf;
                xl += yl;
            }
        }
        if (w < wc) word_cnt[w] = ((xf - cf) & 0xFFFFu) | ((xl - cl) << 16);
        if (lane == 63) {
            word_bsum[blk] = xf;
            word_bsum[word_blocks + blk] = xl;
        }
    }
};

struct FreeCompact {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return P.word_blocks; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        (void)sel;
        const auto st = gbl(B.state);
        const auto fixed_bits = gbl(B.fixed_bits);
        const auto lost_bits = gbl(B.var_part) + (size_t)P.max_chunks * P.w_cap;
        const auto word_cnt = gbl(B.word_cnt);
        const auto word_bsum = gbl(B.word_bsum);
        const uint32_t word_blocks = P.word_blocks;
        const auto free_list = gbl(B.free_list);
        __shared__ uint32_t s_off[2];
        __shared__ uint32_t s_red[2][kBlock];
        const uint32_t wc = st->wc;
        const uint32_t nb = (wc + kFreeWords - 1) / kFreeWords;
        if (blk >= nb) return;
        // intermediate generation: the free list of the last analysis stays
        if (st->inter != P.inter_k) return;
        // offsets of this block and the fixed total, from the block totals
        uint32_t pf = 0, tfix = 0, pl = 0;
        for (uint32_t q = threadIdx.x; q < nb; q += kBlock) {
            const uint32_t a = word_bsum[q], l = word_bsum[word_blocks + q];
            tfix += a;
            if (q < blk) {
                pf += a;
                pl += l;
            }
        }
        s_red[0][threadIdx.x] = pf;
        s_red[1][threadIdx.x] = pl;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) {
                s_red[0][threadIdx.x] += s_red[0][threadIdx.x + s];
                s_red[1][threadIdx.x] += s_red[1][threadIdx.x + s];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            s_off[0] = s_red[0][0];
            s_off[1] = s_red[1][0];
        }
        __syncthreads();
        s_red[0][threadIdx.x] = tfix;
        __syncthreads();
        for (int s = kBlock / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) s_red[0][threadIdx.x] += s_red[0][threadIdx.x + s];
            __syncthreads();
        }
        const uint32_t nfixed = s_red[0][0];
        const uint32_t w = blk * kFreeWords + threadIdx.x;
        if (threadIdx.x < kFreeWords && w < wc) {
            const uint32_t wcnt = word_cnt[w];
            uint32_t pos = s_off[0] + (wcnt & 0xFFFFu);
            uint64_t fx = fixed_bits[w];
            while (fx) {
                free_list[pos++] = w * 64 + (uint32_t)__builtin_ctzll(fx);
                fx &= fx - 1;
            }
            pos = nfixed + s_off[1] + (wcnt >> 16);
            uint64_t ls = lost_bits[w];
            while (ls) {
                free_list[pos++] = w * 64 + (uint32_t)__builtin_ctzll(ls);
                ls &= ls - 1;
            }
        }
        if (blk == nb - 1 && threadIdx.x == 0) {
            uint32_t tl = 0;
            for (uint32_t q = 0; q < nb; ++q) tl += word_bsum[word_blocks + q];
            st->n_fixed = nfixed;
            st->free_count = nfixed + tl;
        }
    }
};

// Column OR/AND straight from a resident population (after an upload).
__global__ void __launch_bounds__(kBlock) k_column_reduce(Params P, const GenState* st,
                                                          const uint64_t* rows,
                                                          uint64_t* var_part,
                                                          uint64_t* fix_part) {
    const uint32_t w = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t wc = st->wc;
    if (w >= wc) return;
    const uint32_t rows_per_chunk = (P.h + gridDim.y - 1) / gridDim.y;
    const uint32_t r0 = blockIdx.y * rows_per_chunk, r1 = min(P.h, r0 + rows_per_chunk);
    uint64_t vo = 0, va = ~(uint64_t)0;
    for (uint32_t r = r0; r < r1; ++r) {
        const uint64_t x = rows[(size_t)r * wc + w];
        vo |= x;
        va &= x;
    }
    var_part[(size_t)blockIdx.y * P.w_cap + w] = vo;
    fix_part[(size_t)blockIdx.y * P.w_cap + w] = va;
}

// ---------------------------------------------------------------------------
// Column compaction: fixed and lost columns of the current population are
// removed, the variable columns keep their order and are renumbered (no
// reference counterpart; the reference reuses free slots instead,
// QS/engine_parallel_pipeline.hpp:337-348, and its CUDA engine only clears
// them, CU/data_spaces/population_space/population_space_helper_kernels.hpp:457-510).
// Per word w the keep mask m_w = ~(fixed | lost) (valid slots only) is the
// same for every row, so its compress plan is built once: dst[w] = kept bits
// before word w, and the six move masks of the parallel bit compress of
// Hacker's Delight (2nd ed., section 7-4, "compress"), after which packing a
// row word is x &= m; x = (x ^ t) | (t >> 2^i) with t = x & mv_i, i = 0..5.
// Scratch: keep masks in fix_part, dst in word_cnt, move masks in pm
// ([6][w_cap]), compacted allele attributes in sorted / bucket / digits; all
// of them are rebuilt by the next generation.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t compress_bits(uint64_t x, const uint64_t (&mv)[6]) {
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const uint64_t t = x & mv[i];
        x = (x ^ t) | (t >> (1 << i));
    }
    return x;
}

__global__ void __launch_bounds__(1024) k_compact_plan(Params P, Buffers B) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_carry;
    const auto st = gbl(B.state);
    const uint32_t W = uniform_read(&st->wp), A = uniform_read(&st->alleles);
    const auto fixed_bits = gbl(B.fixed_bits);
    const auto lost_bits = gbl(B.var_part) + (size_t)P.max_chunks * P.w_cap;
    const auto keep_out = gbl(B.fix_part);
    const auto dst = gbl(B.word_cnt);
    const auto mv = gbl(B.pm);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < W; base += 1024) {
        const uint32_t w = base + tid;
        uint64_t keep = 0;
        if (w < W) {
            const uint32_t lo = w * 64;
            const uint64_t valid =
                A > lo ? (A - lo >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << (A - lo)) - 1)) : 0;
            keep = ~(fixed_bits[w] | lost_bits[w]) & valid;
            keep_out[w] = keep;
            uint64_t m = keep, mk = ~m << 1;
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                uint64_t mp = mk ^ (mk << 1);
                mp ^= mp << 2;
                mp ^= mp << 4;
                mp ^= mp << 8;
                mp ^= mp << 16;
                mp ^= mp << 32;
                const uint64_t v = mp & m;
                mv[(size_t)i * P.w_cap + w] = v;
                m = (m ^ v) | (v >> (1 << i));
                mk &= ~mp;
            }
        }
        const uint32_t cnt = (uint32_t)__popcll(keep);
        uint32_t x = cnt;
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        if (lane == 63) s_wave[wave] = x;
        __syncthreads();
        uint32_t off = s_carry, tot = 0;
        for (int q = 0; q < 16; ++q) {
            if (q < wave) off += s_wave[q];
            tot += s_wave[q];
        }
        if (w < W) dst[w] = off + x - cnt;
        __syncthreads();
        if (tid == 0) s_carry += tot;
        __syncthreads();
    }
    if (tid == 0) st->kept = s_carry;
}

// Kept slot j moves to dst[j/64] + popcount(keep below j).
__global__ void __launch_bounds__(kBlock) k_compact_alleles(Params P, Buffers B) {
    const auto st = gbl(B.state);
    const uint32_t A = uniform_read(&st->alleles);
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= A) return;
    const uint64_t keep = gbl(B.fix_part)[j / 64];
    if (!((keep >> (j % 64)) & 1)) return;
    const uint32_t n =
        gbl(B.word_cnt)[j / 64] + (uint32_t)__popcll(keep & (((uint64_t)1 << (j % 64)) - 1));
    gbl(B.sorted)[n] = gbl(B.keys)[j];
    gbl(B.bucket)[n] = gbl(B.ages)[j];
    const auto eff = gbl(B.effects);
    const auto tmp = gbl((int32_t*)B.digits);
    for (uint32_t t = 0; t < P.traits; ++t) tmp[(size_t)n * P.traits + t] = eff[(size_t)j * P.traits + t];
}

// Rows: wave = one tile of TW <= 64 words of the compacted row over a chunk
// of rows; block = kCompactWaves adjacent tiles over the same chunk.  The
// input words whose kept bits fall in a tile are [wa, wb) (dst is
// non-decreasing); every input word contributes its packed bits to at most
// two output words of the tile, OR-ed into a per-wave LDS row (no global
// atomics).  When the tile draws on at most 64 input words lane l keeps the
// plan of input word wa + l in registers (of wa + l and wa + 64 + l up to 128
// input words) and the wave packs kCompactRows rows per step, their loads in
// flight together; otherwise it packs one row at a time, reading the plan
// per row.  With a fraction f of the columns removed a 64-word tile draws on
// about 64 / (1 - f) + 1 input words.  The few words past the first 64 (up
// to 64 / kCompactRows of them) are packed for all the step's rows in one
// compress, lanes spread over (row, word); more take the two-plan path,
// which compresses a second set of words per row.  Narrower tiles stay on one
// plan but write rows in pieces that do not start on 128-byte lines.
constexpr int kCompactWaves = 4;

// Block x of a grid of nb runs on XCD x mod 8; the returned index is a
// bijection of [0, nb) under which XCD k runs one contiguous range (XCDs
// below nb mod 8 take one block more).
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t x, uint32_t nb) {
    const uint32_t q = nb / 8, r = nb % 8, xcd = x % 8, i = x / 8;
    return xcd < r ? xcd * (q + 1) + i : r * (q + 1) + (xcd - r) * q + i;
}

// One packed input word into a tile's LDS row: x holds the word's n kept
// bits, packed, and d is their tile-relative destination bit.
__device__ __forceinline__ void or_packed(unsigned long long* row, uint64_t x, int d, uint32_t n) {
    if (d >= 0) {
        const uint32_t u0 = (uint32_t)d >> 6, s = (uint32_t)d & 63;
        if (u0 < 64) atomicOr(&row[u0], (unsigned long long)(x << s));
        if (s && s + n > 64 && u0 + 1 < 64) atomicOr(&row[u0 + 1], (unsigned long long)(x >> (64 - s)));
    } else if (-d < (int)n) {
        atomicOr(&row[0], (unsigned long long)(x >> (-d)));
    }
}

// The planned path of k_compact_rows: lane l holds the plan (keep mask,
// compress masks, destination bit) of input words wa + l + 64 i, i < NWL, and
// the wave packs kCompactRows rows per step from row o_first, `step` rows
// apart, up to row r1.  With kSpill (NWL = 1) the tile draws on m = wb - wa -
// 64 more words, m * kCompactRows <= 64, and these are packed for all the
// step's rows at once: lane l holds the plan of word wa + 64 + l % m and packs
// it for row l / m of the step, so the m words cost one compress per step
// instead of one per row.
template <int NWL, bool kSpill, int kCompactRows, class In, class Out, class Keep, class Dst,
          class Mv>
__device__ __forceinline__ void compact_rows_planned(const Params& P, In in, Out out, Keep keepm,
                                                     Dst dst, Mv mvg,
                                                     unsigned long long (*s_row)[64], uint32_t W,
                                                     uint32_t W2, uint32_t tw, uint32_t wa,
                                                     uint32_t wb, uint32_t lo_bit,
                                                     uint32_t o_first, uint32_t r1,
                                                     uint32_t step) {
    static_assert(!kSpill || NWL == 1, "spill words follow a single plan");
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t u = lo_bit / 64 + lane;
    const bool writes = lane < tw && u < W2;
    uint32_t wr[NWL], n[NWL];
    uint64_t keep[NWL], mv[NWL][6];
    int d[NWL];
#pragma unroll
    for (int i = 0; i < NWL; ++i) {
        const uint32_t w = wa + 64 * i + lane;
        const bool active = w < wb;
        wr[i] = active ? w : wa;  // in-bounds word for the unconditional loads
        keep[i] = active ? keepm[wr[i]] : 0;
        n[i] = (uint32_t)__popcll(keep[i]);
#pragma unroll
        for (int j = 0; j < 6; ++j) mv[i][j] = mvg[(size_t)j * P.w_cap + wr[i]];
        d[i] = (int)dst[wr[i]] - (int)lo_bit;  // tile-relative bit of the packed word
    }
    // spill plan (kSpill): lanes past m * kCompactRows keep no bits; their
    // row and word are clamped in bounds for the unconditional loads
    uint32_t sk = 0, sw = wa, sn = 0;
    uint64_t skeep = 0, smv[6] = {};
    int sd = 0;
    if constexpr (kSpill) {
        const uint32_t m = wb - wa - 64;
        sk = min(lane / m, (uint32_t)kCompactRows - 1);
        sw = wa + 64 + lane % m;
        skeep = lane < m * kCompactRows ? keepm[sw] : 0;
        sn = (uint32_t)__popcll(skeep);
#pragma unroll
        for (int j = 0; j < 6; ++j) smv[j] = mvg[(size_t)j * P.w_cap + sw];
        sd = (int)dst[sw] - (int)lo_bit;
    }
    if (o_first >= r1) return;
    // software pipeline: the rows of step s + 1 are loaded (row indices
    // clamped into the chunk, so every load is unconditional and in bounds)
    // before step s is packed, keeping loads in flight through the packing
    uint64_t xin[kCompactRows][NWL], xs = 0;
#pragma unroll
    for (int k = 0; k < kCompactRows; ++k)
#pragma unroll
        for (int i = 0; i < NWL; ++i) xin[k][i] = in[(size_t)min(o_first + k, r1 - 1) * W + wr[i]];
    if constexpr (kSpill) xs = in[(size_t)min(o_first + sk, r1 - 1) * W + sw];
    for (uint32_t o0 = o_first; o0 < r1; o0 += step) {
        uint64_t xnx[kCompactRows][NWL], xsn = 0;
        const uint32_t o1 = min(o0 + step, r1 - 1);
#pragma unroll
        for (int k = 0; k < kCompactRows; ++k)
#pragma unroll
            for (int i = 0; i < NWL; ++i) xnx[k][i] = in[(size_t)min(o1 + k, r1 - 1) * W + wr[i]];
        if constexpr (kSpill) xsn = in[(size_t)min(o1 + sk, r1 - 1) * W + sw];
#pragma unroll
        for (int k = 0; k < kCompactRows; ++k) s_row[k][lane] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int i = 0; i < NWL; ++i) {
            if (!n[i]) continue;
#pragma unroll
            for (int k = 0; k < kCompactRows; ++k)
                or_packed(s_row[k], compress_bits(xin[k][i] & keep[i], mv[i]), d[i], n[i]);
        }
        if constexpr (kSpill) {
            if (sn) or_packed(s_row[sk], compress_bits(xs & skeep, smv), sd, sn);
            xs = xsn;
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        // every packed row into its own registers first, then the stores: a
        // store's data registers may not be overwritten before the store
        // completes, so stores sharing one register pair would each wait out
        // the previous store's write
        uint64_t packed[kCompactRows];
#pragma unroll
        for (int k = 0; k < kCompactRows; ++k) packed[k] = s_row[k][lane];
#pragma unroll
        for (int k = 0; k < kCompactRows; ++k)
            if (o0 + k < r1 && writes) out[(size_t)(o0 + k) * W2 + u] = packed[k];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
        for (int k = 0; k < kCompactRows; ++k)
#pragma unroll
            for (int i = 0; i < NWL; ++i) xin[k][i] = xnx[k][i];
    }
}

template <int kCompactRows>
__global__ void __launch_bounds__(64 * kCompactWaves) k_compact_rows(Params P, Buffers B,
                                                                     uint32_t src, uint32_t tw,
                                                                     uint32_t tiles) {
    __shared__ unsigned long long s_out[kCompactWaves][kCompactRows][64];
    const auto st = gbl(B.state);
    const uint32_t V = uniform_read(&st->kept), W = uniform_read(&st->wp);
    const uint32_t W2 = ((V + 511) / 512) * 8;
    // The input windows of neighbouring output tiles overlap by a word or
    // two, so their edge lines are read by both.  A block's waves take
    // kCompactWaves adjacent tiles over the same rows in the same order,
    // and blocks, which go to the eight XCDs (each with its own L2)
    // round-robin, are numbered so that each XCD runs a contiguous range of
    // (chunk, tile group) blocks: the second read of an edge line finds it
    // in the unit's L1 or the XCD's L2 instead of fetching it again.
    const uint32_t tgroups = (tiles + kCompactWaves - 1) / kCompactWaves;
    const uint32_t vb = xcd_contiguous(blockIdx.x, gridDim.x);
    const uint32_t tg = vb % tgroups, chunk = vb / tgroups;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t tile = tg * kCompactWaves + wave;
    if (tile >= tiles || tile * tw >= W2) return;  // uniform for the wave
    const auto in = gbl(B.rows[src]);
    const auto out = gbl(B.rows[1 - src]);
    const auto keepm = gbl(B.fix_part);
    const auto dst = gbl(B.word_cnt);
    const auto mvg = gbl(B.pm);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t lo_bit = tile * tw * 64, hi_bit = lo_bit + tw * 64;
    // wa = last word with dst <= lo_bit, wb = first word with dst >= hi_bit
    uint32_t a = 0, b = W;
    while (b - a > 1) {
        const uint32_t mid = (a + b) >> 1;
        if (dst[mid] <= lo_bit) a = mid; else b = mid;
    }
    const uint32_t wa = a;
    a = wa;
    b = W;
    while (a < b) {
        const uint32_t mid = (a + b) >> 1;
        if (dst[mid] >= hi_bit) b = mid; else a = mid + 1;
    }
    const uint32_t wb = a;
    const uint32_t chunks = gridDim.x / tgroups;
    const uint32_t rows_per_chunk = (P.h + chunks - 1) / chunks;
    const uint32_t r0 = chunk * rows_per_chunk, r1 = min(P.h, r0 + rows_per_chunk);
    const uint32_t u = tile * tw + lane;
    if (wb - wa <= 64) {
        compact_rows_planned<1, false, kCompactRows>(P, in, out, keepm, dst, mvg, s_out[wave], W,
                                                     W2, tw, wa, wb, lo_bit, r0, r1, kCompactRows);
        return;
    }
    if ((wb - wa - 64) * kCompactRows <= 64) {
        compact_rows_planned<1, true, kCompactRows>(P, in, out, keepm, dst, mvg, s_out[wave], W,
                                                    W2, tw, wa, wb, lo_bit, r0, r1, kCompactRows);
        return;
    }
    if (wb - wa <= 128) {
        compact_rows_planned<2, false, kCompactRows>(P, in, out, keepm, dst, mvg, s_out[wave], W,
                                                     W2, tw, wa, wb, lo_bit, r0, r1, kCompactRows);
        return;
    }
    for (uint32_t o = r0; o < r1; ++o) {
        s_out[wave][0][lane] = 0;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        for (uint32_t base = wa; base < wb; base += 64) {
            const uint32_t w = base + lane;
            if (w >= wb) continue;
            const uint64_t keep = keepm[w];
            const uint32_t n = (uint32_t)__popcll(keep);
            if (!n) continue;
            uint64_t mv[6];
#pragma unroll
            for (int i = 0; i < 6; ++i) mv[i] = mvg[(size_t)i * P.w_cap + w];
            // tile-relative destination bit of the packed word
            or_packed(s_out[wave][0], compress_bits(in[(size_t)o * W + w] & keep, mv),
                      (int)dst[w] - (int)lo_bit, n);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        if (lane < tw && u < W2) out[(size_t)o * W2 + u] = s_out[wave][0][lane];
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    }
}

// Allele attributes back into place (tail zeroed), free-space bits cleared,
// traits and fitness carried to the population's new buffer index.
__global__ void __launch_bounds__(kBlock) k_compact_finish(Params P, Buffers B, uint32_t src) {
    const auto st = gbl(B.state);
    const uint32_t V = uniform_read(&st->kept), A = uniform_read(&st->alleles);
    const uint32_t W = uniform_read(&st->wp);
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < A) {
        gbl(B.keys)[i] = i < V ? gbl(B.sorted)[i] : 0u;
        gbl(B.ages)[i] = i < V ? gbl(B.bucket)[i] : 0u;
        const auto tmp = gbl((int32_t*)B.digits);
        for (uint32_t t = 0; t < P.traits; ++t)
            gbl(B.effects)[i * P.traits + t] = i < V ? tmp[i * P.traits + t] : 0;
    }
    if (i < W) {
        gbl(B.fixed_bits)[i] = 0;
        (gbl(B.var_part) + (size_t)P.max_chunks * P.w_cap)[i] = 0;
    }
    if (i < (size_t)P.h * P.traits) gbl(B.traits[1 - src])[i] = gbl(B.traits[src])[i];
    if (i < P.n) gbl(B.fitness[1 - src])[i] = gbl(B.fitness[src])[i];
}

__global__ void k_compact_state(Buffers B) {
    GenState* st = B.state;
    const uint32_t V = st->kept;
    st->alleles = V;
    st->all_size = V;
    st->wp = st->wc = ((V + 511) / 512) * 8;
    st->free_count = 0;
    st->n_fixed = 0;
    st->free_base = 0;
}

// ---------------------------------------------------------------------------
// Population statistics: column and row margins over a row sample
// (frequency_evaluator<population_space<row_block_alignment>>::operator(),
// include/clotho/data_spaces/analysis/allele_frequency/frequency_evaluator_population_space_row_alignment.hpp:36-61).
// Block = 16 waves over 16 consecutive words (one 128-byte line per row) of
// groups of 64 sampled rows, lane = row.  Row margins: per-word popcounts
// summed across the 16 waves in LDS.  Column margins: the wave's 64 x 64 bit
// block (lane = row, bit = column) is transposed across lanes in six
// butterfly stages, after which lane = column and a popcount counts that
// column's carriers among the 64 rows.  Partials per (row chunk, column) and
// (word group, row) are summed by two small kernels (no atomics).
// ---------------------------------------------------------------------------
constexpr int kStatWaves = (int)kStatBlockWaves;

// x of lane ^ S: DPP quad permutes for S = 1, 2, swizzles for S = 4, 8, 16
// (both stay inside 32-lane halves), a bpermute for S = 32.
template <uint32_t S>
__device__ __forceinline__ uint32_t lane_xor(uint32_t v, uint32_t lane) {
    if constexpr (S == 1) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
    if constexpr (S == 2) return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false);
    if constexpr (S >= 4 && S <= 16)
        return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, (int)((S << 10) | 0x1F));
    return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((lane ^ S) << 2), (int)v);
}

// One butterfly stage: rows l and l ^ S exchange the off-diagonal S x S
// blocks selected by M (the low S bits of every 2S-bit block).  A row with
// bit S clear keeps its M bits and takes its partner's M bits, shifted up,
// into ~M; a row with it set keeps ~M and takes the partner's ~M bits,
// shifted down.  Per lane the stage is fixed by K (the bits kept: M or ~M)
// and the shift pair (S, 0) or (0, S), computed once per kernel.
struct TransposeStage {
    uint32_t k_lo, k_hi, up, down;
};

template <uint32_t S>
__device__ __forceinline__ TransposeStage transpose_stage_of(uint32_t lane, uint64_t M) {
    const uint32_t b = lane & S;
    const uint64_t K = b ? ~M : M;
    return TransposeStage{(uint32_t)K, (uint32_t)(K >> 32), S - b, b};
}

template <uint32_t S>
__device__ __forceinline__ uint64_t transpose_stage(uint64_t x, uint32_t lane,
                                                    const TransposeStage& t) {
    const uint32_t lo = lane_xor<S>((uint32_t)x, lane);
    const uint32_t hi = lane_xor<S>((uint32_t)(x >> 32), lane);
    const uint64_t z = ((((uint64_t)hi << 32) | lo) << t.up) >> t.down;
    const uint32_t xl = (uint32_t)x, xh = (uint32_t)(x >> 32);
    const uint32_t zl = (uint32_t)z, zh = (uint32_t)(z >> 32);
    const uint32_t rl = (xl & t.k_lo) | (zl & ~t.k_lo);
    const uint32_t rh = (xh & t.k_hi) | (zh & ~t.k_hi);
    return ((uint64_t)rh << 32) | rl;
}

struct Transpose64 {
    TransposeStage s32, s16, s8, s4, s2, s1;
    __device__ explicit Transpose64(uint32_t lane)
        : s32(transpose_stage_of<32>(lane, 0x00000000FFFFFFFFull)),
          s16(transpose_stage_of<16>(lane, 0x0000FFFF0000FFFFull)),
          s8(transpose_stage_of<8>(lane, 0x00FF00FF00FF00FFull)),
          s4(transpose_stage_of<4>(lane, 0x0F0F0F0F0F0F0F0Full)),
          s2(transpose_stage_of<2>(lane, 0x3333333333333333ull)),
          s1(transpose_stage_of<1>(lane, 0x5555555555555555ull)) {}
    // 64 x 64 bit transpose across a wave: row l (lane l, bit j = column j)
    // becomes column l (lane l, bit j = row j)
    __device__ uint64_t operator()(uint64_t x, uint32_t lane) const {
        x = transpose_stage<32>(x, lane, s32);
        x = transpose_stage<16>(x, lane, s16);
        x = transpose_stage<8>(x, lane, s8);
        x = transpose_stage<4>(x, lane, s4);
        x = transpose_stage<2>(x, lane, s2);
        return transpose_stage<1>(x, lane, s1);
    }
};

// Carry-save adder on bit planes: (hi, lo) = a + b + c per bit position.
__device__ __forceinline__ void csa(uint64_t& hi, uint64_t& lo, uint64_t a, uint64_t b,
                                    uint64_t c) {
    const uint64_t u = a ^ b;
    hi = (a & b) | (u & c);
    lo = u ^ c;
}

// Block = kStatWaves waves over one tile of kStatTileWords = 64 words (512
// bytes per row) and a chunk of at most kStatChunkRows sampled rows, split
// between the waves and, within a wave, between its halves: lane l holds
// words 2 (l % 32) and 2 (l % 32) + 1 of the tile (one 16-byte load per row)
// for the rows of its half-wave h = l / 32 (at most 255).  Rows are loaded 8
// per half at a time, the next 8 in flight while the current 8 are counted.
// Column margins: each lane keeps, per word, a bit-sliced counter of 8 planes
// (plane i, bit j = bit i of the number of its rows carrying column j), fed 8
// rows at a time through a carry-save (Harley-Seal) tree: 7 adders per 8 rows
// for the ones, twos and fours planes, and the eights carry rippled into the
// planes above.  At the end of the chunk a word's 8 counters (kStatWaves
// waves x 2 halves) x 8 planes are exactly 64 x 64 bits: one transpose across
// a wave leaves lane j = column j holding the 8 counts of column j as bytes,
// summed by two v_sad_u8.  Row margins: each lane's per-row popcounts of its
// two words are summed over the 32 lanes of its half by a reduce-scatter
// butterfly (8 rows -> 4 -> 2 -> 1 value per lane, then two full exchanges).
// Partials per (row chunk, column) and (word tile, row) are summed by two
// small kernels (no atomics).  kSample: rows come from sample_in (all eight
// row indices of a batch are loaded before its rows; a per-row choice
// between the two would make every row load wait for the previous ones).
template <bool kSample>
__global__ void __launch_bounds__(64 * kStatWaves) k_allele_counts(
    const uint64_t* __restrict__ rows_in, uint32_t W, const uint32_t* __restrict__ sample_in,
    uint32_t n, uint32_t rows_per_chunk, uint32_t* __restrict__ col_part_in,
    uint32_t* __restrict__ row_part_in) {
    constexpr uint32_t TW = kStatTileWords, kContrib = 2 * kStatWaves, kPlanes = 8;
    static_assert(kContrib * kPlanes == 64, "a word's counters fill one 64 x 64 transpose");
    __shared__ uint64_t s_pl[TW][kContrib * kPlanes + 1];  // +1: spread the writers' banks
    const auto rows = gbl(rows_in);
    const auto col_part = gbl(col_part_in);
    const auto row_part = gbl(row_part_in);
    const auto sample = gbl(sample_in);
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t h = lane >> 5, pr = lane & 31;
    const uint32_t ntw = stat_word_tiles(W);
    const uint32_t wt = blockIdx.x % ntw, chunk = blockIdx.x / ntw;
    const uint32_t w0 = wt * TW;
    const uint32_t r0 = min(n, chunk * rows_per_chunk), r1 = min(n, r0 + rows_per_chunk);
    const uint32_t per_wave = (r1 - r0 + kStatWaves - 1) / kStatWaves;
    const uint32_t rw0 = min(r1, r0 + wave * per_wave), rw1 = min(r1, rw0 + per_wave);
    const uint32_t wlocal = 2 * pr;
    const bool wok = w0 + wlocal < W;  // W is even: both words are in the row
    const uint32_t wsrc = wok ? w0 + wlocal : 0;
    uint64_t pl0[kPlanes], pl1[kPlanes];
#pragma unroll
    for (uint32_t i = 0; i < kPlanes; ++i) pl0[i] = pl1[i] = 0;
    // the half's rows are rw0 + 2 i + h; batches of 8 per half, every load
    // unconditional (rows past the wave's range re-read its first row and are
    // zeroed when counted)
    const uint32_t half_rows = (rw1 - rw0 + 1) / 2;
    auto load8 = [&](uint32_t i0, i8::u32x4 (&x)[8]) {
        uint32_t r[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
            const uint32_t k = rw0 + 2 * (i0 + t) + h;
            const uint32_t kc = k < rw1 ? k : rw0;
            r[t] = kSample ? sample[kc] : kc;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t)
            x[t] = *(const __attribute__((address_space(1))) i8::u32x4*)(rows + (size_t)r[t] * W +
                                                                         wsrc);
    };
    i8::u32x4 x[8];
    if (half_rows) load8(0, x);
    for (uint32_t i0 = 0; i0 < half_rows; i0 += 8) {
        i8::u32x4 xn[8];
        if (i0 + 8 < half_rows) load8(i0 + 8, xn);
#pragma unroll
        for (int t = 0; t < 8; ++t)
            if (!(rw0 + 2 * (i0 + t) + h < rw1 && wok)) x[t] = i8::u32x4{0, 0, 0, 0};
        // row margins: popcounts of the lane's two words, summed over the half
        uint32_t v[8];
#pragma unroll
        for (int t = 0; t < 8; ++t)
            v[t] = (uint32_t)(__builtin_popcount(x[t].x) + __builtin_popcount(x[t].y) +
                              __builtin_popcount(x[t].z) + __builtin_popcount(x[t].w));
        {
            const bool b16 = pr & 16, b8 = pr & 8, b4 = pr & 4;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint32_t keep = b16 ? v[j + 4] : v[j], send = b16 ? v[j] : v[j + 4];
                v[j] = keep + lane_xor<16>(send, lane);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const uint32_t keep = b8 ? v[j + 2] : v[j], send = b8 ? v[j] : v[j + 2];
                v[j] = keep + lane_xor<8>(send, lane);
            }
            {
                const uint32_t keep = b4 ? v[1] : v[0], send = b4 ? v[0] : v[1];
                v[0] = keep + lane_xor<4>(send, lane);
            }
            v[0] += lane_xor<2>(v[0], lane);
            v[0] += lane_xor<1>(v[0], lane);
            const uint32_t t = (b16 ? 4u : 0u) + (b8 ? 2u : 0u) + (b4 ? 1u : 0u);
            const uint32_t k = rw0 + 2 * (i0 + t) + h;
            if ((pr & 3) == 0 && k < rw1) row_part[(size_t)wt * n + k] = v[0];
        }
        // column margins: 8 rows into the bit-sliced counters of each word
#pragma unroll
        for (int wd = 0; wd < 2; ++wd) {
            uint64_t d[8];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                d[t] = wd ? (((uint64_t)x[t].w << 32) | x[t].z) : (((uint64_t)x[t].y << 32) | x[t].x);
            uint64_t* pl = wd ? pl1 : pl0;
            uint64_t tA, tB, fA, fB, e;
            csa(tA, pl[0], pl[0], d[0], d[1]);
            csa(tB, pl[0], pl[0], d[2], d[3]);
            csa(fA, pl[1], pl[1], tA, tB);
            csa(tA, pl[0], pl[0], d[4], d[5]);
            csa(tB, pl[0], pl[0], d[6], d[7]);
            csa(fB, pl[1], pl[1], tA, tB);
            csa(e, pl[2], pl[2], fA, fB);
#pragma unroll
            for (uint32_t i = 3; i < kPlanes; ++i) {
                const uint64_t c = pl[i] & e;
                pl[i] ^= e;
                e = c;
            }
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) x[t] = xn[t];
    }
    // a word's counters into LDS: row (contributor c = 2 wave + h, plane i)
    const uint32_t c = 2 * wave + h;
#pragma unroll
    for (uint32_t i = 0; i < kPlanes; ++i) {
        s_pl[wlocal][c * kPlanes + i] = pl0[i];
        s_pl[wlocal + 1][c * kPlanes + i] = pl1[i];
    }
    __syncthreads();
    const Transpose64 transpose64(lane);
    for (uint32_t wl = wave; wl < TW; wl += kStatWaves) {
        const uint64_t tr = transpose64(s_pl[wl][lane], lane);  // byte c: counter c's count
        const uint32_t cnt = __builtin_amdgcn_sad_u8((uint32_t)tr, 0u, 0u) +
                             __builtin_amdgcn_sad_u8((uint32_t)(tr >> 32), 0u, 0u);
        if (w0 + wl < W) col_part[(size_t)chunk * W * 64 + (size_t)(w0 + wl) * 64 + lane] = cnt;
    }
}

__global__ void __launch_bounds__(kBlock) k_sum_columns(const uint32_t* __restrict__ part,
                                                        uint32_t chunks, uint32_t W,
                                                        uint32_t A, uint32_t* __restrict__ out) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= A) return;
    const auto p = gbl(part) + j;
    const size_t stride = (size_t)W * 64;
    uint32_t s = 0, c = 0;
    for (; c + 8 <= chunks; c += 8) {  // eight independent loads in flight
        uint32_t v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = p[(c + u) * stride];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += v[u];
    }
    for (; c < chunks; ++c) s += p[c * stride];
    gbl(out)[j] = s;
}

__global__ void __launch_bounds__(kBlock) k_sum_rows(const uint32_t* __restrict__ part,
                                                     uint32_t ngroups, uint32_t n,
                                                     uint32_t* __restrict__ out) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    uint32_t s = 0;
    for (uint32_t g = 0; g < ngroups; ++g) s += gbl(part)[(size_t)g * n + k];
    gbl(out)[k] = s;
}

// ---------------------------------------------------------------------------
// Fitness: GeneralFitness over individuals, z = genome(2i) + genome(2i+1)
// ---------------------------------------------------------------------------
// x = the first trait value (0 without traits), ss = the sum of squares of
// all trait values in trait order (the multivariate kind's only input).
__device__ inline double fitness_of(const Params& P, double x, double ss) {
    switch (P.fitness_kind) {
        case CLOTHO_FIT_GAUSSIAN: {
            const double ex = (x * x) * P.fit_inv;
            return P.fit_coeff * det_exp(-ex);
        }
        case CLOTHO_FIT_QUADRATIC: {
            double r = x * P.fit_inv;
            r = r * r;
            return r >= 1.0 ? 0.0 : 1.0 - r;
        }
        case CLOTHO_FIT_GAUSSIAN_FIXED:
            return det_exp(-((x * x) * P.fit_inv));
        case CLOTHO_FIT_MV_STABILIZING:
            return det_exp(-(ss * P.fit_inv));
        default:
            return 1.0;
    }
}

// kCurrent: the current population [sel] after its generation was committed
// (noise of the generation that produced it, gen - 1); else the offspring
// population [1 - sel] of the generation in progress.
template <bool kCurrent>
struct FitnessOp {
    static constexpr int kThreads = kBlock;
    __host__ __device__ static uint32_t blocks(const Params& P) { return (P.n + kBlock - 1) / kBlock; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        const auto st = gbl(B.state);
        const uint32_t which = kCurrent ? sel : 1 - sel;
        const auto traits = gbl(B.traits[which]);
        const auto fit = gbl(B.fitness[which]);
        const uint32_t i = blk * kBlock + threadIdx.x;
        if (i >= P.n) return;
        const uint32_t gen = kCurrent ? st->gen - 1 : st->gen;
        // trait values in trait order, kept as the first value and the sum of
        // squares (no per-thread array: indexed by a loop counter it lives in
        // scratch memory)
        double x0 = 0.0, ss = 0.0;
        for (uint32_t t = 0; t < P.traits; ++t) {
            const int64_t g = traits[(size_t)(2 * i) * P.traits + t] +
                              traits[(size_t)(2 * i + 1) * P.traits + t];
            double z = (double)g * P.eff_quantum;
            if (P.env_sigma > 0.0) {
                const Draw4 d = philox_draw(P.seed, gen, S_ENV, P.replicate, i, t);
                z = z + P.env_sigma * normal_ppf(uopen(d.x[0], d.x[1]));
            }
            if (t == 0) x0 = z;
            ss = ss + z * z;
        }
        fit[i] = fitness_of(P, x0, ss);
    }
};
using Fitness = FitnessOp<false>;
using FitnessCurrent = FitnessOp<true>;


// Synthetic standing variation: bit j of row r set when Philox word
// (seed, gen 0, stream 7, replicate, r, j/4)[j%4] < thr[j].
__global__ void __launch_bounds__(kBlock) k_synth_rows(Params P, uint64_t seed,
                                                       const uint32_t* thr, uint32_t alleles,
                                                       uint32_t words, uint64_t* rows) {
    const size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= (size_t)P.h * words) return;
    const uint32_t r = (uint32_t)(id / words), w = (uint32_t)(id % words);
    uint64_t x = 0;
    for (uint32_t q = 0; q < 16; ++q) {
        const uint32_t j0 = w * 64 + q * 4;
        if (j0 >= alleles) break;
        const Draw4 d = philox_draw(seed, 0, 7, P.replicate, r, j0 / 4);
        for (uint32_t k = 0; k < 4; ++k)
            if (j0 + k < alleles && d.x[k] < thr[j0 + k]) x |= (uint64_t)1 << (q * 4 + k);
    }
    rows[id] = x;
}

// End of a generation.  With intermediate generations
// (QS/engine_parallel_pipeline_batched.hpp:390-407) the free slots the new
// mutations took are dropped from the front of the free list until the last
// generation of the batch, whose analysis wrote a fresh list.
struct Commit {
    static constexpr int kThreads = 64;
    __host__ __device__ static uint32_t blocks(const Params&) { return 1u; }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t,
                                               uint32_t) {
        if (threadIdx.x) return;
        const auto st = gbl(B.state);
        st->alleles = st->all_size;
        st->wp = st->wc;
        st->gen = st->gen + 1;
        st->given = 0;
        if (P.inter_k) {
            if (st->inter < P.inter_k) {
                const uint32_t F = st->free_count, used = min(st->m, F);
                st->free_base = st->free_base + used;
                st->free_count = F - used;
                st->n_fixed = st->n_fixed > used ? st->n_fixed - used : 0u;
                st->inter = st->inter + 1;
            } else {
                st->free_base = 0;
                st->inter = 0;
            }
        }
    }
};

__global__ void __launch_bounds__(kBlock) k_any_nonzero(const int32_t* effects, size_t n,
                                                        uint32_t* flag_out) {
    __shared__ int s[kBlock];
    int any = 0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kBlock)
        any |= effects[i] != 0;
    s[threadIdx.x] = any;
    __syncthreads();
    for (int q = kBlock / 2; q > 0; q >>= 1) {
        if ((int)threadIdx.x < q) s[threadIdx.x] |= s[threadIdx.x + q];
        __syncthreads();
    }
    if (threadIdx.x == 0) flag_out[blockIdx.x] = s[0];
}

// Exact genome traits of population `sel` (i8::traits_i8_block); the effect
// digit planes are packed first (i8::pack_i8_block).
struct PackI8 {
    static constexpr int kThreads = 256;
    __host__ __device__ static uint32_t ng_cap(const Params& P) {
        return (P.w_cap + i8::kGroupWords - 1) / i8::kGroupWords;
    }
    __host__ __device__ static uint32_t blocks(const Params& P) {
        return P.traits ? i8::kernel_tiles(P.traits) * ng_cap(P) * 4 : 0u;  // 1024 items per (nt, G)
    }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t,
                                               uint32_t blk) {
        i8::pack_i8_block(B.effects, P.traits, i8::kernel_tiles(P.traits), &B.state->wc,
                          &B.state->all_size, ng_cap(P), B.digits, blk);
    }
};

// Up to RT row tiles of 16 rows per wave.  kEven: P.trait_blocks blocks
// share the row tiles evenly (i8::wave_tiles_even), the grid a multiple of
// the compute-unit count, and a wave one tile short runs the RT - 1 body;
// otherwise RT tiles per wave in block order, the grid following from Params
// alone (as the batched launches need).
template <int RT, int NT, bool kEven>
struct TraitsI8 {
    static constexpr int kThreads = 256;
    static constexpr uint32_t kRowTiles = RT;
    __host__ __device__ static uint32_t blocks(const Params& P) {
        if (!P.traits) return 0u;
        return kEven ? P.trait_blocks : (P.h + 64 * RT - 1) / (64 * RT);
    }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        const uint32_t wave = threadIdx.x >> 6;
        const i8::WaveTiles wt = kEven ? i8::wave_tiles_even(blk, wave, P.h, P.trait_blocks)
                                       : i8::wave_tiles_fixed(blk, wave, RT);
        i8::traits_i8_block<RT, NT, kEven>(B.rows[sel], P.h, &B.state->wc, P.traits, B.digits,
                                           B.traits[sel], wt);
    }
};

// Grid of an even traits launch over h rows: the fewest blocks that keep
// every wave at RT tiles or fewer, rounded up to a multiple of the compute
// units once there are more blocks than units; 0 when that grid exceeds one
// round of `resident` blocks per unit (the fixed-tile kernel is used then).
uint32_t even_trait_blocks(uint32_t h, uint32_t rt, uint32_t cus, uint32_t resident) {
    const uint32_t tiles = (h + 15u) / 16u;
    const uint32_t bmin = (tiles + 4u * rt - 1u) / (4u * rt);
    if (bmin <= cus) return bmin;
    const uint32_t b = (bmin + cus - 1u) / cus * cus;
    return b <= resident * cus ? b : 0u;
}

// Two independent ops in one launch: blocks [0, A::blocks) run A, the rest
// run Bo (neither reads what the other writes).  One launch instead of two
// saves the fixed cost of a dependent kernel in the sequence.
template <class A, class Bo>
struct Fused {
    static_assert(A::kThreads == Bo::kThreads, "fused ops share the block size");
    static constexpr int kThreads = A::kThreads;
    __host__ __device__ static uint32_t blocks(const Params& P) {
        return A::blocks(P) + Bo::blocks(P);
    }
    __device__ static __forceinline__ void run(const Params& P, const Buffers& B, uint32_t sel,
                                               uint32_t blk) {
        const uint32_t na = A::blocks(P);
        if (blk < na)
            A::run(P, B, sel, blk);
        else
            Bo::run(P, B, sel, blk - na);
    }
};

// ---------------------------------------------------------------------------
// Kernels.  Single population: one block per unit of Op::blocks, arguments by
// value.  Batched: a persistent grid walks the concatenated block ranges of
// all replicates (offs[r] = first virtual block of replicate r), with the
// replicate's Params and Buffers read from device memory; a barrier separates
// consecutive virtual blocks so that shared memory is reused safely.
// ---------------------------------------------------------------------------
template <class Op>
struct MinWaves {
    static constexpr int value = 1;
};
template <bool W, int K>
struct MinWaves<Recombine<W, K>> {
    static constexpr int value = 8;
};
// four traits waves per SIMD (<= 128 registers): the batched launch at 136
// registers held three 256-thread blocks per CU where its LDS allows four
template <>
struct MinWaves<TraitsI8<4, 1, false>> {
    static constexpr int value = 4;
};

#define CLOTHO_SINGLE(name, Op)                                                                 \
    __global__ void __launch_bounds__(Op::kThreads)                                            \
        __attribute__((amdgpu_waves_per_eu(MinWaves<Op>::value)))                               \
        name(Params P, Buffers B, uint32_t sel) {                                               \
        Op::run(P, B, sel, blockIdx.x);                                                         \
    }

CLOTHO_SINGLE(k_prepare, Prepare)
CLOTHO_SINGLE(k_sel_blockscan, SelBlockscan)
CLOTHO_SINGLE(k_parent_blockscan, ParentBlockscan)
using ParentScatter = ParentScatterOp<true>;
CLOTHO_SINGLE(k_parent_scatter, ParentScatter)
CLOTHO_SINGLE(k_word_sort, WordSort)
// the draw phase of a single population: counts and fitness statistics, then
// attributes and selection scan, then gametes and the per-word order
// statistics, each pair in one launch
using DrawCounts = Fused<MutCounts, FitnessStats>;
using DrawAttrs = Fused<MutAttrs, SelScan>;
using DrawGametes = Fused<GametesOp<kTotalsScan>, WordSort>;
CLOTHO_SINGLE(k_draw_counts, DrawCounts)
CLOTHO_SINGLE(k_draw_attrs, DrawAttrs)
CLOTHO_SINGLE(k_draw_gametes, DrawGametes)
using RecombineNarrow = Recombine<false, kSibsOrdered>;  // at most 64 events per gamete
using RecombineWide = Recombine<true, 2>;  // two sets of 128 event keys stay in registers
using BatchRecombineNarrow = Recombine<false, 1>;
using BatchRecombineWide = Recombine<true, 1>;
// The recombination's blocks are numbered so that each XCD runs a contiguous
// range of (row chunk, column tile) blocks, i.e. whole row chunks one tile
// after another: the chunk's gamete records and event keys, which every tile
// block of the chunk reads, then stay in that XCD's L2 (with the hardware's
// round-robin order every tile block of a chunk ran on another XCD and
// fetched them again).
#define CLOTHO_SINGLE_XCD(name, Op)                                                             \
    __global__ void __launch_bounds__(Op::kThreads)                                            \
        __attribute__((amdgpu_waves_per_eu(MinWaves<Op>::value)))                               \
        name(Params P, Buffers B, uint32_t sel) {                                               \
        Op::run(P, B, sel, xcd_contiguous(blockIdx.x, gridDim.x));                              \
    }
CLOTHO_SINGLE_XCD(k_recombine, RecombineNarrow)
CLOTHO_SINGLE_XCD(k_recombine_wide, RecombineWide)
#undef CLOTHO_SINGLE_XCD
CLOTHO_SINGLE(k_pack_i8_pop, PackI8)
using Traits4Even = TraitsI8<5, 1, true>;  // T <= 4, grid within one round of resident blocks
using Traits4 = TraitsI8<4, 1, false>;      // T <= 4
using Traits4Two = TraitsI8<2, 1, false>;   // T <= 4, two row tiles per wave (autotuner)
using Traits8Four = TraitsI8<4, 2, false>;  // T <= 8, four row tiles per wave (autotuner)
using Traits32Two = TraitsI8<2, 8, false>;  // T <= 32, two row tiles per wave (autotuner)
using Traits8 = TraitsI8<2, 2, false>;      // T <= 8
using Traits16 = TraitsI8<4, 4, false>;     // T <= 16 (four row tiles: 5% faster than two at H = 1e6)
using BatchTraits16 = TraitsI8<2, 4, false>;  // batched T <= 16: the rows per block of Traits8
using Traits32 = TraitsI8<1, 8, false>;     // T <= 32
CLOTHO_SINGLE(k_traits_i8_pop, Traits4Even)
CLOTHO_SINGLE(k_traits_i8_pop_fixed, Traits4)
CLOTHO_SINGLE(k_traits_i8_pop_t8, Traits8)
CLOTHO_SINGLE(k_traits_i8_pop_t16, Traits16)
CLOTHO_SINGLE(k_traits_i8_pop_t32, Traits32)
CLOTHO_SINGLE(k_traits_i8_pop_two, Traits4Two)
CLOTHO_SINGLE(k_traits_i8_pop_t8_four, Traits8Four)
CLOTHO_SINGLE(k_traits_i8_pop_t16_two, BatchTraits16)
CLOTHO_SINGLE(k_traits_i8_pop_t32_two, Traits32Two)
CLOTHO_SINGLE(k_free_reduce, FreeReduce)
CLOTHO_SINGLE(k_free_compact, FreeCompact)
CLOTHO_SINGLE(k_fitness, Fitness)
CLOTHO_SINGLE(k_fitness_current, FitnessCurrent)
CLOTHO_SINGLE(k_commit, Commit)
// more shared launches of a single-population generation: the effect digit
// planes (new effects are final after the draws) with the parent-order counts,
// and the free-list compaction with the offspring fitness
using DrawParents = Fused<ParentScan, PackI8>;
using CompactFitness = Fused<FreeCompact, Fitness>;
CLOTHO_SINGLE(k_draw_parents, DrawParents)
CLOTHO_SINGLE(k_compact_fitness, CompactFitness)
#undef CLOTHO_SINGLE

// Occupancy floor of a batched kernel: the recombination kernel needs two
// resident 1024-thread blocks per CU (8 waves per SIMD, <= 64 VGPRs) to keep
// enough parent loads in flight.
constexpr uint32_t kBatchOffsLds = 1024;  // replicate offsets kept in LDS up to this R

// Ops whose batched offsets count live blocks only (k_batch_live_offsets).
// Their block ranges at capacity cover the allele capacity, but the child
// rows are wc words wide this generation, so only the blocks of the first wc
// words do any work.  kind indexes the live offsets; block() maps the local
// live index u of a replicate to the Op's own block number.
template <class Op>
struct LiveBlocks {
    static constexpr int kind = -1;
};
template <bool W, int K>
struct LiveBlocks<Recombine<W, K>> {
    // live tiles lt = ceil(wc / 64) x row chunks: u -> tile u mod lt, chunk u / lt
    static constexpr int kind = kLiveRecombine;
    __host__ __device__ static uint32_t count(const Params& P, uint32_t wc) {
        return min((wc + 63) / 64, P.tiles) * P.chunks;
    }
    __device__ static uint32_t block(const Params& P, uint32_t wc, uint32_t u) {
        const uint32_t lt = (wc + 63) / 64;  // >= 1 when u exists
        return (u / lt) * P.tiles + u % lt;
    }
};
template <>
struct LiveBlocks<WordSort> {
    // kWaves words per block: the live blocks are the first ceil(wc / kWaves)
    static constexpr int kind = kLiveWordSort;
    __host__ __device__ static uint32_t count(const Params& P, uint32_t wc) {
        return min((wc + kWaves - 1) / kWaves, WordSort::blocks(P));
    }
    __device__ static uint32_t block(const Params&, uint32_t, uint32_t u) { return u; }
};
template <>
struct LiveBlocks<PackI8> {
    // four blocks per (trait tile, 16-word group); live groups NG = ceil(wc / 16)
    static constexpr int kind = kLivePack;
    __host__ __device__ static uint32_t count(const Params& P, uint32_t wc) {
        if (!P.traits) return 0u;
        const uint32_t ng = min((wc + i8::kGroupWords - 1) / i8::kGroupWords, PackI8::ng_cap(P));
        return i8::kernel_tiles(P.traits) * ng * 4u;
    }
    __device__ static uint32_t block(const Params& P, uint32_t wc, uint32_t u) {
        const uint32_t per = ((wc + i8::kGroupWords - 1) / i8::kGroupWords) * 4u;  // >= 4 here
        return (u / per) * PackI8::ng_cap(P) * 4u + u % per;
    }
};

// Batched ops whose persistent blocks take virtual blocks round robin: their
// virtual blocks differ in cost (event counts, live widths), and contiguous
// ranges measured slower for them (gamete draws 0.29 -> 0.41 ms on the
// 256-replicate workload).  The others take contiguous ranges (mutation
// counts 55 -> 38 us, fitness 59 -> 53 us).
template <class Op>
struct RoundRobin {
    static constexpr bool value = false;
};
template <bool W, int K>
struct RoundRobin<Recombine<W, K>> {
    static constexpr bool value = true;
};
template <int RT, int NT, bool E>
struct RoundRobin<TraitsI8<RT, NT, E>> {
    static constexpr bool value = true;
};
template <int K>
struct RoundRobin<GametesOp<K>> {
    static constexpr bool value = true;
};
template <>
struct RoundRobin<PackI8> {
    static constexpr bool value = true;
};
template <>
struct RoundRobin<WordSort> {
    static constexpr bool value = true;
};
template <>
struct RoundRobin<FreeReduce> {
    static constexpr bool value = true;
};
template <>
struct RoundRobin<FreeCompact> {
    static constexpr bool value = true;
};

// Live virtual blocks of the batched recombination, word sort and digit
// packing (LiveBlocks): one block scans each kind's live block counts over the
// replicates into out[kind][0 .. R] (exclusive, out[kind][R] = total) from the
// child widths the draws set, so the persistent blocks of those ops visit live
// blocks only.
__global__ void __launch_bounds__(1024) k_batch_live_offsets(const RepView* __restrict__ views,
                                                             uint32_t R, uint32_t* __restrict__ out) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_carry;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int kind = 0; kind < kLiveKinds; ++kind) {
        uint32_t* const o = out + (size_t)kind * (R + 1);
        if (tid == 0) s_carry = 0;
        __syncthreads();
        for (uint32_t base = 0; base < R; base += 1024) {
            const uint32_t r = base + tid;
            uint32_t n = 0;
            if (r < R) {
                const RepView& rv = views[r];
                const uint32_t wc = gbl(rv.B.state)->wc;
                n = kind == kLiveRecombine ? LiveBlocks<RecombineNarrow>::count(rv.P, wc)
                    : kind == kLiveWordSort ? LiveBlocks<WordSort>::count(rv.P, wc)
                                            : LiveBlocks<PackI8>::count(rv.P, wc);
            }
            uint32_t x = n;  // inclusive scan within the wave
            for (uint32_t d = 1; d < 64; d <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)x, d, 64);
                if (lane >= d) x += y;
            }
            if (lane == 63) s_wave[wave] = x;
            __syncthreads();
            if (wave == 0) {
                uint32_t t = lane < 16 ? s_wave[lane] : 0u;
                for (uint32_t d = 1; d < 16; d <<= 1) {
                    const uint32_t y = (uint32_t)__shfl_up((int)t, d, 64);
                    if (lane >= d) t += y;
                }
                if (lane < 16) s_wave[lane] = t;  // inclusive wave totals
            }
            __syncthreads();
            if (r < R) o[r] = s_carry + (wave ? s_wave[wave - 1] : 0u) + x - n;
            __syncthreads();
            if (tid == 0) s_carry += s_wave[15];
            __syncthreads();
        }
        if (tid == 0) o[R] = s_carry;
        __syncthreads();
    }
}

template <class Op>
__global__ void __launch_bounds__(Op::kThreads)
    __attribute__((amdgpu_waves_per_eu(MinWaves<Op>::value))) k_batched(const RepView* __restrict__ views,
                                                          const uint32_t* __restrict__ offs,
                                                          uint32_t R, uint32_t sel) {
    // the replicate offsets in LDS when they fit: every virtual block
    // searches them, and a search in global memory is a chain of dependent loads
    __shared__ uint32_t s_offs[kBatchOffsLds + 1];
    const bool in_lds = R <= kBatchOffsLds;
    if (in_lds)
        for (uint32_t i = threadIdx.x; i <= R; i += Op::kThreads) s_offs[i] = offs[i];
    __syncthreads();
    const uint32_t total = offs[R];
    auto off = [&](uint32_t r) { return in_lds ? s_offs[r] : offs[r]; };
    // Virtual blocks v, v + grid, ... (RoundRobin ops: their virtual blocks
    // differ in cost) or one contiguous range per block: consecutive virtual
    // blocks then mostly belong to one replicate, so the replicate index
    // advances by a step or two and the replicate's descriptor and state
    // reads hit the caches.
    uint32_t v = blockIdx.x, vend = total, step = gridDim.x;
    if constexpr (!RoundRobin<Op>::value) {
        const uint32_t per = (total + gridDim.x - 1) / gridDim.x;
        v = min(total, blockIdx.x * per);
        vend = min(total, v + per);
        step = 1;
    }
    uint32_t lo = 0;
    for (bool first = true; v < vend; v += step, first = false) {
        // replicate: the largest r < R with offs[r] <= v (empty ranges are skipped)
        if (RoundRobin<Op>::value || first) {
            uint32_t a = 0, hi = R;
            while (hi - a > 1) {
                const uint32_t mid = (a + hi) >> 1;
                if (off(mid) <= v)
                    a = mid;
                else
                    hi = mid;
            }
            lo = a;
        } else {
            while (lo + 1 < R && off(lo + 1) <= v) ++lo;
        }
        lo = __builtin_amdgcn_readfirstlane(lo);  // uniform: descriptor reads stay scalar
        const RepView& rv = views[lo];
        uint32_t blk = v - off(lo);
        if constexpr (LiveBlocks<Op>::kind >= 0)
            blk = LiveBlocks<Op>::block(rv.P, uniform_read(&gbl(rv.B.state)->wc), blk);
        Op::run(rv.P, rv.B, sel, blk);
        __syncthreads();
    }
}

}  // namespace

// ===========================================================================
// launch sequences
// ===========================================================================
#define LAUNCH_CHECK()                              \
    do {                                            \
        hipError_t e_ = hipGetLastError();          \
        if (e_ != hipSuccess) return e_;            \
    } while (0)

namespace {

template <class Op>
hipError_t launch_single(void (*kernel)(Params, Buffers, uint32_t), clotho_ctx* c, uint32_t sel) {
    const uint32_t nb = Op::blocks(c->p);
    if (!nb) return hipSuccess;
    hipLaunchKernelGGL(kernel, dim3(nb), dim3(Op::kThreads), 0, c->stream, c->p, c->b, sel);
    return hipGetLastError();
}

#define TRY_LAUNCH(expr)                     \
    do {                                     \
        const hipError_t e_ = (expr);        \
        if (e_ != hipSuccess) return e_;     \
    } while (0)

}  // namespace

// Draws of the generation, the per-word order statistics of the child columns
// (which need only the new mutations' keys) and the effect digit planes of
// the traits pass.
hipError_t launch_draw(clotho_ctx* c) {
    const uint32_t par = (uint32_t)c->cur;  // parent = current population
    TRY_LAUNCH(launch_single<DrawCounts>(k_draw_counts, c, par));
    TRY_LAUNCH(launch_single<Prepare>(k_prepare, c, par));
    TRY_LAUNCH(launch_single<DrawAttrs>(k_draw_attrs, c, par));
    // up to kTotalsLds block totals, the gamete and scatter blocks scan them
    // themselves; beyond, a one-block scan runs first
    const bool scan_apart = parent_blocks(c->p) > kTotalsLds;
    if (scan_apart) TRY_LAUNCH(launch_single<SelBlockscan>(k_sel_blockscan, c, par));
    TRY_LAUNCH(launch_single<DrawGametes>(k_draw_gametes, c, par));
    if (!c->b.by_parent) return launch_single<PackI8>(k_pack_i8_pop, c, par);
    TRY_LAUNCH(launch_single<DrawParents>(k_draw_parents, c, par));
    if (scan_apart) TRY_LAUNCH(launch_single<ParentBlockscan>(k_parent_blockscan, c, par));
    return launch_single<ParentScatter>(k_parent_scatter, c, par);
}

// Exact genome traits of population [which] (width st->wc, alleles st->all_size).
static hipError_t launch_traits_of(clotho_ctx* c, int which, bool pack) {
    if (!c->p.traits) return hipSuccess;
    const uint32_t sel = (uint32_t)which;
    if (pack) TRY_LAUNCH(launch_single<PackI8>(k_pack_i8_pop, c, sel));
    const uint32_t nt = (c->p.traits + 3) / 4;
    const uint32_t v = c->trait_variant;  // 0: default shape; the autotuner may pick another
    if (nt <= 1) {
        if (v == 2) return launch_single<Traits4>(k_traits_i8_pop_fixed, c, sel);
        if (v == 3) return launch_single<Traits4Two>(k_traits_i8_pop_two, c, sel);
        // one round of the even kernel where it fits, else the fixed tiles
        if (!c->trait_resident) TRY_LAUNCH(query_trait_residency(c));
        const uint32_t cus = (uint32_t)std::max(1, c->prop.multiProcessorCount);
        c->p.trait_blocks = even_trait_blocks(c->p.h, Traits4Even::kRowTiles, cus, c->trait_resident);
        if (c->p.trait_blocks) return launch_single<Traits4Even>(k_traits_i8_pop, c, sel);
        return launch_single<Traits4>(k_traits_i8_pop_fixed, c, sel);
    }
    if (nt <= 2) {
        if (v == 2) return launch_single<Traits8Four>(k_traits_i8_pop_t8_four, c, sel);
        return launch_single<Traits8>(k_traits_i8_pop_t8, c, sel);
    }
    if (nt <= 4) {
        if (v == 2) return launch_single<BatchTraits16>(k_traits_i8_pop_t16_two, c, sel);
        return launch_single<Traits16>(k_traits_i8_pop_t16, c, sel);
    }
    if (v == 2) return launch_single<Traits32Two>(k_traits_i8_pop_t32_two, c, sel);
    return launch_single<Traits32>(k_traits_i8_pop_t32, c, sel);
}

// Launch-shape candidates of the traits pass for the context's trait count
// (values of clotho_ctx::trait_variant; 1 is the default shape).
uint32_t trait_variants(const clotho_ctx* c) { return (c->p.traits + 3) / 4 <= 1 ? 3u : 2u; }

hipError_t launch_traits_current(clotho_ctx* c) { return launch_traits_of(c, c->cur, true); }

hipError_t query_trait_residency(clotho_ctx* c) {
    int per_cu = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(k_traits_i8_pop), Traits4Even::kThreads, 0);
    if (e != hipSuccess) return e;
    c->trait_resident = (uint32_t)std::max(1, per_cu);
    return hipSuccess;
}

hipError_t launch_word_sort(clotho_ctx* c) {
    return launch_single<WordSort>(k_word_sort, c, (uint32_t)c->cur);
}

hipError_t launch_recombine(clotho_ctx* c) {
    const uint32_t par = (uint32_t)c->cur;
    if (c->p.ev_cap > 64) return launch_single<RecombineWide>(k_recombine_wide, c, par);
    return launch_single<RecombineNarrow>(k_recombine, c, par);
}

// the digit planes were packed by the draw phase
hipError_t launch_traits(clotho_ctx* c) { return launch_traits_of(c, 1 - c->cur, false); }

hipError_t launch_free_space(clotho_ctx* c) {
    TRY_LAUNCH(launch_single<FreeReduce>(k_free_reduce, c, 0));
    return launch_single<FreeCompact>(k_free_compact, c, 0);
}

hipError_t launch_free_reduce(clotho_ctx* c) {
    return launch_single<FreeReduce>(k_free_reduce, c, 0);
}
hipError_t launch_compact_fitness(clotho_ctx* c) {
    return launch_single<CompactFitness>(k_compact_fitness, c, (uint32_t)c->cur);
}
hipError_t launch_fitness(clotho_ctx* c) {
    return launch_single<Fitness>(k_fitness, c, (uint32_t)c->cur);
}

hipError_t launch_fitness_current(clotho_ctx* c) {
    return launch_single<FitnessCurrent>(k_fitness_current, c, (uint32_t)c->cur);
}

hipError_t launch_commit(clotho_ctx* c) { return launch_single<Commit>(k_commit, c, 0); }

// After an upload into population [cur]: traits, free-space classification.
hipError_t launch_population_analysis(clotho_ctx* c) {
    const Params& P = c->p;
    Buffers& B = c->b;
    const int cur = c->cur;
    TRY_LAUNCH(launch_traits_of(c, cur, true));
    hipLaunchKernelGGL(k_column_reduce, dim3((P.w_cap + kBlock - 1) / kBlock, P.chunks),
                       dim3(kBlock), 0, c->stream, P, B.state, B.rows[cur], B.var_part,
                       B.fix_part);
    TRY_LAUNCH(hipGetLastError());
    return launch_free_space(c);
}

// ---------------------------------------------------------------------------
// Batched replicates: one persistent launch per kernel of the generation.
// ---------------------------------------------------------------------------
void batch_op_blocks(const Params& P, uint32_t* out) {
    out[kOpMutCounts] = MutCounts::blocks(P);
    out[kOpFitnessStats] = FitnessStats::blocks(P);
    out[kOpPrepare] = Prepare::blocks(P);
    out[kOpMutAttrs] = MutAttrs::blocks(P);
    out[kOpSelScan] = SelScan::blocks(P);
    out[kOpSelBlockscan] = SelBlockscan::blocks(P);
    out[kOpGametes] = Gametes::blocks(P);
    out[kOpWordSort] = WordSort::blocks(P);
    out[kOpRecombine] = RecombineNarrow::blocks(P);
    out[kOpPackI8] = PackI8::blocks(P);
    const uint32_t nt = (P.traits + 3) / 4;
    out[kOpTraits] = nt <= 1 ? Traits4::blocks(P)
                   : nt <= 4 ? Traits8::blocks(P)  // same rows per block as BatchTraits16
                             : Traits32::blocks(P);
    out[kOpFreeReduce] = FreeReduce::blocks(P);
    out[kOpFreeCompact] = FreeCompact::blocks(P);
    out[kOpFitness] = Fitness::blocks(P);
    out[kOpCommit] = Commit::blocks(P);
}

namespace {

template <class Op>
hipError_t launch_batched(const BatchLaunch& L, int op, uint32_t sel, hipStream_t s) {
    const uint32_t total = L.total[op];  // at capacity: the most live blocks
    if (!total) return hipSuccess;
    // blocks per compute unit.  A grid past residency starts its last blocks
    // late: for the traits blocks (four resident per CU) that measured faster
    // than a resident grid, since the late blocks take the virtual blocks
    // left over, but the recombination (two resident per CU, LDS-bound) lost
    // 7% at twice its resident grid, so its grid stays resident.
    uint32_t per_cu = std::max<uint32_t>(
        1u, (Op::kThreads >= 1024 ? 2u : 8u) * L.grid_scale_num / L.grid_scale_den);
    if constexpr (LiveBlocks<Op>::kind == kLiveRecombine) {
        static const uint32_t resident = [] {
            int n = 0;
            const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &n, reinterpret_cast<const void*>(k_batched<Op>), Op::kThreads, 0);
            return e == hipSuccess && n > 0 ? (uint32_t)n : 0u;
        }();
        if (resident) per_cu = std::min(per_cu, resident);
    }
    // An odd grid: round-robin virtual block v runs on block v mod grid, and
    // with replicate ranges that are multiples of a power of two (the column
    // tiles at capacity, before the live-block offsets) an even grid handed
    // every tile-0 block to the same quarter of the grid.
    uint32_t grid = std::min<uint32_t>(total, per_cu * L.cus);
    if (grid > 1 && !(grid & 1u)) --grid;
    const uint32_t* offs = LiveBlocks<Op>::kind >= 0
                               ? L.live_offs + (size_t)LiveBlocks<Op>::kind * (L.R + 1)
                               : L.offs + (size_t)op * (L.R + 1);
    hipLaunchKernelGGL(k_batched<Op>, dim3(grid), dim3(Op::kThreads), 0, s, L.views, offs, L.R, sel);
    return hipGetLastError();
}

// The live-block offsets of this generation (child widths are set by the
// draws' Prepare launch), for the word sort, recombination and digit packing.
hipError_t launch_live_offsets(const BatchLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_live_offsets, dim3(1), dim3(1024), 0, s, L.views, L.R, L.live_offs);
    return hipGetLastError();
}

hipError_t launch_batched_recombine(const BatchLaunch& L, uint32_t par, hipStream_t s) {
    if (L.wide_events) return launch_batched<BatchRecombineWide>(L, kOpRecombine, par, s);
    return launch_batched<BatchRecombineNarrow>(L, kOpRecombine, par, s);
}

}  // namespace

// The batch's recombination and traits over the last generation's gametes,
// parents [par], children into [1 - par] (the autotuner's timed replay).
hipError_t launch_batch_replay(const BatchLaunch& L, uint32_t par, hipStream_t s) {
    TRY_LAUNCH(launch_live_offsets(L, s));
    TRY_LAUNCH(launch_batched_recombine(L, par, s));
    if (!L.traits) return hipSuccess;
    const uint32_t chi = 1 - par;
    TRY_LAUNCH(launch_batched<PackI8>(L, kOpPackI8, chi, s));
    const uint32_t nt = (L.traits + 3) / 4;
    if (nt <= 1) return launch_batched<Traits4>(L, kOpTraits, chi, s);
    if (nt <= 2) return launch_batched<Traits8>(L, kOpTraits, chi, s);
    if (nt <= 4) return launch_batched<BatchTraits16>(L, kOpTraits, chi, s);
    return launch_batched<Traits32>(L, kOpTraits, chi, s);
}

hipError_t launch_batch_generation(const BatchLaunch& L, uint32_t par, hipStream_t s,
                                   hipEvent_t* marks) {
    auto mark = [&](int i) { return marks ? hipEventRecord(marks[i], s) : hipSuccess; };
    TRY_LAUNCH(mark(0));
    TRY_LAUNCH(launch_batched<MutCounts>(L, kOpMutCounts, par, s));
    TRY_LAUNCH(launch_batched<FitnessStats>(L, kOpFitnessStats, par, s));
    TRY_LAUNCH(launch_batched<Prepare>(L, kOpPrepare, par, s));
    TRY_LAUNCH(launch_batched<MutAttrs>(L, kOpMutAttrs, par, s));
    TRY_LAUNCH(launch_batched<SelScan>(L, kOpSelScan, par, s));
    TRY_LAUNCH(launch_batched<SelBlockscan>(L, kOpSelBlockscan, par, s));
    TRY_LAUNCH(launch_batched<Gametes>(L, kOpGametes, par, s));
    TRY_LAUNCH(mark(1));
    TRY_LAUNCH(launch_live_offsets(L, s));
    TRY_LAUNCH(launch_batched<WordSort>(L, kOpWordSort, par, s));
    TRY_LAUNCH(launch_batched_recombine(L, par, s));
    TRY_LAUNCH(mark(2));
    if (L.traits) {
        const uint32_t chi = 1 - par;
        TRY_LAUNCH(launch_batched<PackI8>(L, kOpPackI8, chi, s));
        const uint32_t nt = (L.traits + 3) / 4;
        if (nt <= 1)
            TRY_LAUNCH(launch_batched<Traits4>(L, kOpTraits, chi, s));
        else if (nt <= 2)
            TRY_LAUNCH(launch_batched<Traits8>(L, kOpTraits, chi, s));
        else if (nt <= 4)
            TRY_LAUNCH(launch_batched<BatchTraits16>(L, kOpTraits, chi, s));
        else
            TRY_LAUNCH(launch_batched<Traits32>(L, kOpTraits, chi, s));
    }
    TRY_LAUNCH(mark(3));
    TRY_LAUNCH(launch_batched<FreeReduce>(L, kOpFreeReduce, par, s));
    TRY_LAUNCH(launch_batched<FreeCompact>(L, kOpFreeCompact, par, s));
    TRY_LAUNCH(mark(4));
    TRY_LAUNCH(launch_batched<Fitness>(L, kOpFitness, par, s));
    TRY_LAUNCH(launch_batched<Commit>(L, kOpCommit, par, s));
    TRY_LAUNCH(mark(5));
    return hipSuccess;
}

// Rows of the compaction, kCompactWaves waves of c->compact_rows rows per
// step over (column tiles of c->compact_tile_words words at capacity) x
// compact_chunks blocks (default: about eight blocks per compute unit).
static uint32_t compact_tiles_of(const clotho_ctx* c) {
    return (c->p.w_cap + c->compact_tile_words - 1) / c->compact_tile_words;
}

// blocks of kCompactWaves adjacent tiles per row chunk
static uint32_t compact_tile_groups_of(const clotho_ctx* c) {
    return (compact_tiles_of(c) + kCompactWaves - 1) / kCompactWaves;
}

uint32_t compact_chunks_of(const clotho_ctx* c) {
    if (c->compact_chunks) return c->compact_chunks;
    const uint32_t cus = (uint32_t)std::max(1, c->prop.multiProcessorCount);
    const uint32_t groups = compact_tile_groups_of(c);
    const uint32_t want = (8u * cus + groups - 1) / groups;
    return std::max<uint32_t>(1, std::min<uint32_t>(want, (c->p.h + 255) / 256));
}

hipError_t launch_compact_rows(clotho_ctx* c, int src) {
    const Params& P = c->p;
    const uint32_t tw = c->compact_tile_words, tiles = compact_tiles_of(c);
    const dim3 grid(compact_tile_groups_of(c) * compact_chunks_of(c)), block(64 * kCompactWaves);
    const uint32_t s = (uint32_t)src;
    if (c->compact_rows == 4)
        hipLaunchKernelGGL(k_compact_rows<4>, grid, block, 0, c->stream, P, c->b, s, tw, tiles);
    else if (c->compact_rows == 16)
        hipLaunchKernelGGL(k_compact_rows<16>, grid, block, 0, c->stream, P, c->b, s, tw, tiles);
    else
        hipLaunchKernelGGL(k_compact_rows<8>, grid, block, 0, c->stream, P, c->b, s, tw, tiles);
    return hipGetLastError();
}

hipError_t launch_compact_plan(clotho_ctx* c) {
    hipLaunchKernelGGL(k_compact_plan, dim3(1), dim3(1024), 0, c->stream, c->p, c->b);
    return hipGetLastError();
}

hipError_t launch_compact(clotho_ctx* c, int src) {
    const Params& P = c->p;
    const uint32_t s = (uint32_t)src;
    hipLaunchKernelGGL(k_compact_plan, dim3(1), dim3(1024), 0, c->stream, P, c->b);
    TRY_LAUNCH(hipGetLastError());
    hipLaunchKernelGGL(k_compact_alleles, dim3((P.s_cap + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       c->stream, P, c->b);
    TRY_LAUNCH(hipGetLastError());
    TRY_LAUNCH(launch_compact_rows(c, src));
    const size_t n = std::max<size_t>(std::max<size_t>(P.s_cap, (size_t)P.h * P.traits),
                                      std::max<size_t>(P.n, P.w_cap));
    hipLaunchKernelGGL(k_compact_finish, dim3((uint32_t)((n + kBlock - 1) / kBlock)), dim3(kBlock),
                       0, c->stream, P, c->b, s);
    TRY_LAUNCH(hipGetLastError());
    hipLaunchKernelGGL(k_compact_state, dim3(1), dim3(1), 0, c->stream, c->b);
    return hipGetLastError();
}

hipError_t launch_allele_counts(clotho_ctx* c, const uint64_t* rows, uint32_t W, uint32_t A,
                                const uint32_t* d_sample, uint32_t n, uint32_t* d_cols,
                                uint32_t* d_rows, uint32_t* d_col_part, uint32_t* d_row_part,
                                uint32_t chunks) {
    const uint32_t ngroups = stat_word_tiles(W);
    const uint32_t per = n ? (n + chunks - 1) / chunks : 0u;
    if (per > kStatChunkRows) return hipErrorInvalidValue;  // counters hold 255 rows per half
    if (W && n) {
        if (d_sample)
            hipLaunchKernelGGL(k_allele_counts<true>, dim3(ngroups * chunks), dim3(64 * kStatWaves),
                               0, c->stream, rows, W, d_sample, n, per, d_col_part, d_row_part);
        else
            hipLaunchKernelGGL(k_allele_counts<false>, dim3(ngroups * chunks),
                               dim3(64 * kStatWaves), 0, c->stream, rows, W, d_sample, n, per,
                               d_col_part, d_row_part);
        TRY_LAUNCH(hipGetLastError());
    }
    if (A) {
        hipLaunchKernelGGL(k_sum_columns, dim3((A + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, d_col_part, W && n ? chunks : 0u, W, A, d_cols);
        TRY_LAUNCH(hipGetLastError());
    }
    if (n) {
        hipLaunchKernelGGL(k_sum_rows, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0,
                           c->stream, d_row_part, W ? ngroups : 0u, n, d_rows);
        TRY_LAUNCH(hipGetLastError());
    }
    return hipSuccess;
}

// Pairwise differences popcount(r_i xor r_j) over the W words of sample
// rows i < j (pairwise_difference<association_matrix>::eval with the
// row-grouped evaluator,
// include/clotho/data_spaces/analysis/pairwise_difference/association_matrix_pairwise_difference.hpp:78-102,
// row_grouped_pairwise_difference.hpp:25-55).  Block = a 16 x 16 tile of pairs
// (tiles on or above the diagonal, enumerated row by row), one pair per
// thread; 64-word chunks of the tile's 16 + 16 rows are staged in LDS with a
// 65-word row pitch, so the 16 rows a wave reads per word fall in distinct banks.
constexpr uint32_t kPairTile = 16;
constexpr uint32_t kPairChunk = 64;

__global__ void __launch_bounds__(256) k_pairwise(const uint64_t* __restrict__ rows_in, uint32_t W,
                                                  const uint32_t* __restrict__ sample, uint32_t n,
                                                  uint32_t tiles, uint32_t* __restrict__ out) {
    const auto rows = gbl(rows_in);
    __shared__ uint64_t sa[kPairTile][kPairChunk + 1], sb[kPairTile][kPairChunk + 1];
    uint32_t t = blockIdx.x, bi = 0;
    while (t >= tiles - bi) {  // uniform: at most `tiles` steps
        t -= tiles - bi;
        ++bi;
    }
    const uint32_t bj = bi + t;
    const uint32_t ti = threadIdx.x / kPairTile, tj = threadIdx.x % kPairTile;
    const uint32_t i = bi * kPairTile + ti, j = bj * kPairTile + tj;
    uint32_t acc = 0;
    for (uint32_t w0 = 0; w0 < W; w0 += kPairChunk) {
        for (uint32_t e = threadIdx.x; e < kPairTile * kPairChunk; e += 256) {
            const uint32_t r = e / kPairChunk, c = e % kPairChunk, w = w0 + c;
            const uint32_t ia = bi * kPairTile + r, ib = bj * kPairTile + r;
            sa[r][c] = (ia < n && w < W) ? rows[(size_t)sample[ia] * W + w] : 0;
            sb[r][c] = (ib < n && w < W) ? rows[(size_t)sample[ib] * W + w] : 0;
        }
        __syncthreads();
#pragma unroll 8
        for (uint32_t c = 0; c < kPairChunk; ++c) acc += (uint32_t)__popcll(sa[ti][c] ^ sb[tj][c]);
        __syncthreads();
    }
    if (i < j && j < n) out[(size_t)i * n - (size_t)i * (i + 1) / 2 + (j - i - 1)] = acc;
}

hipError_t launch_pairwise(clotho_ctx* c, const uint64_t* rows, uint32_t W,
                           const uint32_t* d_sample, uint32_t n, uint32_t* d_out) {
    const uint32_t tiles = (n + kPairTile - 1) / kPairTile;
    const uint32_t blocks = tiles * (tiles + 1) / 2;
    if (!blocks) return hipSuccess;
    hipLaunchKernelGGL(k_pairwise, dim3(blocks), dim3(256), 0, c->stream, rows, W, d_sample, n,
                       tiles, d_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Host-directed operations of the generic operator families: row re-stride,
// column clears, mutations at given rows and slots, per-gamete recombination
// of given rows.  None is on the generation's path.
// ---------------------------------------------------------------------------
namespace {

__global__ void __launch_bounds__(kBlock) k_restride(const uint64_t* __restrict__ src, uint32_t w0,
                                                     uint64_t* __restrict__ dst, uint32_t w1,
                                                     uint32_t h) {
    const size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= (size_t)h * w1) return;
    const uint32_t r = (uint32_t)(id / w1), w = (uint32_t)(id % w1);
    dst[id] = w < w0 ? src[(size_t)r * w0 + w] : 0;
}

__global__ void __launch_bounds__(kBlock) k_clear_columns(uint64_t* __restrict__ rows, uint32_t w,
                                                          uint32_t h,
                                                          const uint64_t* __restrict__ mask) {
    const size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= (size_t)h * w) return;
    rows[id] &= ~mask[id % w];
}

// Mutation k: allele attributes into slot slots[k], bit slots[k] of row
// rows[k] set.  Slots are distinct, so the bits are; two mutations of one
// row word meet in the atomic OR.
__global__ void __launch_bounds__(kBlock) k_add_mutations(
    uint64_t* __restrict__ pop, uint32_t w, const uint32_t* __restrict__ rows,
    const uint32_t* __restrict__ slots, const uint32_t* __restrict__ keys_in,
    const int32_t* __restrict__ eff_in, uint32_t n, uint32_t traits, uint32_t age,
    uint32_t* __restrict__ keys, int32_t* __restrict__ effects, uint32_t* __restrict__ ages) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t slot = slots[k];
    keys[slot] = keys_in[k];
    ages[slot] = age;
    for (uint32_t t = 0; t < traits; ++t)
        effects[(size_t)slot * traits + t] = eff_in[(size_t)k * traits + t];
    atomicOr((unsigned long long*)(pop + (size_t)rows[k] * w + slot / 64),
             (unsigned long long)1 << (slot % 64));
}

// One (gamete, word) per thread: mask bit j = odd #{events e <= key_j} over
// the heterozygous bits j (block_crossover with region_classifier's
// upper_bound), out = (top & ~mask) | (bottom & mask); words past the input
// width are zero.
__global__ void __launch_bounds__(kBlock) k_crossover_rows(
    const uint32_t* __restrict__ keys, const uint64_t* __restrict__ top,
    const uint64_t* __restrict__ bottom, uint32_t w, const uint32_t* __restrict__ nev,
    const uint32_t* __restrict__ ev, uint32_t stride, uint32_t gametes,
    uint64_t* __restrict__ out, uint32_t ow) {
    const size_t id = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (id >= (size_t)gametes * ow) return;
    const uint32_t g = (uint32_t)(id / ow), x = (uint32_t)(id % ow);
    if (x >= w) {
        out[id] = 0;
        return;
    }
    const uint64_t t = top[(size_t)g * w + x], b = bottom[(size_t)g * w + x];
    const uint32_t ne = nev[g];
    const uint32_t* e = ev + (size_t)g * stride;
    uint64_t het = t ^ b, mask = 0;
    while (het) {
        const uint32_t j = (uint32_t)__builtin_ctzll(het);
        het &= het - 1;
        const uint32_t key = keys[x * 64 + j];
        uint32_t lo = 0, hi = ne;
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (e[mid] <= key)
                lo = mid + 1;
            else
                hi = mid;
        }
        if (lo & 1) mask |= (uint64_t)1 << j;
    }
    out[id] = (t & ~mask) | (b & mask);
}

uint32_t grid_of(size_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

// analyze_free_indices over reduced column masks
// (include/clotho/data_spaces/free_space/free_space_details.hpp:359-408):
// fixed = set bits of the AND mask, lost = set bits of the complemented OR
// mask, both below M, fixed ascending then lost ascending.  One block walks
// the words 256 at a time, a word's counts (fixed << 32 | lost) scanned
// across the block and carried over the chunks.
__global__ void __launch_bounds__(kBlock) k_classify_columns(
    const uint64_t* __restrict__ fix_and, const uint64_t* __restrict__ var_or, uint32_t words,
    uint32_t M, uint32_t* __restrict__ fixed, uint32_t* __restrict__ lost,
    uint32_t* __restrict__ counts) {
    __shared__ uint64_t lds[kWaves];
    uint64_t carry = 0;
    for (uint32_t base = 0; base < words; base += kBlock) {
        const uint32_t w = base + threadIdx.x;
        uint64_t fx = 0, ls = 0;
        if (w < words) {
            const uint64_t valid = (uint64_t)w * 64 >= M       ? 0
                                 : (uint64_t)w * 64 + 64 <= M ? ~(uint64_t)0
                                                              : (((uint64_t)1 << (M - w * 64)) - 1);
            fx = fix_and[w] & valid;
            ls = ~var_or[w] & valid;
        }
        const uint64_t c = ((uint64_t)__popcll(fx) << 32) | (uint64_t)__popcll(ls);
        uint64_t total;
        const uint64_t ex = carry + block_inclusive_scan<uint64_t>(c, lds, total) - c;
        uint32_t f = (uint32_t)(ex >> 32), l = (uint32_t)ex;
        while (fx) {
            fixed[f++] = w * 64 + (uint32_t)__builtin_ctzll(fx);
            fx &= fx - 1;
        }
        while (ls) {
            lost[l++] = w * 64 + (uint32_t)__builtin_ctzll(ls);
            ls &= ls - 1;
        }
        carry += total;
    }
    if (threadIdx.x == 0) {
        counts[0] = (uint32_t)(carry >> 32);
        counts[1] = (uint32_t)carry;
    }
}

}  // namespace

hipError_t launch_classify_columns(const uint64_t* d_fix, const uint64_t* d_var, uint32_t words,
                                   uint32_t M, uint32_t* d_fixed, uint32_t* d_lost,
                                   uint32_t* d_counts, hipStream_t s) {
    hipLaunchKernelGGL(k_classify_columns, dim3(1), dim3(kBlock), 0, s, d_fix, d_var, words, M,
                       d_fixed, d_lost, d_counts);
    return hipGetLastError();
}

hipError_t launch_restride(clotho_ctx* c, int src, uint32_t w0, uint32_t w1) {
    const size_t n = (size_t)c->p.h * w1;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_restride, dim3(grid_of(n)), dim3(kBlock), 0, c->stream, c->b.rows[src],
                       w0, c->b.rows[1 - src], w1, c->p.h);
    return hipGetLastError();
}

hipError_t launch_clear_columns(clotho_ctx* c, int which, uint32_t w, const uint64_t* d_mask) {
    const size_t n = (size_t)c->p.h * w;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_clear_columns, dim3(grid_of(n)), dim3(kBlock), 0, c->stream,
                       c->b.rows[which], w, c->p.h, d_mask);
    return hipGetLastError();
}

hipError_t launch_add_mutations(clotho_ctx* c, int which, uint32_t w, const uint32_t* d_rows,
                                const uint32_t* d_slots, const uint32_t* d_keys,
                                const int32_t* d_eff, uint32_t n, uint32_t age) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_add_mutations, dim3(grid_of(n)), dim3(kBlock), 0, c->stream,
                       c->b.rows[which], w, d_rows, d_slots, d_keys, d_eff, n, c->p.traits, age,
                       c->b.keys, c->b.effects, c->b.ages);
    return hipGetLastError();
}

hipError_t launch_crossover_rows(clotho_ctx* c, const uint64_t* d_top, const uint64_t* d_bottom,
                                 uint32_t w, const uint32_t* d_nev, const uint32_t* d_ev,
                                 uint32_t stride, uint32_t gametes, uint64_t* d_out, uint32_t ow) {
    const size_t n = (size_t)gametes * ow;
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_crossover_rows, dim3(grid_of(n)), dim3(kBlock), 0, c->stream, c->b.keys,
                       d_top, d_bottom, w, d_nev, d_ev, stride, gametes, d_out, ow);
    return hipGetLastError();
}

hipError_t launch_all_neutral(clotho_ctx* c, uint32_t* d_flag) {
    const size_t n = (size_t)c->p.s_cap * c->p.traits;
    hipLaunchKernelGGL(k_any_nonzero, dim3(64), dim3(kBlock), 0, c->stream, c->b.effects, n,
                       d_flag);
    LAUNCH_CHECK();
    return hipSuccess;
}

hipError_t launch_synth_rows(clotho_ctx* c, uint64_t seed, const uint32_t* d_thr, uint32_t alleles,
                             uint32_t words) {
    const size_t n = (size_t)c->p.h * words;
    const uint32_t blocks = (uint32_t)((n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_synth_rows, dim3(blocks), dim3(kBlock), 0, c->stream, c->p, seed, d_thr,
                       alleles, words, c->b.rows[c->cur]);
    LAUNCH_CHECK();
    return hipSuccess;
}

}  // namespace clotho
