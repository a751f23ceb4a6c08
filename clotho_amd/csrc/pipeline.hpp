// pipeline.hpp — device-resident state and launch interface of the
// per-generation offspring pipeline (one population pair per context).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/clotho_amd.h"

namespace clotho {

constexpr int kMaxCdf = 128;      // Poisson inversion table capacity
constexpr int kMaxTraits = 32;    // trait capacity of the fused paths
constexpr int kScanBlock = 256;   // items per block of every two-level scan
// |quantised effect| bound: four balanced base-256 digits represent it exactly
constexpr double kEffectLimit = 1073741823.0;  // 2^30 - 1
// allele-slot capacity limit: int8 digit sums over a row stay within int32
constexpr uint32_t kMaxAlleleSlots = 1u << 24;
// words per column tile of the population-statistics pass (row partials are
// kept per tile), its waves per block, and the most rows of one of its row
// chunks (4 waves x 2 half-waves x 255 rows: its 8-plane bit-sliced counters
// count to 255)
constexpr uint32_t kStatTileWords = 64;
constexpr uint32_t kStatBlockWaves = 4;
constexpr uint32_t kStatChunkRows = 2040;
__host__ __device__ inline uint32_t stat_word_tiles(uint32_t words) {
    return (words + kStatTileWords - 1) / kStatTileWords;
}
// row chunks of the statistics pass over n rows of `words` words: at most
// 2048 blocks (two rounds of four resident blocks per compute unit: a few
// blocks past a round would run alone), each chunk at most kStatChunkRows
// rows and at least one row
inline uint32_t stat_chunks(uint32_t words, uint32_t n) {
    const uint32_t tiles = stat_word_tiles(words) ? stat_word_tiles(words) : 1u;
    uint32_t chunks = 2048u / tiles;
    if (chunks < 1u) chunks = 1u;
    const uint32_t need = (n + kStatChunkRows - 1) / kStatChunkRows;
    if (chunks < need) chunks = need;
    const uint32_t most = n ? n : 1u;
    return chunks < most ? chunks : most;
}

// The per-word order-statistic tables of the recombination (sorted keys,
// bucket counts, prefix masks) are stored tile-major: the table of a 64-word
// column tile is one contiguous [row][64 words] image, the layout a
// recombination block keeps in LDS, so a block fills its LDS copy with
// coalesced reads (a word-major table puts each lane of a read on its own
// cache line).  Entry k of word w, and prefix mask c of word w:
__host__ __device__ inline size_t table_at(uint32_t w, uint32_t k) {
    return (size_t)(w >> 6) * (64u * 64u) + k * 64u + (w & 63u);
}
__host__ __device__ inline size_t prefix_at(uint32_t w, uint32_t c) {
    return (size_t)(w >> 6) * (65u * 64u) + c * 64u + (w & 63u);
}
// elements of the sorted / bucket tables and of the prefix-mask table
inline size_t table_elems(uint32_t w_cap) { return (size_t)((w_cap + 63) / 64) * 64u * 64u; }
inline size_t prefix_elems(uint32_t w_cap) { return (size_t)((w_cap + 63) / 64) * 65u * 64u; }

// Scalars of the generation, kept in device memory so that a captured
// hipGraph replays successive generations without host involvement.
struct GenState {
    uint32_t gen;          // stream generation counter (reference m_generation)
    uint32_t alleles;      // allele-table size of the parent population (A)
    uint32_t wp;           // parent words per row
    uint32_t free_count;   // F = fixed + lost of the parent
    uint32_t n_fixed;      // fixed part of the free list
    uint32_t m;            // new mutations this generation
    uint32_t all_size;     // allele-table size of the child
    uint32_t wc;           // child words per row
    uint32_t fit_constant; // parent fitness all equal -> uniform selection
    int32_t sel_shift;     // fitness quantisation exponent
    uint32_t error;        // sticky: 1 = allele capacity exceeded
    uint32_t error_gen;    // generation of the first error
    uint64_t sel_total;    // sum of quantised fitness
    uint32_t kept;         // variable columns kept by the last compaction plan
    uint32_t free_base;    // free-list entries used by earlier intermediate generations
    uint32_t inter;        // position in the intermediate batch, 0 .. inter_k
    uint32_t given;        // 1: the next draw takes Buffers::given_parents (one generation)
};

// Kernel-side constants that do not change between generations.
struct Params {
    uint64_t seed;
    uint32_t replicate;
    uint32_t n;            // individuals of the current population (the child
                           // while a generation is produced)
    uint32_t h;            // haplotypes = 2n
    uint32_t n_par;        // individuals of the parent population of the
                           // generation being produced
    uint32_t n_cap;        // individual capacity of the buffers
    uint32_t inter_k;      // intermediate generations between free-space analyses
    uint32_t traits;       // T
    uint32_t s_cap;        // allele capacity
    uint32_t w_cap;        // word capacity per row
    uint32_t ev_cap;       // event slots per gamete
    uint32_t n_cdf_rho, n_cdf_mu;
    double bias, neutral_p, trait_mean, trait_sigma, env_sigma;
    double eff_scale;      // 2^-q
    double eff_quantum;    // 2^q
    uint32_t fitness_kind;
    double fit_coeff, fit_inv;
    uint32_t chunks;       // row chunks of the recombination grid
    uint32_t tiles;        // 64-word column tiles at capacity
    uint32_t mut_blocks;   // scan blocks over gametes
    uint32_t word_blocks;  // 64-word free-space blocks at capacity
    uint32_t max_chunks;   // row-chunk capacity of the partial buffers
    uint32_t trait_blocks; // grid of the single-population traits launch
};

// Device buffers.  Index [c] is the population that is current after the
// last generation ("child"); the other one receives the next generation.
struct Buffers {
    uint64_t* rows[2];        // [h][w] dense, row stride = that population's W
    int64_t* traits[2];       // [h][T] quantised genome trait sums
    double* fitness[2];       // [n]
    uint32_t* keys;           // [s_cap] position * 2^32
    int32_t* effects;         // [s_cap][T] quantised effects
    uint32_t* ages;           // [s_cap]
    uint32_t* free_list;      // [s_cap] fixed ascending then lost ascending; entries
                              // from free_base on are unused
    uint64_t* fixed_bits;     // [w_cap] AND of all rows, idx < all_size
    // per-word order statistics in tile-major order (table_at / prefix_at)
    uint32_t* sorted;         // [tiles][64 k][64 w] position keys of each word, ascending
    uint64_t* pm;             // [tiles][65 c][64 w] bits of the c smallest keys of each word
    uint32_t* bucket;         // [tiles][64 q][64 w] per word: #{keys < b<<24}, four b per dword
    uint4* digits;            // [ceil(T/4)][w_cap][64] int8 digit planes of the effects
    uint32_t* mut_count;      // [h]
    uint32_t* mut_local;      // [h] exclusive scan inside a block
    uint32_t* mut_bsum;       // [mut_blocks] block totals -> exclusive offsets
    uint32_t* top;            // [h] parent row used as top strand
    uint32_t* bottom;         // [h]
    uint32_t* n_events;       // [h]
    uint32_t* events;         // [h][ev_cap] event keys
    // Gamete records in parent order (the children of one parent individual
    // adjacent): {child row, top row, n_events | mutations << 8, first
    // mutation}.  The recombination loads a parent row pair once for up to
    // kSibs of its children; nullptr = child order (the batched launches).
    uint4* by_parent;         // [h]
    uint32_t* parent_cnt;     // [n] children per parent -> exclusive in-block offsets
    uint32_t* prank;          // [h] rank of the gamete among its parent's picks
    uint32_t* parent_bsum;    // [ceil(n/256)] block totals -> exclusive offsets
    uint32_t* given_parents;  // [h] parent individual of each gamete when GenState::given
    uint64_t* var_part;       // [chunks][w_cap] OR partials
    uint64_t* fix_part;       // [chunks][w_cap] AND partials
    uint32_t* word_cnt;       // [w_cap] (fixed count << 16 | lost count) per word
    uint32_t* word_bsum;      // [2*word_blocks] block totals of fixed / lost counts
                              // (partials are sized for max_chunks rows)
    double* cum_local;        // [n] quantised fitness inclusive scan inside a block
    double* cum_block;        // [sel blocks] inclusive scan of block totals
    double* fit_stats;        // [2*sel blocks] per-block max, all-equal flag
    double* cdf_rho;          // [kMaxCdf]
    double* cdf_mu;           // [kMaxCdf]
    GenState* state;
};

// One replicate as the batched kernels see it (device memory, read-only).
struct RepView {
    Params P;
    Buffers B;
};

// Kernels of one generation, in launch order; the batched launch of kernel k
// walks offs[k][0 .. R] (first virtual block of each replicate).
enum BatchOp : int {
    kOpMutCounts,
    kOpFitnessStats,
    kOpPrepare,
    kOpMutAttrs,
    kOpSelScan,
    kOpSelBlockscan,
    kOpGametes,
    kOpWordSort,
    kOpRecombine,
    kOpPackI8,
    kOpTraits,
    kOpFreeReduce,
    kOpFreeCompact,
    kOpFitness,
    kOpCommit,
    kOpCount
};

// Batched ops whose virtual blocks are counted from the child width each
// generation (live blocks), and their rows of BatchLaunch::live_offs.
constexpr int kLiveRecombine = 0, kLiveWordSort = 1, kLivePack = 2, kLiveKinds = 3;

// Row chunks of a replicate's batched recombination: its own launch shape's
// chunk count divided by this (two measured best of 1, 2, 4, 8 on the
// 256-replicate workload at generation 200: 1.12 against 1.20 ms).
constexpr uint32_t kBatchChunkDiv = 2;

struct BatchLaunch {
    const RepView* views = nullptr;  // device [R]
    const uint32_t* offs = nullptr;  // device [kOpCount][R + 1]
    uint32_t* live_offs = nullptr;   // device [kLiveKinds][R + 1]: live blocks of three ops
    uint32_t R = 0;
    uint32_t total[kOpCount] = {};   // offs[k][R]
    uint32_t traits = 0;             // trait count shared by every replicate
    bool wide_events = false;        // some replicate keeps more than 64 events per gamete
    uint32_t cus = 1;                // compute units of the device
    // persistent grid per compute unit: the default (2 blocks of 1024
    // threads, 8 of 256) scaled by num / den (clotho_batch_autotune)
    uint32_t grid_scale_num = 1, grid_scale_den = 1;
};

struct Timing {
    bool enabled = false;
    double ms[6] = {0, 0, 0, 0, 0, 0};
    int generations = 0;
};

// Events of a timed section, destroyed on every return path.
struct EventSet {
    std::vector<hipEvent_t> ev;
    EventSet() = default;
    EventSet(const EventSet&) = delete;
    EventSet& operator=(const EventSet&) = delete;
    ~EventSet() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    hipError_t add(hipEvent_t* out) {
        *out = nullptr;
        const hipError_t r = hipEventCreate(out);
        if (r == hipSuccess) ev.push_back(*out);
        return r;
    }
};

struct GraphPair {
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    bool valid = false;
};

}  // namespace clotho

struct clotho_ctx {
    clotho_config cfg;
    int device = 0;
    hipStream_t stream = nullptr;
    clotho::Params p;
    clotho::Buffers b;
    int cur = 0;                // index of the current (child) population
    bool has_population = false;
    clotho::Timing timing;
    clotho::GraphPair graphs;
    std::string err;
    hipDeviceProp_t prop;
    size_t device_bytes = 0;    // sum of the context's device allocations
    uint32_t next_n = 0;        // offspring individuals of following generations (0: keep)
    uint32_t trait_resident = 0;  // resident blocks per unit of the even traits kernel (0: unknown)
    // launch shapes the autotuner may change (0 / defaults: the built-in choice)
    uint32_t trait_variant = 0;   // traits-pass kernel shape, 1 .. trait_variants()
    uint32_t compact_rows = 8;    // rows per wave step of the column compaction: 4, 8, 16
    uint32_t compact_chunks = 0;  // row chunks of the compaction grid (0: about 8 blocks per CU)
    uint32_t compact_tile_words = 64;  // compacted words per compaction wave: 48, 56, 60 or 64
};

namespace clotho {
// Launch sequences (pipeline.hip).
hipError_t launch_draw(clotho_ctx* c);       // draws + per-word order statistics
hipError_t launch_word_sort(clotho_ctx* c);  // order statistics alone (autotuner replays)
hipError_t launch_recombine(clotho_ctx* c);
hipError_t launch_traits(clotho_ctx* c);
hipError_t launch_free_space(clotho_ctx* c);
hipError_t launch_fitness(clotho_ctx* c);
// fitness of the current population from its traits (noise of generation gen - 1)
hipError_t launch_fitness_current(clotho_ctx* c);
// rows of population [src] re-strided from w0 to w1 >= w0 words into [1 - src]
hipError_t launch_restride(clotho_ctx* c, int src, uint32_t w0, uint32_t w1);
// clear the columns of `mask` (w words) in every row of population [which]
hipError_t launch_clear_columns(clotho_ctx* c, int which, uint32_t w, const uint64_t* d_mask);
// host-directed mutations into population [which] (width w): mutation k sets
// slot slots[k] in row rows[k] and writes the slot's key, effects and age
hipError_t launch_add_mutations(clotho_ctx* c, int which, uint32_t w, const uint32_t* d_rows,
                                const uint32_t* d_slots, const uint32_t* d_keys,
                                const int32_t* d_eff, uint32_t n, uint32_t age);
// ordered fixed / lost slots of reduced column masks (AND, OR) below M;
// counts = {fixed, lost}
hipError_t launch_classify_columns(const uint64_t* d_fix, const uint64_t* d_var, uint32_t words,
                                   uint32_t M, uint32_t* d_fixed, uint32_t* d_lost,
                                   uint32_t* d_counts, hipStream_t s);
// per-gamete recombination of rows against the allele keys (sorted events per gamete)
hipError_t launch_crossover_rows(clotho_ctx* c, const uint64_t* d_top, const uint64_t* d_bottom,
                                 uint32_t w, const uint32_t* d_nev, const uint32_t* d_ev,
                                 uint32_t stride, uint32_t gametes, uint64_t* d_out, uint32_t ow);
// the generation's late launches: the reduction alone, then the free-list
// compaction and the offspring fitness in one launch
hipError_t launch_free_reduce(clotho_ctx* c);
hipError_t launch_compact_fitness(clotho_ctx* c);
hipError_t launch_commit(clotho_ctx* c);
hipError_t launch_population_analysis(clotho_ctx* c);  // after an upload
// traits pass over the current population (digit planes repacked); the
// candidates of its launch shape; the compaction's plan and row pass alone
// (the autotuner times them into the spare population buffer)
hipError_t launch_traits_current(clotho_ctx* c);
uint32_t trait_variants(const clotho_ctx* c);
hipError_t launch_compact_plan(clotho_ctx* c);
hipError_t launch_compact_rows(clotho_ctx* c, int src);
uint32_t compact_chunks_of(const clotho_ctx* c);
hipError_t query_trait_residency(clotho_ctx* c);  // sets c->trait_resident
// Column compaction of population [src] into [1 - src] (fixed and lost
// columns removed, slots renumbered in order); the caller flips cur.
hipError_t launch_compact(clotho_ctx* c, int src);
// Column and row margins of a row sample (d_sample == nullptr: rows 0..n-1);
// partial buffers: d_col_part [chunks][W*64], d_row_part [stat_word_tiles(W)][n].
hipError_t launch_allele_counts(clotho_ctx* c, const uint64_t* rows, uint32_t W, uint32_t A,
                                const uint32_t* d_sample, uint32_t n, uint32_t* d_cols,
                                uint32_t* d_rows, uint32_t* d_col_part, uint32_t* d_row_part,
                                uint32_t chunks);
// popcount(row_a xor row_b) for the sample pairs a < b, packed row by row
// (n (n - 1) / 2 values).
hipError_t launch_pairwise(clotho_ctx* c, const uint64_t* rows, uint32_t W,
                           const uint32_t* d_sample, uint32_t n, uint32_t* d_out);
hipError_t launch_all_neutral(clotho_ctx* c, uint32_t* d_flag);
hipError_t launch_synth_rows(clotho_ctx* c, uint64_t seed, const uint32_t* d_thr, uint32_t alleles,
                             uint32_t words);
// Context helpers (capi.hip).
void destroy_graphs(clotho_ctx* c);
int read_state(clotho_ctx* c, GenState* st);
int write_state(clotho_ctx* c, const GenState& st);
int capture_graphs(clotho_ctx* c);  // one generation per parity, on the context stream
void begin_generation(clotho_ctx* c);  // parent = current, offspring at the requested size
void end_generation(clotho_ctx* c);    // offspring population becomes current
int after_generation(const clotho_ctx* c, int cur);  // population index after a generation
// Batched replicates (all kernels of one generation for every replicate).
void batch_op_blocks(const Params& P, uint32_t* out);  // kOpCount block counts
hipError_t launch_batch_generation(const BatchLaunch& L, uint32_t par, hipStream_t s,
                                   hipEvent_t* marks);  // marks: 6 events or null
hipError_t launch_batch_replay(const BatchLaunch& L, uint32_t par, hipStream_t s);
}  // namespace clotho
