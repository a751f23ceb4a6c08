/*
 * clotho_amd.h — C ABI of the MI355X offspring pipeline.
 *
 * One context owns one population pair (parent/child), its allele table, the
 * free-space classification and the fitness vector, all resident in HBM of one
 * GPU.  Every entry point is plain C: status codes (0 = ok, <0 = error, text in
 * clotho_last_error), caller-owned host buffers, context-owned device buffers,
 * work ordered on the context's single HIP stream, one host thread per context.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference root; DS/ = include/clotho/data_spaces/, QS/ = examples/QTLSim/):
 *
 *   clotho_create / clotho_destroy
 *       Engine<..., parallel_pipeline>::Engine / ~Engine / init
 *       (QS/engine_parallel_pipeline.hpp:109-170, 306-323)
 *   clotho_generation / clotho_run
 *       Engine<..., parallel_pipeline>::simulate
 *       (QS/engine_parallel_pipeline.hpp:172-239)
 *   clotho_phase_draw
 *       mutation_allocator::allocate + generate_child_mutations +
 *       make_selection_generator + the per-gamete draws of common_crossover
 *       (DS/mutation/mutation_allocator.hpp:41-47;
 *        QS/engine_parallel_pipeline.hpp:327-367;
 *        DS/selection/selection_generator_factory.hpp:33-67;
 *        DS/crossover/common_crossover_row_block_alignment.hpp:53-69)
 *   clotho_phase_recombine_mutate
 *       removeFixedAlleles + offspring_generator::{crossover,mutate,phenotype,fixed}
 *       (QS/engine_parallel_pipeline.hpp:409-426;
 *        DS/offspring_generator/offspring_generator_population_space_row_alignment.hpp:84-260;
 *        DS/crossover/block_crossover.hpp:46-63)
 *   clotho_phase_traits
 *       offspring_generator::phenotype -> common_phenotype_accumulator::operator()
 *       + population_space::updateGenomeWeights
 *       (DS/offspring_generator/offspring_generator_population_space_row_alignment.hpp:234-251;
 *        DS/phenotype_evaluator/common_phenotype_accumulator_row_block_alignment.hpp:47-66)
 *   clotho_phase_free_space
 *       free_space_accumulator_mt::finalize -> free_space_details::analyze_free_indices
 *       (DS/free_space/free_space_accumulator.hpp:142-151;
 *        DS/free_space/free_space_details.hpp:359-408)
 *   clotho_phase_fitness
 *       GeneralFitness::operator()(population_space<row_block_alignment>*, s, e)
 *       (DS/fitness/general_fitness.hpp:148-177) with the ifitness metrics of
 *       src/clotho/fitness/{normal,quadratic,constant}_fitness_metric.cpp
 *   clotho_upload_* / clotho_download_*
 *       population_space::begin_genome/end_genome, AlleleSpace::getPositions,
 *       trait_space_vector, free_space_details::{fixed,lost}_{begin,end},
 *       GeneralFitness::begin/end (state access used by the reference's
 *       state_getter, QS/engine_parallel_pipeline.hpp:488-521)
 *   clotho_trait_contract
 *       common_phenotype_accumulator::operator() over a whole population
 *       (DS/phenotype_evaluator/common_phenotype_accumulator_row_block_alignment.hpp:47-66)
 *       in three arithmetic modes.
 */
#ifndef CLOTHO_AMD_H_
#define CLOTHO_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLOTHO_OK 0
#define CLOTHO_ERR_ARG (-1)      /* invalid argument or configuration            */
#define CLOTHO_ERR_HIP (-2)      /* a HIP runtime call failed                    */
#define CLOTHO_ERR_CAPACITY (-3) /* population or allele table exceeds capacity  */
#define CLOTHO_ERR_STATE (-4)    /* call out of order (e.g. no population)      */

/* Fitness metric kinds (fitness.name of the reference's fitness_toolkit). */
#define CLOTHO_FIT_CONSTANT 0       /* "constant": 1.0                                   */
#define CLOTHO_FIT_GAUSSIAN 1       /* "gaussian": (1/(s*sqrt(2pi))) exp(-x0^2/(2s^2)),
                                       s = sqrt(4 N mu), first trait only                */
#define CLOTHO_FIT_QUADRATIC 2      /* "quadratic": max(0, 1-(x0/s)^2), s = scale*sqrt(4 N mu) */
#define CLOTHO_FIT_GAUSSIAN_FIXED 3 /* exp(-x0^2/(2 scale^2)); no coefficient            */
#define CLOTHO_FIT_MV_STABILIZING 4 /* exp(-|x|^2/(2 scale^2)) over all traits           */

/* Trait contraction arithmetic for clotho_trait_contract. */
#define CLOTHO_TRAITS_PLAIN_FP32 0 /* dense bit x effect FMAs, FP32 (lane = haplotype)   */
#define CLOTHO_TRAITS_MFMA_BF16 1  /* bits -> BF16, effects split into BF16 terms, MFMA */
#define CLOTHO_TRAITS_EXACT_FP64 2 /* set-bit walk, FP64 sums                          */
#define CLOTHO_TRAITS_WALK_FP32 3  /* set-bit walk, FP32 sums                          */
#define CLOTHO_TRAITS_EXACT_I8 4   /* the pipeline's exact path: effects quantised to   */
                                   /* k 2^-q (largest q with |k| < 2^30), int8 digit    */
                                   /* planes on int8 MFMA, exact int64 sums x 2^-q      */

typedef struct clotho_config {
    uint64_t seed;            /* random_number.seed (Philox key)                         */
    uint32_t replicate;       /* replicate id, part of every stream counter              */
    uint32_t n_individuals;   /* N, diploid founder individuals (the population size
                                 until clotho_set_population_size changes it)            */
    uint32_t trait_count;     /* traits.N                                                */
    uint32_t fitness_kind;    /* CLOTHO_FIT_*                                            */
    double mu;                /* mutation.mu, expected new mutations per gamete          */
    double rho;               /* recombination.rho, expected crossovers per gamete       */
    double bias;              /* recombination.sequence_bias.p, strand-swap probability  */
    double neutral_p;         /* neutral.p, probability that an effect is 0              */
    double trait_mean;        /* traits.mean                                             */
    double trait_sigma;       /* traits.sigma                                            */
    double fitness_scale;     /* quadratic scale, or sigma of the fixed Gaussians        */
    double env_sigma;         /* s.d. of environmental noise added per individual (0=off)*/
    int32_t effect_quantum_log2; /* effects are integer multiples of 2^q (e.g. -24)      */
    uint32_t max_alleles;     /* allele-slot capacity S_cap, at most 2^24                */
    uint32_t compact_every;   /* 1: compact fixed and lost columns away after every
                                 generation (clotho_compact); 0: keep the reference's
                                 slot reuse                                              */
    uint32_t max_individuals; /* individual capacity for population growth (0: N)        */
    uint32_t intermediates;   /* generations.intermediates: k generations between free-
                                 space analyses (QS/engine_parallel_pipeline_batched.hpp);
                                 0 = analysis after every generation; not with
                                 compact_every                                          */
    uint32_t flags;           /* CLOTHO_FLAG_*                                           */
} clotho_config;

/* clotho_config.flags */
#define CLOTHO_FLAG_HOST_MUTATIONS 1u /* generations draw no mutations of their own: new
                                         alleles come from clotho_add_mutations (the
                                         reference's base engine, where MutationGenerator
                                         adds mutation_allocator's count after crossover);
                                         mu still sets the fitness constants            */

/* Shape and bookkeeping of the current (child) population after a generation. */
typedef struct clotho_shape {
    uint32_t generation;      /* generations completed + 1 (reference's m_generation)   */
    uint32_t n_individuals;   /* N                                                      */
    uint32_t haplotypes;      /* H = 2N rows                                            */
    uint32_t words;           /* W = ceil(alleles/512)*8 uint64 words per row           */
    uint32_t alleles;         /* allele-table size (getMaxAlleles)                      */
    uint32_t new_mutations;   /* M of the last generation                               */
    uint32_t fixed_count;     /* free_space_details::fixed_size                         */
    uint32_t lost_count;      /* free_space_details::lost_size                          */
    uint32_t free_count;      /* fixed + lost                                           */
    uint32_t variable_count;  /* alleles - free                                          */
    uint32_t all_neutral;     /* trait_space_vector::isAllNeutral                       */
    uint32_t fitness_constant;/* every fitness value equal (selection is uniform)       */
    uint32_t error;           /* sticky device-side error flag (capacity)               */
    uint32_t reserved[3];
} clotho_shape;

typedef struct clotho_ctx clotho_ctx;

/* ---- lifetime ---------------------------------------------------------- */
int clotho_create(const clotho_config* cfg, int device, clotho_ctx** out);
void clotho_destroy(clotho_ctx* ctx);
const char* clotho_last_error(const clotho_ctx* ctx);
const char* clotho_version(void);

/* Poisson inversion table used for crossover and mutation counts: cdf[k] =
 * P(X <= k), last entry forced to 1.0.  Returns the table length (<= cap) or <0. */
int clotho_poisson_table(double lambda, double* cdf, int cap);

/* ---- state upload / download (host buffers, synchronous) --------------- */
/* Founders: rows[H][W] with H = 2N, W a multiple of 8; positions[A] in [0,1)
 * (multiples of 2^-32); effects[A][T] (multiples of 2^q); traits/fitness may
 * be NULL (then computed on the device; fitness NULL means constant 1.0). */
int clotho_set_population(clotho_ctx* ctx, const uint64_t* rows, uint32_t words,
                          const double* positions, const double* effects,
                          uint32_t alleles, const double* fitness);
/* Synthetic standing variation, drawn on the device (no reference counterpart:
 * the reference starts from empty founders).  Bit j of row r is set when the
 * 32-bit Philox word (seed, generation 0, stream 7, replicate, index r,
 * sub j/4)[j%4] is below floor(freqs[j] * 2^32).  Positions/effects as in
 * clotho_set_population; fitness starts at 1.0. */
int clotho_set_synthetic_population(clotho_ctx* ctx, uint32_t alleles, const double* freqs,
                                    const double* positions, const double* effects,
                                    uint64_t seed);
int clotho_get_shape(clotho_ctx* ctx, clotho_shape* out);
int clotho_download_population(clotho_ctx* ctx, uint64_t* rows, size_t cap_words);
int clotho_download_alleles(clotho_ctx* ctx, double* positions, double* effects,
                            uint32_t* ages, size_t cap_alleles);
int clotho_download_free_space(clotho_ctx* ctx, uint32_t* fixed, uint32_t* lost,
                               size_t cap);
int clotho_download_traits(clotho_ctx* ctx, double* haplotype_traits, size_t cap);
int clotho_download_fitness(clotho_ctx* ctx, double* fitness, size_t cap);
/* Per offspring gamete o of the last generation: parent haplotype row that was
 * the top strand, crossover event count, mutation count (any pointer may be NULL). */
int clotho_download_gametes(clotho_ctx* ctx, uint32_t* top_row, uint32_t* bottom_row,
                            uint32_t* events, uint32_t* mutations, size_t cap);

/* ---- population statistics -------------------------------------------- */
/* Row sample for the per-period statistics (random_sample_generator,
 * include/clotho/data_spaces/generators/random_sample_generator.hpp:133-160):
 * k >= population gives every row 0..population-1 in order; otherwise k
 * distinct rows in draw order, each draw uniform on [0, population) from
 * Philox stream 8 (seed, generation, replicate) and rejected when already
 * taken.  Host-only; out holds min(k, population) entries. */
int clotho_sample_rows(uint64_t seed, uint32_t generation, uint32_t replicate,
                       uint32_t population, uint32_t k, uint32_t* out);
/* Column and row margins of the current population over a row subset
 * (frequency_evaluator<population_space<row_block_alignment>>::operator(),
 * include/clotho/data_spaces/analysis/allele_frequency/frequency_evaluator_population_space_row_alignment.hpp:36-61):
 * column_counts[j] = number of rows of the subset carrying slot j, for every
 * slot j < alleles (cap_columns >= alleles); row_counts[i] = alleles carried
 * by rows[i].  rows == NULL takes every haplotype (n_rows ignored, row_counts
 * holds H entries).  kernel_ms (may be NULL) receives the device time of the
 * counting kernels.  Synchronous. */
int clotho_allele_counts(clotho_ctx* ctx, const uint32_t* rows, uint32_t n_rows,
                         uint32_t* column_counts, size_t cap_columns, uint32_t* row_counts,
                         double* kernel_ms);

/* Pairwise differences of a row sample (pairwise_difference<association_matrix>
 * with the row-grouped evaluator,
 * include/clotho/data_spaces/analysis/pairwise_difference/association_matrix_pairwise_difference.hpp:78-102):
 * out[p] = popcount(row a xor row b) over the current population's words for
 * the pairs a < b of rows[0..n_rows) in the given order, packed row by row
 * (p = a n - a (a + 1) / 2 + b - a - 1; n (n - 1) / 2 values, cap at least
 * that).  The reference passes its sample through a bitset (unique rows in
 * increasing order); the C++ adapter does the same before calling.  At most
 * CLOTHO_MAX_PAIRWISE_ROWS rows.  kernel_ms (may be null) receives the kernel time. */
#define CLOTHO_MAX_PAIRWISE_ROWS 16384
int clotho_pairwise_differences(clotho_ctx* ctx, const uint32_t* rows, uint32_t n_rows,
                                uint32_t* out, size_t cap, double* kernel_ms);

/* Device memory held by the context (both populations, allele table,
 * free-space and selection buffers, scratch), in bytes.  No reference
 * counterpart (the reference's footprint is its std::vector/new[] storage). */
int clotho_device_bytes(const clotho_ctx* ctx, uint64_t* bytes);

/* ---- one generation ---------------------------------------------------- */
/* Enqueue n generations on the context stream and wait; returns the sticky
 * device error as CLOTHO_ERR_CAPACITY if one occurred.  use_graph != 0
 * replays a captured hipGraph of the generation instead of launching kernels. */
int clotho_run(clotho_ctx* ctx, uint32_t n_generations, int use_graph);
int clotho_generation(clotho_ctx* ctx);

/* Individuals of the offspring population of every following generation: the
 * pN of the reference's population-growth model, pN = growth(parent N,
 * generation) (QS/engine_parallel_pipeline.hpp:181-197;
 * src/clotho/genetics/linear_population_growth.cpp:23-25,
 * src/clotho/genetics/predefined_population_growth.cpp:25-27).  Parents are
 * drawn from the current population whatever its size.  n must not exceed the
 * capacity max(n_individuals, max_individuals) (CLOTHO_ERR_CAPACITY).  A run
 * whose first generation changes the size uses eager launches. */
int clotho_set_population_size(clotho_ctx* ctx, uint32_t n);

/* Phase-level entry points, in generation order (clotho_generation = these).
 * clotho_phase_draw also builds the per-word order statistics that
 * clotho_phase_recombine_mutate reads and the effect digit planes that
 * clotho_phase_traits reads, so the phases run in this order. */
int clotho_phase_draw(clotho_ctx* ctx);
int clotho_phase_recombine_mutate(clotho_ctx* ctx);
int clotho_phase_traits(clotho_ctx* ctx);
int clotho_phase_free_space(clotho_ctx* ctx);
int clotho_phase_fitness(clotho_ctx* ctx);
int clotho_phase_commit(clotho_ctx* ctx);
int clotho_sync(clotho_ctx* ctx);
/* Remove every fixed or lost column of the current population: variable
 * slots keep their order and are renumbered 0..V-1, rows shrink to
 * ceil(V/512)*8 words, positions/effects/ages follow their slots, the free
 * lists become empty; traits and fitness are unchanged.  Synchronous.  No
 * reference counterpart (the reference reuses free slots,
 * QS/engine_parallel_pipeline.hpp:337-348); cfg.compact_every runs it after
 * every generation of clotho_run / clotho_generation (phase-level callers
 * call it themselves). */
int clotho_compact(clotho_ctx* ctx);

/* ---- the generic operator families ------------------------------------ */
/* Parent individual of every offspring gamete of the next generation, in
 * place of the fitness-proportional draw (Crossover<RNG,GS>::update(parents,
 * mate_pairs, offspring), DS/crossover/crossover.hpp:49-72, mate pair i =
 * (parents[2i], parents[2i+1])): gamete o recombines the two rows of
 * individual parents[o], events and strand swap drawn as usual.  gametes
 * must be twice the next generation's individuals and every entry below the
 * current population size.  Applies to the next generation only. */
int clotho_set_parents(clotho_ctx* ctx, const uint32_t* parents, uint32_t gametes);
/* n new alleles on the current population
 * (MutationGenerator<RNG,GS>::operator()(child, N), DS/mutation/mutation_generator.hpp:41-52):
 * mutation k, in order, takes the next free slot (the free list, fixed
 * ascending then lost ascending, then appended slots), sets it in row rows[k]
 * and gives it position positions[k] (k 2^-32 in [0,1)), effects
 * effects[k][T] (integer multiples of 2^q) and the age of the generation
 * that produced the population.  A reused fixed slot is first cleared from
 * every row (its allele leaves the population, as removeFixedAlleles does);
 * rows widen when appended slots need it.  Traits and the free-space lists
 * are recomputed, fitness is not (clotho_evaluate_fitness).  slots (may be
 * NULL) receives each mutation's slot.  Synchronous. */
int clotho_add_mutations(clotho_ctx* ctx, const uint32_t* rows, const double* positions,
                         const double* effects, uint32_t n, uint32_t* slots);
/* Fitness of every individual of the current population from its genome
 * traits, phenotype = trait(2i) + trait(2i+1) (+ environmental noise of the
 * generation that produced the population), with the configured metric
 * (Fitness<phenotype_evaluator<...>>::update(pop, eval),
 * DS/fitness/phenotype_evaluator_fitness.hpp:51-69).  Synchronous. */
int clotho_evaluate_fitness(clotho_ctx* ctx);
/* Per-gamete recombination of host rows against the current population's
 * allele positions (common_crossover::operator()(p0b, p0e, p1b, p1e, ob, oe),
 * DS/crossover/common_crossover_row_block_alignment.hpp:53-93, with the strand
 * choice made by the caller): out[g] takes top[g] at slots with an even
 * number of events e <= position, bottom[g] at odd ones, and is zero past
 * `words` up to out_words; no events copies top.  event_keys[g * event_stride
 * + e] (e < n_events[g], any order) are positions k 2^-32 as uint32.  Bits of
 * slots at or beyond the allele count must be clear.  Synchronous. */
int clotho_crossover_rows(clotho_ctx* ctx, const uint64_t* top, const uint64_t* bottom,
                          uint32_t words, const uint32_t* n_events, const uint32_t* event_keys,
                          uint32_t event_stride, uint32_t gametes, uint64_t* out,
                          uint32_t out_words);

/* Ordered free-space classification of reduced column masks, for rows
 * accumulated on the host (free_space_accumulator_mt::finalize ->
 * free_space_details::analyze_free_indices,
 * DS/free_space/free_space_accumulator.hpp:142-151,
 * DS/free_space/free_space_details.hpp:359-408): fixed = set bits of
 * fixed_and, lost = set bits of ~variable_or, both below max_alleles, each
 * ascending; counts = {fixed, lost}.  fixed and lost hold up to
 * min(64 words, max_alleles) entries each.  Context-free, synchronous. */
int clotho_classify_free_space(const uint64_t* fixed_and, const uint64_t* variable_or,
                               uint32_t words, uint32_t max_alleles, uint32_t* fixed,
                               uint32_t* lost, uint32_t* counts, int device);

/* Kernel timing over the last clotho_run: milliseconds summed over the
 * generations for [0] whole generations, [1] draw, [2] recombine+mutate,
 * [3] traits, [4] free space, [5] fitness + commit; returns the number of
 * generations timed (eager runs; a graph run fills [0] only). */
int clotho_timings(clotho_ctx* ctx, double* ms, int cap);
void clotho_set_timing(clotho_ctx* ctx, int enabled);
/* Launch-shape autotuner (no reference counterpart; the reference's batch
 * shape is ceil(N/(T+1)) rows per host thread, QS/engine_parallel_pipeline.hpp:205-216):
 * times the coming generation's recombination for candidate row-chunk counts
 * on the resident population and keeps the fastest.  State is unchanged. */
int clotho_autotune(clotho_ctx* ctx, uint32_t* chosen_chunks);
/* The launch shapes in use (clotho_autotune tunes the recombination's row
 * chunks, the traits pass's kernel shape among the candidates for the trait
 * count, and, for contexts with compact_every, the compaction's rows per wave
 * step and row chunks; no reference counterpart). */
typedef struct clotho_launch_shape {
    uint32_t recombine_chunks; /* row chunks of the recombination grid               */
    uint32_t traits_variant;   /* traits-pass kernel shape, 1 = the built-in choice   */
    uint32_t traits_variants;  /* shapes available for the trait count               */
    uint32_t compact_rows;     /* rows per wave step of the column compaction        */
    uint32_t compact_chunks;   /* row chunks of the compaction grid                  */
    uint32_t compact_tile_words; /* compacted words per compaction wave: 48, 56, 60, 64 (default 64; not autotuned) */
    uint32_t reserved[2];
} clotho_launch_shape;
int clotho_get_launch_shape(clotho_ctx* ctx, clotho_launch_shape* out);
/* Apply launch shapes (e.g. tuned in an earlier run); compact_chunks 0 and
 * compact_tile_words 0 take the defaults; traits_variants is ignored.
 * Invalid shapes: CLOTHO_ERR_ARG. */
int clotho_set_launch_shape(clotho_ctx* ctx, const clotho_launch_shape* in);

/* ---- replicate batches ------------------------------------------------- */
/* Many independent populations of uneven size and parameters advanced
 * together: each kernel of the generation runs once for the whole batch as a
 * persistent launch over the concatenated blocks of every replicate, and a
 * batched generation can be captured as one hipGraph.  No reference
 * counterpart: the reference runs one Engine<...>::simulate per process
 * (QS/engine_parallel_pipeline.hpp:172-239, driven by QS/qtlsim_main.cpp:117-171);
 * a batch runs R of them, replicate r drawing the Philox streams of
 * cfgs[r].replicate, so its results equal a lone context with that config.
 * Every replicate must have the same trait count and compact_every = 0.  Replicate contexts are
 * owned by the batch; borrow one with clotho_batch_replicate for uploads,
 * downloads and shapes (not for clotho_run while it belongs to the batch). */
typedef struct clotho_batch clotho_batch;
#define CLOTHO_BATCH_GRAPH 0      /* batched kernels, one hipGraph per generation   */
#define CLOTHO_BATCH_EAGER 1      /* batched kernels launched directly              */
#define CLOTHO_BATCH_SEQUENTIAL 2 /* each replicate's own generation graph in turn  */
int clotho_batch_create(const clotho_config* cfgs, uint32_t count, int device,
                        clotho_batch** out);
void clotho_batch_destroy(clotho_batch* batch);
const char* clotho_batch_last_error(const clotho_batch* batch);
uint32_t clotho_batch_size(const clotho_batch* batch);
clotho_ctx* clotho_batch_replicate(clotho_batch* batch, uint32_t r);
/* Advance every replicate by n generations and wait; elapsed_ms (may be NULL)
 * receives the device time of the n generations (HIP events on the batch
 * stream, graph capture excluded).  CLOTHO_ERR_CAPACITY names the first
 * replicate whose allele table overflowed. */
int clotho_batch_run(clotho_batch* batch, uint32_t n_generations, int mode, double* elapsed_ms);
/* Per-phase kernel milliseconds of the last eager batched run, as
 * clotho_timings; returns the number of generations timed. */
int clotho_batch_timings(clotho_batch* batch, double* ms, int cap);
void clotho_batch_set_timing(clotho_batch* batch, int enabled);
/* Size of the persistent grids: times the batch's recombination and traits
 * launches over the last generation's gametes (written into the spare
 * population buffers; visible state unchanged) for grids of 1/2, 1 and 2
 * times the default blocks per compute unit and keeps the fastest;
 * percent_of_default receives the choice (50, 100 or 200). */
int clotho_batch_autotune(clotho_batch* batch, uint32_t* percent_of_default);

/* ---- standalone trait contraction -------------------------------------- */
/* z[h][t] = sum_j bit(h,j) * effects[j][t] over a device-resident genotype
 * matrix.  Host-buffer variant: uploads, runs `repeats` times, downloads z
 * (FP64 for every mode), reports the average kernel milliseconds.  words is a
 * multiple of 8 (whole 512-bit row blocks). */
int clotho_trait_contract(const uint64_t* rows, uint32_t haplotypes, uint32_t words,
                          const double* effects, uint32_t sites, uint32_t traits,
                          int mode, int split_terms, int repeats, double* z,
                          double* avg_ms, int device);

/* ---- exact trait contraction without allocation ------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): clotho_trait_contract's exact int8 path over caller-owned
 * device buffers, enqueued on `stream` (a hipStream_t; NULL is the default
 * stream) and not waited for.  rows[haplotypes][words] (words a multiple of
 * 8), quantised effects k[sites][traits] (int32, |k| <= 2^30 - 1; the effect
 * is k 2^q for the caller's q) and the output sums[haplotypes][traits]
 * (int64, the exact sum of k over each row's set bits below `sites`) are
 * device pointers.  The effect digit planes live in a fixed device array of
 * CLOTHO_NOALLOC_SCRATCH_BYTES on each GPU the library runs on, held until
 * the process ends and shared by every call on that GPU: calls that overlap
 * on one GPU race.  words * 64 up to CLOTHO_NOALLOC_MAX_SITES and traits up
 * to CLOTHO_NOALLOC_MAX_TRAITS; beyond either the call returns
 * CLOTHO_ERR_CAPACITY without enqueuing anything.  Calls no hipMalloc,
 * hipFree or allocating routine. */
#define CLOTHO_NOALLOC_MAX_SITES 131072
#define CLOTHO_NOALLOC_MAX_TRAITS 16
#define CLOTHO_NOALLOC_SCRATCH_BYTES (256 + 8388608)
int clotho_trait_contract_exact_noalloc(const uint64_t* rows, uint32_t haplotypes,
                                        uint32_t words, const int32_t* effects, uint32_t sites,
                                        uint32_t traits, int64_t* sums, void* stream);

/* ---- diagnostics ------------------------------------------------------ */
/* STREAM-style device-to-device copy of `bytes` (read + write counted);
 * no reference counterpart: it measures the HBM bandwidth the roofline is
 * quoted against and calibrates hardware byte counters. */
int clotho_stream_copy(size_t bytes, int repeats, double* gbps, int device);
/* Free and total device memory of `device` in bytes (hipMemGetInfo); no
 * reference counterpart (sizes the largest population that fits). */
int clotho_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);

/* Same contraction over a genotype matrix drawn on the device: bit j of row
 * r is set when Philox word (seed, generation 0, stream 7, replicate 0,
 * index r, sub j/4)[j%4] < floor(freqs[j] * 2^32); rows are
 * ceil(sites/512)*8 words wide.  No reference counterpart (synthetic
 * workloads for measuring the contraction at scale). */
int clotho_trait_contract_synthetic(uint32_t haplotypes, uint32_t sites, const double* freqs,
                                    uint64_t seed, const double* effects, uint32_t traits,
                                    int mode, int split_terms, int repeats, double* z,
                                    double* avg_ms, int device);

#ifdef __cplusplus
}
#endif

#endif /* CLOTHO_AMD_H_ */
