/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference's per-generation offspring pipeline
 * (Engine<parallel_pipeline>::simulate, examples/QTLSim/engine_parallel_pipeline.hpp:172-239,
 * with offspring_generator<population_space<row_block_alignment<uint64>>>),
 * fed by the same counter-based event streams the GPU library draws.  It is
 * the checker for the HIP path: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 */
#ifndef CLOTHO_ORACLE_H_
#define CLOTHO_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#include "../include/clotho_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block: ctr[4], key[2] -> out[4]. */
void or_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
/* One stream draw: key = seed, counter = (index, sub, generation, stream | replicate<<8). */
void or_draw(uint64_t seed, uint32_t generation, uint32_t stream, uint32_t replicate,
             uint32_t index, uint32_t sub, uint32_t* out);

/* Deterministic elementary functions (fixed +,-,* sequences). */
double or_det_log(double x);
double or_det_exp(double x);
double or_det_sqrt(double x);
double or_det_recip(double x);
double or_normal_ppf(double p);
int or_poisson_table(double lambda, double* cdf, int cap);

/* Synthetic genotype rows [r0, r1) of the device generator (stream 7). */
void or_synth_rows(uint64_t seed, uint32_t replicate, const uint32_t* thr, uint32_t sites,
                   uint32_t words, uint32_t r0, uint32_t r1, uint64_t* out);

/* Per-phase restatements on caller buffers. */
/* block_crossover_method over one gamete: child[Wc] from top/bottom rows of
 * width Wp, sorted or unsorted event positions, allele positions. */
void or_crossover(const uint64_t* top, const uint64_t* bottom, uint32_t wp, uint64_t* child,
                  uint32_t wc, const double* positions, uint32_t alleles,
                  const double* events, uint32_t n_events);
/* free_space_accumulator_mt + analyze_free_indices over rows[H][W]; returns
 * fixed count in *n_fixed and lost count in *n_lost. */
void or_free_space(const uint64_t* rows, uint32_t h, uint32_t w, uint32_t max_alleles,
                   uint32_t* fixed, uint32_t* n_fixed, uint32_t* lost, uint32_t* n_lost);
/* common_phenotype_accumulator over rows[H][W]: traits[H][T]. */
void or_phenotype(const uint64_t* rows, uint32_t h, uint32_t w, const double* effects,
                  uint32_t alleles, uint32_t t, double* traits);
/* Parent draw: uniform when fitness is constant, else quantised FP64 CDF search. */
void or_select(const double* fitness, uint32_t n, uint64_t seed, uint32_t generation,
               uint32_t replicate, uint32_t* parents /* 2n */);
double or_fitness_value(const clotho_config* cfg, uint32_t n, const double* z);
/* Environmental noise env_sigma * Phi^-1(u) of (generation, individual, trait),
   the draw the simulation adds to the diploid phenotype. */
double or_env_noise(const clotho_config* cfg, uint32_t generation, uint32_t individual,
                    uint32_t trait);

/* Whole-simulation oracle. */
typedef struct or_sim or_sim;
or_sim* or_sim_create(const clotho_config* cfg);
void or_sim_destroy(or_sim* s);
int or_sim_set_population(or_sim* s, const uint64_t* rows, uint32_t words,
                          const double* positions, const double* effects, uint32_t alleles,
                          const double* fitness);
void or_sim_set_threads(or_sim* s, int threads);
/* Individuals of the offspring population of every following generation. */
int or_sim_set_population_size(or_sim* s, uint32_t n);
int or_sim_step(or_sim* s);
/* Given parent individual per offspring gamete of the next generation. */
int or_sim_set_parents(or_sim* s, const uint32_t* parents, uint32_t gametes);
/* n new alleles on the current population, mutation k in the next free slot. */
int or_sim_add_mutations(or_sim* s, const uint32_t* rows, const double* positions,
                         const double* effects, uint32_t n, uint32_t* slots);
/* Fitness of the current population from its genome traits. */
void or_sim_evaluate_fitness(or_sim* s);
/* Remove every fixed or lost column of the current population (slots keep
 * their order and are renumbered; the free lists become empty). */
int or_sim_compact(or_sim* s);
/* Column and row margins of rows[sample[i]] (i < n), W words per row, slots < A. */
void or_allele_counts(const uint64_t* rows, uint32_t W, uint32_t A, const uint32_t* sample,
                      uint32_t n, uint32_t* column_counts, uint32_t* row_counts);
/* popcount(row a xor row b) for the sample pairs a < b, packed row by row. */
void or_pairwise_differences(const uint64_t* rows, uint32_t W, const uint32_t* sample, uint32_t n,
                             uint32_t* out);
void or_sim_shape(const or_sim* s, clotho_shape* out);
void or_sim_rows(const or_sim* s, uint64_t* out);
void or_sim_alleles(const or_sim* s, double* positions, double* effects, uint32_t* ages);
void or_sim_free_space(const or_sim* s, uint32_t* fixed, uint32_t* lost);
void or_sim_traits(const or_sim* s, double* traits);
void or_sim_fitness(const or_sim* s, double* fitness);
void or_sim_gametes(const or_sim* s, uint32_t* top_row, uint32_t* bottom_row,
                    uint32_t* events, uint32_t* mutations);
/* Phase wall-clock seconds of the last step: crossover, mutate, phenotype,
 * fixed, fitness, finalize, draws (mirrors the reference's performance.* keys). */
void or_sim_phase_seconds(const or_sim* s, double* out7);

#ifdef __cplusplus
}
#endif

#endif
