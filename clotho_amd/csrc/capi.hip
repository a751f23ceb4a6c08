// capi.hip — C ABI of the offspring pipeline (include/clotho_amd.h): context
// lifetime, device buffers sized for one population pair, state transfer,
// generation sequencing (eager launches or a captured hipGraph per parity),
// and per-phase HIP-event timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "device_common.hpp"
#include "pipeline.hpp"
#include "traits.hpp"

using namespace clotho;

namespace {

#define HIP_TRY(ctx, expr)                                                          \
    do {                                                                            \
        hipError_t e_ = (expr);                                                     \
        if (e_ != hipSuccess) {                                                     \
            (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);        \
            return CLOTHO_ERR_HIP;                                                  \
        }                                                                           \
    } while (0)

int fail(clotho_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

// Poisson(lambda) inversion table, cdf[k] = P(X <= k), last entry forced to 1.
int poisson_table(double lambda, double* cdf, int cap) {
    if (!(lambda >= 0.0) || lambda > 64.0 || cap < 2) return -1;
    double pk = std::exp(-lambda);
    double c = pk;
    int k = 0;
    cdf[0] = c;
    while (k + 1 < cap && c < 1.0 - 1e-17 && !(k > lambda && pk < 1e-18)) {
        ++k;
        pk = pk * lambda / (double)k;
        c = c + pk;
        cdf[k] = c;
    }
    cdf[k] = 1.0;
    return k + 1;
}

template <typename T>
int dalloc(clotho_ctx* c, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(c, hipMalloc((void**)p, count * sizeof(T)));
    c->device_bytes += count * sizeof(T);
    return CLOTHO_OK;
}

void free_buffers(clotho_ctx* c) {
    Buffers& B = c->b;
    void* ptrs[] = {B.rows[0],    B.rows[1],   B.traits[0], B.traits[1],  B.fitness[0],
                    B.fitness[1], B.keys,      B.effects,   B.ages,       B.free_list,
                    B.fixed_bits, B.mut_count, B.mut_local, B.mut_bsum,   B.top,
                    B.sorted,     B.pm,        B.digits,    B.bucket,
                    B.bottom,     B.n_events,  B.events,    B.var_part,   B.fix_part,
                    B.word_cnt,   B.word_bsum, B.cum_local, B.cum_block,  B.fit_stats,
                    B.cdf_rho,    B.cdf_mu,    B.state,     B.by_parent,  B.parent_cnt,
                    B.parent_bsum, B.prank,    B.given_parents};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    std::memset(&B, 0, sizeof(B));
}

}  // namespace

namespace clotho {

void destroy_graphs(clotho_ctx* c) {
    for (int i = 0; i < 2; ++i)
        if (c->graphs.exec[i]) {
            (void)hipGraphExecDestroy(c->graphs.exec[i]);
            c->graphs.exec[i] = nullptr;
        }
    c->graphs.valid = false;
}

int read_state(clotho_ctx* c, GenState* st) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(st, c->b.state, sizeof(GenState), hipMemcpyDeviceToHost));
    return CLOTHO_OK;
}

int write_state(clotho_ctx* c, const GenState& st) {
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->b.state, &st, sizeof(GenState), hipMemcpyHostToDevice));
    return CLOTHO_OK;
}

// Free space of an uploaded population: the analysis runs as at the end of
// an intermediate batch, and the next generation starts a batch.
int analyse_upload(clotho_ctx* c, GenState st) {
    st.inter = c->p.inter_k;
    st.free_base = 0;
    int rc = write_state(c, st);
    if (rc) return rc;
    HIP_TRY(c, launch_population_analysis(c));
    rc = read_state(c, &st);
    if (rc) return rc;
    st.inter = 0;
    return write_state(c, st);
}

// ifitness_generator::generate(N) (src/clotho/fitness/normal_fitness_generator.cpp:36-52,
// src/clotho/fitness/quadratic_fitness_generator.cpp:42-54): constants for the
// P.n individuals being evaluated.
void set_fitness_constants(Params& P, const clotho_config& cfg) {
    P.fitness_kind = cfg.fitness_kind;
    P.fit_coeff = 1.0;
    P.fit_inv = 1.0;
    if (cfg.fitness_kind == CLOTHO_FIT_GAUSSIAN) {
        const double sigma = std::sqrt(4.0 * (double)P.n * cfg.mu);
        P.fit_coeff = (1.0 / sigma) * 0.398942280401432677939946059934;
        P.fit_inv = 1.0 / (2.0 * sigma * sigma);
    } else if (cfg.fitness_kind == CLOTHO_FIT_QUADRATIC) {
        const double s = cfg.fitness_scale * std::sqrt(4.0 * (double)P.n * cfg.mu);
        P.fit_inv = 1.0 / s;
    } else if (cfg.fitness_kind == CLOTHO_FIT_GAUSSIAN_FIXED ||
               cfg.fitness_kind == CLOTHO_FIT_MV_STABILIZING) {
        P.fit_inv = 1.0 / (2.0 * cfg.fitness_scale * cfg.fitness_scale);
    }
}

// Sizes of the generation about to be enqueued: the current population is the
// parent, the offspring population takes the requested size
// (Engine::simulate, QS/engine_parallel_pipeline.hpp:181-197).  Launch
// arguments change with the sizes, so captured graphs are dropped.
void begin_generation(clotho_ctx* c) {
    Params& P = c->p;
    const uint32_t n = c->next_n ? c->next_n : P.n;
    if (P.n_par == P.n && n == P.n) return;
    P.n_par = P.n;
    P.n = n;
    P.h = 2 * n;
    P.mut_blocks = (P.h + kScanBlock - 1) / kScanBlock;
    set_fitness_constants(P, c->cfg);
    destroy_graphs(c);
}

// After a generation the offspring population is the current one.
void end_generation(clotho_ctx* c) {
    if (c->p.n_par == c->p.n) return;
    c->p.n_par = c->p.n;
    destroy_graphs(c);
}

int enqueue_generation(clotho_ctx* c, std::vector<hipEvent_t>* marks) {
    auto mark = [&](int i) -> hipError_t {
        if (!marks) return hipSuccess;
        return hipEventRecord((*marks)[i], c->stream);
    };
    HIP_TRY(c, mark(0));
    HIP_TRY(c, launch_draw(c));
    HIP_TRY(c, mark(1));
    HIP_TRY(c, launch_recombine(c));
    HIP_TRY(c, mark(2));
    HIP_TRY(c, launch_traits(c));
    HIP_TRY(c, mark(3));
    HIP_TRY(c, launch_free_reduce(c));
    HIP_TRY(c, mark(4));
    HIP_TRY(c, launch_compact_fitness(c));
    HIP_TRY(c, launch_commit(c));
    // the child (population 1 - cur) is compacted back into population cur
    if (c->cfg.compact_every) HIP_TRY(c, launch_compact(c, 1 - c->cur));
    HIP_TRY(c, mark(5));
    return CLOTHO_OK;
}

// Population index current after one generation from parent index cur.
int after_generation(const clotho_ctx* c, int cur) { return c->cfg.compact_every ? cur : 1 - cur; }

int capture_graphs(clotho_ctx* c) {
    if (c->graphs.valid) return CLOTHO_OK;
    const int saved = c->cur;
    for (int k = 0; k < 2; ++k) {
        c->cur = k;
        hipGraph_t g = nullptr;
        HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_generation(c, nullptr);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != CLOTHO_OK) {
            c->cur = saved;
            return rc;
        }
        if (e != hipSuccess) {
            c->cur = saved;
            return fail(c, CLOTHO_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        }
        e = hipGraphInstantiate(&c->graphs.exec[k], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->cur = saved;
            return fail(c, CLOTHO_ERR_HIP, std::string("instantiate: ") + hipGetErrorString(e));
        }
    }
    c->cur = saved;
    c->graphs.valid = true;
    return CLOTHO_OK;
}

}  // namespace clotho

extern "C" {

const char* clotho_version(void) { return "clotho_amd 0.1 (gfx950)"; }

const char* clotho_last_error(const clotho_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

int clotho_poisson_table(double lambda, double* cdf, int cap) {
    return poisson_table(lambda, cdf, cap);
}

int clotho_create(const clotho_config* cfg, int device, clotho_ctx** out) {
    if (!cfg || !out) return CLOTHO_ERR_ARG;
    *out = nullptr;
    if (cfg->n_individuals == 0 || cfg->n_individuals > (1u << 28)) return CLOTHO_ERR_ARG;
    if (cfg->max_individuals > (1u << 28)) return CLOTHO_ERR_ARG;
    if (cfg->intermediates > (1u << 16) || (cfg->intermediates && cfg->compact_every))
        return CLOTHO_ERR_ARG;
    if (cfg->trait_count > (uint32_t)kMaxTraits) return CLOTHO_ERR_ARG;
    if (!(cfg->bias >= 0.0 && cfg->bias <= 1.0)) return CLOTHO_ERR_ARG;
    if (!(cfg->neutral_p >= 0.0 && cfg->neutral_p <= 1.0)) return CLOTHO_ERR_ARG;
    if (cfg->fitness_kind > CLOTHO_FIT_MV_STABILIZING) return CLOTHO_ERR_ARG;
    if (cfg->effect_quantum_log2 > 0 || cfg->effect_quantum_log2 < -60) return CLOTHO_ERR_ARG;
    // the exact traits accumulate balanced int8 digits (|d| <= 128) over a whole
    // row in int32: at most 2^24 slots keep every partial sum in range
    if (cfg->max_alleles == 0 || cfg->max_alleles > kMaxAlleleSlots) return CLOTHO_ERR_ARG;
    if ((cfg->fitness_kind == CLOTHO_FIT_GAUSSIAN || cfg->fitness_kind == CLOTHO_FIT_QUADRATIC) &&
        !(cfg->mu > 0.0))
        return CLOTHO_ERR_ARG;
    if ((cfg->fitness_kind == CLOTHO_FIT_QUADRATIC ||
         cfg->fitness_kind == CLOTHO_FIT_GAUSSIAN_FIXED ||
         cfg->fitness_kind == CLOTHO_FIT_MV_STABILIZING) &&
        !(cfg->fitness_scale > 0.0))
        return CLOTHO_ERR_ARG;

    clotho_ctx* c = new (std::nothrow) clotho_ctx();
    if (!c) return CLOTHO_ERR_ARG;
    c->cfg = *cfg;
    c->device = device;
    if (cfg->flags & ~CLOTHO_FLAG_HOST_MUTATIONS) return CLOTHO_ERR_ARG;
    double cdf_rho[kMaxCdf], cdf_mu[kMaxCdf];
    const int nr = poisson_table(cfg->rho, cdf_rho, kMaxCdf);
    // host mutations: the per-gamete counts are Poisson(0), always zero
    const double mu_drawn = (cfg->flags & CLOTHO_FLAG_HOST_MUTATIONS) ? 0.0 : cfg->mu;
    const int nm = poisson_table(mu_drawn, cdf_mu, kMaxCdf);
    if (nr < 0 || nm < 0) {
        delete c;
        return CLOTHO_ERR_ARG;
    }
    auto bail = [&](int rc) {
        free_buffers(c);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(CLOTHO_ERR_HIP);
    if (hipGetDeviceProperties(&c->prop, device) != hipSuccess) return bail(CLOTHO_ERR_HIP);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(CLOTHO_ERR_HIP);
    if (query_trait_residency(c) != hipSuccess) return bail(CLOTHO_ERR_HIP);

    Params& P = c->p;
    std::memset(&P, 0, sizeof(P));
    P.seed = cfg->seed;
    P.replicate = cfg->replicate;
    P.n = cfg->n_individuals;
    P.h = 2 * P.n;
    P.n_par = P.n;
    P.n_cap = std::max(P.n, cfg->max_individuals);
    P.inter_k = cfg->intermediates;
    const uint32_t h_cap = 2 * P.n_cap;
    P.traits = cfg->trait_count;
    P.s_cap = ((cfg->max_alleles + 511) / 512) * 512;
    P.w_cap = P.s_cap / 64;
    P.n_cdf_rho = (uint32_t)nr;
    P.n_cdf_mu = (uint32_t)nm;
    P.ev_cap = (uint32_t)(((std::max(nr - 1, 1)) + 3) / 4 * 4);
    P.bias = cfg->bias;
    P.neutral_p = cfg->neutral_p;
    P.trait_mean = cfg->trait_mean;
    P.trait_sigma = cfg->trait_sigma;
    P.env_sigma = cfg->env_sigma;
    P.eff_scale = std::ldexp(1.0, -cfg->effect_quantum_log2);
    P.eff_quantum = std::ldexp(1.0, cfg->effect_quantum_log2);
    set_fitness_constants(P, *cfg);
    P.tiles = (P.w_cap + 63) / 64;
    {
        // default launch shape: ~4096 recombination blocks, but at least ~32
        // rows per wave (a block's order-statistic tables are loaded once per
        // block); clotho_autotune may pick another row-chunk count up to max_chunks
        P.max_chunks = std::max<uint32_t>(1, std::min<uint32_t>(
                                                 (8192 + P.tiles - 1) / P.tiles, h_cap / 8));
        uint32_t ch = (4096 + P.tiles - 1) / P.tiles;
        ch = std::min<uint32_t>(ch, std::max<uint32_t>(1, (P.h + 511) / 512));
        P.chunks = std::min<uint32_t>(P.max_chunks, std::max<uint32_t>(1, ch));
    }
    P.mut_blocks = (P.h + kScanBlock - 1) / kScanBlock;
    P.word_blocks = (P.w_cap + 63) / 64;  // 64-word free-space blocks
    const uint32_t sel_blocks = (P.n_cap + kScanBlock - 1) / kScanBlock;  // at capacity
    const uint32_t mut_blocks_cap = (h_cap + kScanBlock - 1) / kScanBlock;
    const uint32_t T1 = std::max<uint32_t>(1, P.traits);

    Buffers& B = c->b;
    std::memset(&B, 0, sizeof(B));
    int rc = CLOTHO_OK;
    for (int k = 0; k < 2 && rc == CLOTHO_OK; ++k) {
        rc |= dalloc(c, &B.rows[k], (size_t)h_cap * P.w_cap);
        rc |= dalloc(c, &B.traits[k], (size_t)h_cap * T1);
        rc |= dalloc(c, &B.fitness[k], P.n_cap);
    }
    rc |= dalloc(c, &B.keys, P.s_cap);
    rc |= dalloc(c, &B.effects, (size_t)P.s_cap * T1);
    rc |= dalloc(c, &B.ages, P.s_cap);
    rc |= dalloc(c, &B.free_list, P.s_cap + 1);
    rc |= dalloc(c, &B.fixed_bits, P.w_cap);
    rc |= dalloc(c, &B.sorted, table_elems(P.w_cap));
    rc |= dalloc(c, &B.pm, prefix_elems(P.w_cap));
    rc |= dalloc(c, &B.bucket, table_elems(P.w_cap));
    rc |= dalloc(c, &B.digits, i8_packed_elems(std::max<uint32_t>(1, P.traits), P.w_cap));
    rc |= dalloc(c, &B.mut_count, h_cap);
    rc |= dalloc(c, &B.mut_local, h_cap);
    rc |= dalloc(c, &B.mut_bsum, mut_blocks_cap);
    rc |= dalloc(c, &B.top, h_cap);
    rc |= dalloc(c, &B.bottom, h_cap);
    rc |= dalloc(c, &B.n_events, h_cap);
    rc |= dalloc(c, &B.events, (size_t)h_cap * P.ev_cap);
    rc |= dalloc(c, &B.by_parent, h_cap);
    rc |= dalloc(c, &B.parent_cnt, P.n_cap);
    rc |= dalloc(c, &B.parent_bsum, sel_blocks);
    rc |= dalloc(c, &B.prank, h_cap);
    rc |= dalloc(c, &B.given_parents, h_cap);
    rc |= dalloc(c, &B.var_part, (size_t)(P.max_chunks + 1) * P.w_cap);
    rc |= dalloc(c, &B.fix_part, (size_t)P.max_chunks * P.w_cap);
    rc |= dalloc(c, &B.word_cnt, P.w_cap);
    rc |= dalloc(c, &B.word_bsum, 2 * P.word_blocks);
    rc |= dalloc(c, &B.cum_local, P.n_cap);
    rc |= dalloc(c, &B.cum_block, sel_blocks);
    rc |= dalloc(c, &B.fit_stats, 2 * sel_blocks);
    rc |= dalloc(c, &B.cdf_rho, kMaxCdf);
    rc |= dalloc(c, &B.cdf_mu, kMaxCdf);
    rc |= dalloc(c, &B.state, 1);
    if (rc != CLOTHO_OK) return bail(rc);

    // first generation: empty founders, fitness 1.0 (GeneralFitness::resize)
    bool ok = hipMemcpy(B.cdf_rho, cdf_rho, sizeof(double) * nr, hipMemcpyHostToDevice) ==
                  hipSuccess &&
              hipMemcpy(B.cdf_mu, cdf_mu, sizeof(double) * nm, hipMemcpyHostToDevice) ==
                  hipSuccess;
    for (int k = 0; k < 2 && ok; ++k)
        ok = hipMemset(B.traits[k], 0, sizeof(int64_t) * h_cap * T1) == hipSuccess;
    ok = ok && hipMemset(B.keys, 0, sizeof(uint32_t) * P.s_cap) == hipSuccess &&
         hipMemset(B.effects, 0, sizeof(int32_t) * P.s_cap * T1) == hipSuccess &&
         hipMemset(B.ages, 0, sizeof(uint32_t) * P.s_cap) == hipSuccess &&
         hipMemset(B.fixed_bits, 0, sizeof(uint64_t) * P.w_cap) == hipSuccess &&
         hipMemset(B.top, 0, sizeof(uint32_t) * h_cap) == hipSuccess &&
         hipMemset(B.bottom, 0, sizeof(uint32_t) * h_cap) == hipSuccess &&
         hipMemset(B.n_events, 0, sizeof(uint32_t) * h_cap) == hipSuccess &&
         hipMemset(B.mut_count, 0, sizeof(uint32_t) * h_cap) == hipSuccess &&
         hipMemset(B.mut_local, 0, sizeof(uint32_t) * h_cap) == hipSuccess &&
         hipMemset(B.mut_bsum, 0, sizeof(uint32_t) * mut_blocks_cap) == hipSuccess;
    if (ok) {
        std::vector<double> ones(P.n_cap, 1.0);
        ok = hipMemcpy(B.fitness[0], ones.data(), 8 * (size_t)P.n_cap, hipMemcpyHostToDevice) ==
             hipSuccess;
    }
    if (ok) {  // child order and empty gametes until the first draw
        std::vector<uint4> rec(h_cap);
        for (uint32_t o = 0; o < h_cap; ++o) rec[o] = make_uint4(o, 0, 0, 0);
        ok = hipMemcpy(B.by_parent, rec.data(), sizeof(uint4) * (size_t)h_cap,
                       hipMemcpyHostToDevice) == hipSuccess;
    }
    GenState st;
    std::memset(&st, 0, sizeof(st));
    // Engine::init leaves m_generation at 1 before the first simulate; with k
    // intermediate generations population i of batch g is generation
    // g (k + 1) + i (gen_offset, QS/engine_parallel_pipeline_batched.hpp:390-396)
    st.gen = 1 + P.inter_k;
    st.fit_constant = 1;
    ok = ok && hipMemcpy(B.state, &st, sizeof(st), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) return bail(CLOTHO_ERR_HIP);
    c->cur = 0;
    *out = c;
    return CLOTHO_OK;
}

void clotho_destroy(clotho_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    destroy_graphs(ctx);
    free_buffers(ctx);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int clotho_set_population(clotho_ctx* c, const uint64_t* rows, uint32_t words,
                          const double* positions, const double* effects, uint32_t alleles,
                          const double* fitness) {
    if (!c) return CLOTHO_ERR_ARG;
    const Params& P = c->p;
    if (words % 8 || words > P.w_cap || (size_t)words * 64 < alleles || alleles > P.s_cap)
        return fail(c, CLOTHO_ERR_ARG, "set_population: words/alleles exceed capacity");
    if ((words && !rows) || (alleles && !positions) || (alleles && P.traits && !effects))
        return fail(c, CLOTHO_ERR_ARG, "set_population: null buffer");
    std::vector<uint32_t> keys(P.s_cap, 0);
    std::vector<int32_t> eq((size_t)P.s_cap * std::max<uint32_t>(1, P.traits), 0);
    for (uint32_t j = 0; j < alleles; ++j) {
        const double x = positions[j];
        const double k = x * 4294967296.0;
        if (!(x >= 0.0 && x < 1.0) || k != std::floor(k))
            return fail(c, CLOTHO_ERR_ARG, "set_population: positions must be k*2^-32 in [0,1)");
        keys[j] = (uint32_t)k;
        for (uint32_t t = 0; t < P.traits; ++t) {
            const double w = effects[(size_t)j * P.traits + t] * P.eff_scale;
            if (w != std::floor(w) || std::fabs(w) > kEffectLimit)
                return fail(c, CLOTHO_ERR_ARG,
                            "set_population: effects must be integer multiples of 2^q");
            eq[(size_t)j * P.traits + t] = (int32_t)w;
        }
    }
    // the bit of a slot >= alleles must be clear in every row
    if (words && alleles < (size_t)words * 64) {
        const uint32_t wfull = alleles / 64;
        for (uint32_t r = 0; r < P.h; ++r) {
            const uint64_t* row = rows + (size_t)r * words;
            for (uint32_t w = wfull; w < words; ++w) {
                const uint32_t lo = w * 64;
                const uint64_t allowed =
                    alleles > lo ? (alleles - lo >= 64 ? ~0ull : ((1ull << (alleles - lo)) - 1)) : 0;
                if (row[w] & ~allowed)
                    return fail(c, CLOTHO_ERR_ARG, "set_population: bit set beyond allele count");
            }
        }
    }
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const Buffers& B = c->b;
    const int cur = c->cur;
    if (words)
        HIP_TRY(c, hipMemcpy(B.rows[cur], rows, sizeof(uint64_t) * P.h * words,
                             hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(B.keys, keys.data(), 4 * keys.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(B.effects, eq.data(), 4 * eq.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(B.ages, 0, 4 * P.s_cap));
    std::vector<double> fit(P.n, 1.0);
    if (fitness) std::copy(fitness, fitness + P.n, fit.begin());
    HIP_TRY(c, hipMemcpy(B.fitness[cur], fit.data(), 8 * P.n, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(B.traits[cur], 0, sizeof(int64_t) * P.h * std::max<uint32_t>(1, P.traits)));
    st.alleles = alleles;
    st.all_size = alleles;
    st.wc = words;
    st.wp = words;
    st.m = 0;
    st.error = 0;
    rc = analyse_upload(c, st);
    if (rc) return rc;
    c->has_population = true;
    return CLOTHO_OK;
}

int clotho_set_synthetic_population(clotho_ctx* c, uint32_t alleles, const double* freqs,
                                    const double* positions, const double* effects,
                                    uint64_t seed) {
    if (!c) return CLOTHO_ERR_ARG;
    const Params& P = c->p;
    const uint32_t words = ((alleles + 511) / 512) * 8;
    if (alleles > P.s_cap || words > P.w_cap)
        return fail(c, CLOTHO_ERR_ARG, "synthetic population exceeds capacity");
    if (alleles && (!freqs || !positions || (P.traits && !effects)))
        return fail(c, CLOTHO_ERR_ARG, "synthetic population: null buffer");
    std::vector<uint32_t> keys(P.s_cap, 0), thr(std::max<uint32_t>(alleles, 1), 0);
    std::vector<int32_t> eq((size_t)P.s_cap * std::max<uint32_t>(1, P.traits), 0);
    for (uint32_t j = 0; j < alleles; ++j) {
        const double x = positions[j], k = x * 4294967296.0;
        if (!(x >= 0.0 && x < 1.0) || k != std::floor(k))
            return fail(c, CLOTHO_ERR_ARG, "positions must be k*2^-32 in [0,1)");
        keys[j] = (uint32_t)k;
        const double f = std::min(std::max(freqs[j], 0.0), 1.0);
        thr[j] = f >= 1.0 ? 0xFFFFFFFFu : (uint32_t)std::floor(f * 4294967296.0);
        for (uint32_t t = 0; t < P.traits; ++t) {
            const double w = effects[(size_t)j * P.traits + t] * P.eff_scale;
            if (w != std::floor(w) || std::fabs(w) > kEffectLimit)
                return fail(c, CLOTHO_ERR_ARG, "effects must be integer multiples of 2^q");
            eq[(size_t)j * P.traits + t] = (int32_t)w;
        }
    }
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const Buffers& B = c->b;
    uint32_t* d_thr = nullptr;
    HIP_TRY(c, hipMalloc((void**)&d_thr, 4 * thr.size()));
    hipError_t e = hipMemcpy(d_thr, thr.data(), 4 * thr.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && words) e = launch_synth_rows(c, seed, d_thr, alleles, words);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d_thr);
    if (e != hipSuccess) return fail(c, CLOTHO_ERR_HIP, hipGetErrorString(e));
    HIP_TRY(c, hipMemcpy(B.keys, keys.data(), 4 * keys.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(B.effects, eq.data(), 4 * eq.size(), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(B.ages, 0, 4 * P.s_cap));
    std::vector<double> fit(P.n, 1.0);
    HIP_TRY(c, hipMemcpy(B.fitness[c->cur], fit.data(), 8 * P.n, hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(B.traits[c->cur], 0,
                         sizeof(int64_t) * P.h * std::max<uint32_t>(1, P.traits)));
    st.alleles = alleles;
    st.all_size = alleles;
    st.wc = words;
    st.wp = words;
    st.m = 0;
    st.error = 0;
    rc = analyse_upload(c, st);
    if (rc) return rc;
    c->has_population = true;
    return CLOTHO_OK;
}

int clotho_get_shape(clotho_ctx* c, clotho_shape* out) {
    if (!c || !out) return CLOTHO_ERR_ARG;
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    std::memset(out, 0, sizeof(*out));
    out->generation = st.gen;
    out->n_individuals = c->p.n;
    out->haplotypes = c->p.h;
    out->words = st.wc;
    out->alleles = st.all_size;
    out->new_mutations = st.m;
    out->fixed_count = st.n_fixed;
    out->lost_count = st.free_count - st.n_fixed;
    out->free_count = st.free_count;
    out->variable_count = st.all_size - st.free_count;
    out->error = st.error;
    // all-neutral over the allele table, fitness constancy of the current population
    uint32_t* d_flags = nullptr;
    HIP_TRY(c, hipMalloc((void**)&d_flags, 64 * sizeof(uint32_t)));
    hipError_t e = launch_all_neutral(c, d_flags);
    uint32_t flags[64] = {0};
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(flags, d_flags, sizeof(flags), hipMemcpyDeviceToHost);
    (void)hipFree(d_flags);
    if (e != hipSuccess) return fail(c, CLOTHO_ERR_HIP, hipGetErrorString(e));
    uint32_t any = 0;
    for (uint32_t f : flags) any |= f;
    out->all_neutral = any ? 0 : 1;
    std::vector<double> fit(c->p.n);
    HIP_TRY(c, hipMemcpy(fit.data(), c->b.fitness[c->cur], 8 * fit.size(), hipMemcpyDeviceToHost));
    bool constant = true;
    for (double f : fit)
        if (!(f == fit[0])) {
            constant = false;
            break;
        }
    out->fitness_constant = constant ? 1 : 0;
    return CLOTHO_OK;
}

int clotho_download_population(clotho_ctx* c, uint64_t* rows, size_t cap_words) {
    if (!c || !rows) return CLOTHO_ERR_ARG;
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const size_t n = (size_t)c->p.h * st.wc;
    if (cap_words < n) return fail(c, CLOTHO_ERR_ARG, "download_population: buffer too small");
    if (n)
        HIP_TRY(c, hipMemcpy(rows, c->b.rows[c->cur], n * 8, hipMemcpyDeviceToHost));
    return CLOTHO_OK;
}

int clotho_download_alleles(clotho_ctx* c, double* positions, double* effects, uint32_t* ages,
                            size_t cap) {
    if (!c) return CLOTHO_ERR_ARG;
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const uint32_t A = st.all_size, T = c->p.traits;
    if (cap < A) return fail(c, CLOTHO_ERR_ARG, "download_alleles: buffer too small");
    if (positions && A) {
        std::vector<uint32_t> k(A);
        HIP_TRY(c, hipMemcpy(k.data(), c->b.keys, 4 * A, hipMemcpyDeviceToHost));
        for (uint32_t j = 0; j < A; ++j) positions[j] = (double)k[j] * 0x1p-32;
    }
    if (effects && A && T) {
        std::vector<int32_t> q((size_t)A * T);
        HIP_TRY(c, hipMemcpy(q.data(), c->b.effects, 4 * q.size(), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < q.size(); ++j) effects[j] = (double)q[j] * c->p.eff_quantum;
    }
    if (ages && A) HIP_TRY(c, hipMemcpy(ages, c->b.ages, 4 * A, hipMemcpyDeviceToHost));
    return CLOTHO_OK;
}

int clotho_download_free_space(clotho_ctx* c, uint32_t* fixed, uint32_t* lost, size_t cap) {
    if (!c) return CLOTHO_ERR_ARG;
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const uint32_t nf = st.n_fixed, nl = st.free_count - st.n_fixed;
    if (cap < std::max(nf, nl)) return fail(c, CLOTHO_ERR_ARG, "download_free_space: too small");
    const uint32_t* list = c->b.free_list + st.free_base;  // entries not yet handed out
    if (fixed && nf) HIP_TRY(c, hipMemcpy(fixed, list, 4 * nf, hipMemcpyDeviceToHost));
    if (lost && nl) HIP_TRY(c, hipMemcpy(lost, list + nf, 4 * nl, hipMemcpyDeviceToHost));
    return CLOTHO_OK;
}

int clotho_download_traits(clotho_ctx* c, double* out, size_t cap) {
    if (!c || !out) return CLOTHO_ERR_ARG;
    const size_t n = (size_t)c->p.h * c->p.traits;
    if (cap < n) return fail(c, CLOTHO_ERR_ARG, "download_traits: buffer too small");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (!n) return CLOTHO_OK;
    std::vector<int64_t> q(n);
    HIP_TRY(c, hipMemcpy(q.data(), c->b.traits[c->cur], 8 * n, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < n; ++i) out[i] = (double)q[i] * c->p.eff_quantum;
    return CLOTHO_OK;
}

int clotho_download_fitness(clotho_ctx* c, double* out, size_t cap) {
    if (!c || !out) return CLOTHO_ERR_ARG;
    if (cap < c->p.n) return fail(c, CLOTHO_ERR_ARG, "download_fitness: buffer too small");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, c->b.fitness[c->cur], 8 * c->p.n, hipMemcpyDeviceToHost));
    return CLOTHO_OK;
}

int clotho_download_gametes(clotho_ctx* c, uint32_t* top, uint32_t* bottom, uint32_t* events,
                            uint32_t* mutations, size_t cap) {
    if (!c) return CLOTHO_ERR_ARG;
    const size_t h = c->p.h;
    if (cap < h) return fail(c, CLOTHO_ERR_ARG, "download_gametes: buffer too small");
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (top) HIP_TRY(c, hipMemcpy(top, c->b.top, 4 * h, hipMemcpyDeviceToHost));
    if (bottom) HIP_TRY(c, hipMemcpy(bottom, c->b.bottom, 4 * h, hipMemcpyDeviceToHost));
    if (events) HIP_TRY(c, hipMemcpy(events, c->b.n_events, 4 * h, hipMemcpyDeviceToHost));
    if (mutations) HIP_TRY(c, hipMemcpy(mutations, c->b.mut_count, 4 * h, hipMemcpyDeviceToHost));
    return CLOTHO_OK;
}

int clotho_phase_draw(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    begin_generation(c);
    HIP_TRY(c, launch_draw(c));
    return CLOTHO_OK;
}
int clotho_phase_recombine_mutate(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, launch_recombine(c));
    return CLOTHO_OK;
}
int clotho_phase_traits(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, launch_traits(c));
    return CLOTHO_OK;
}
int clotho_phase_free_space(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, launch_free_space(c));
    return CLOTHO_OK;
}
int clotho_phase_fitness(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, launch_fitness(c));
    return CLOTHO_OK;
}
int clotho_phase_commit(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, launch_commit(c));
    c->cur = 1 - c->cur;
    end_generation(c);
    return CLOTHO_OK;
}

int clotho_set_population_size(clotho_ctx* c, uint32_t n) {
    if (!c) return CLOTHO_ERR_ARG;
    if (n == 0) return fail(c, CLOTHO_ERR_ARG, "population size must be positive");
    if (n > c->p.n_cap)
        return fail(c, CLOTHO_ERR_CAPACITY,
                    "population size " + std::to_string(n) + " exceeds the capacity of " +
                        std::to_string(c->p.n_cap) + " individuals (max_individuals)");
    c->next_n = n;
    return CLOTHO_OK;
}
int clotho_sample_rows(uint64_t seed, uint32_t generation, uint32_t replicate, uint32_t population,
                       uint32_t k, uint32_t* out) {
    if (!out || population == 0) return CLOTHO_ERR_ARG;
    if (k >= population) {  // every row, in order
        for (uint32_t i = 0; i < population; ++i) out[i] = i;
        return CLOTHO_OK;
    }
    std::vector<bool> taken(population, false);
    uint32_t got = 0;
    for (uint64_t d = 0; got < k; ++d) {
        const Draw4 r = philox_draw(seed, generation, kStreamSample, replicate,
                                    (uint32_t)(d / 2), (uint32_t)(d >> 33));
        const uint64_t u = (d & 1) ? k53(r.x[2], r.x[3]) : k53(r.x[0], r.x[1]);
        const uint32_t t = (uint32_t)(((unsigned __int128)u * population) >> 53);
        if (taken[t]) continue;  // rejection, as random_sample_generator
        taken[t] = true;
        out[got++] = t;
    }
    return CLOTHO_OK;
}

int clotho_allele_counts(clotho_ctx* c, const uint32_t* rows, uint32_t n_rows,
                         uint32_t* column_counts, size_t cap_columns, uint32_t* row_counts,
                         double* kernel_ms) {
    if (!c || !column_counts || !row_counts) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const uint32_t W = st.wc, A = st.all_size, H = c->p.h;
    const uint32_t n = rows ? n_rows : H;
    if (cap_columns < A) return fail(c, CLOTHO_ERR_ARG, "allele_counts: column buffer too small");
    if (rows)
        for (uint32_t i = 0; i < n; ++i)
            if (rows[i] >= H) return fail(c, CLOTHO_ERR_ARG, "allele_counts: row out of range");
    const uint32_t ngroups = stat_word_tiles(W);
    const uint32_t chunks = stat_chunks(W, n);
    uint32_t *d_sample = nullptr, *d_cols = nullptr, *d_rows = nullptr, *d_cp = nullptr,
             *d_rp = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto cleanup = [&]() {
        for (uint32_t* p : {d_sample, d_cols, d_rows, d_cp, d_rp})
            if (p) (void)hipFree(p);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    };
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    if (rows) step(hipMalloc((void**)&d_sample, 4ull * std::max<uint32_t>(n, 1)));
    step(hipMalloc((void**)&d_cols, 4ull * std::max<uint32_t>(A, 1)));
    step(hipMalloc((void**)&d_rows, 4ull * std::max<uint32_t>(n, 1)));
    step(hipMalloc((void**)&d_cp, 4ull * std::max<size_t>((size_t)chunks * W * 64, 1)));
    step(hipMalloc((void**)&d_rp, 4ull * std::max<size_t>((size_t)ngroups * n, 1)));
    step(hipEventCreate(&e0));
    step(hipEventCreate(&e1));
    if (rows && e == hipSuccess)
        step(hipMemcpyAsync(d_sample, rows, 4ull * n, hipMemcpyHostToDevice, c->stream));
    step(hipEventRecord(e0, c->stream));
    if (e == hipSuccess)
        step(launch_allele_counts(c, c->b.rows[c->cur], W, A, d_sample, n, d_cols, d_rows, d_cp,
                                  d_rp, chunks));
    step(hipEventRecord(e1, c->stream));
    if (A) step(hipMemcpyAsync(column_counts, d_cols, 4ull * A, hipMemcpyDeviceToHost, c->stream));
    if (n) step(hipMemcpyAsync(row_counts, d_rows, 4ull * n, hipMemcpyDeviceToHost, c->stream));
    step(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (e == hipSuccess) step(hipEventElapsedTime(&ms, e0, e1));
    cleanup();
    if (e != hipSuccess) return fail(c, CLOTHO_ERR_HIP, hipGetErrorString(e));
    if (kernel_ms) *kernel_ms = ms;
    return CLOTHO_OK;
}

int clotho_pairwise_differences(clotho_ctx* c, const uint32_t* rows, uint32_t n_rows,
                                uint32_t* out, size_t cap, double* kernel_ms) {
    if (!c || (!rows && n_rows) || (!out && n_rows > 1)) return CLOTHO_ERR_ARG;
    if (n_rows > CLOTHO_MAX_PAIRWISE_ROWS)
        return fail(c, CLOTHO_ERR_ARG, "pairwise_differences: at most " +
                                           std::to_string(CLOTHO_MAX_PAIRWISE_ROWS) + " rows");
    const size_t pairs = (size_t)n_rows * (n_rows ? n_rows - 1 : 0) / 2;
    if (cap < pairs) return fail(c, CLOTHO_ERR_ARG, "pairwise_differences: output too small");
    HIP_TRY(c, hipSetDevice(c->device));
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const uint32_t W = st.wc, H = c->p.h;
    for (uint32_t i = 0; i < n_rows; ++i)
        if (rows[i] >= H) return fail(c, CLOTHO_ERR_ARG, "pairwise_differences: row out of range");
    if (kernel_ms) *kernel_ms = 0.0;
    if (pairs == 0) return CLOTHO_OK;
    uint32_t *d_sample = nullptr, *d_out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    step(hipMalloc((void**)&d_sample, 4ull * n_rows));
    step(hipMalloc((void**)&d_out, 4ull * pairs));
    step(hipEventCreate(&e0));
    step(hipEventCreate(&e1));
    if (e == hipSuccess)
        step(hipMemcpyAsync(d_sample, rows, 4ull * n_rows, hipMemcpyHostToDevice, c->stream));
    if (e == hipSuccess) step(hipEventRecord(e0, c->stream));
    if (e == hipSuccess) step(launch_pairwise(c, c->b.rows[c->cur], W, d_sample, n_rows, d_out));
    if (e == hipSuccess) step(hipEventRecord(e1, c->stream));
    if (e == hipSuccess)
        step(hipMemcpyAsync(out, d_out, 4ull * pairs, hipMemcpyDeviceToHost, c->stream));
    step(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    if (e == hipSuccess) step(hipEventElapsedTime(&ms, e0, e1));
    for (uint32_t* p : {d_sample, d_out})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (e != hipSuccess) return fail(c, CLOTHO_ERR_HIP, hipGetErrorString(e));
    if (kernel_ms) *kernel_ms = ms;
    return CLOTHO_OK;
}

int clotho_set_parents(clotho_ctx* c, const uint32_t* parents, uint32_t gametes) {
    if (!c || (!parents && gametes)) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t next_n = c->next_n ? c->next_n : c->p.n;
    if (gametes != 2 * next_n)
        return fail(c, CLOTHO_ERR_ARG, "set_parents: expected " + std::to_string(2 * next_n) +
                                           " gamete parents (two per offspring individual)");
    for (uint32_t o = 0; o < gametes; ++o)
        if (parents[o] >= c->p.n)
            return fail(c, CLOTHO_ERR_ARG, "set_parents: parent " + std::to_string(parents[o]) +
                                               " outside the current population");
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    HIP_TRY(c, hipMemcpy(c->b.given_parents, parents, 4ull * gametes, hipMemcpyHostToDevice));
    st.given = 1;
    return write_state(c, st);
}

int clotho_add_mutations(clotho_ctx* c, const uint32_t* rows, const double* positions,
                         const double* effects, uint32_t n, uint32_t* slots_out) {
    if (!c) return CLOTHO_ERR_ARG;
    if (n == 0) return CLOTHO_OK;
    const Params& P = c->p;
    if (!rows || !positions || (P.traits && !effects))
        return fail(c, CLOTHO_ERR_ARG, "add_mutations: null buffer");
    HIP_TRY(c, hipSetDevice(c->device));
    const uint32_t T = P.traits;
    std::vector<uint32_t> keys(n);
    std::vector<int32_t> eq((size_t)n * std::max<uint32_t>(1, T), 0);
    for (uint32_t k = 0; k < n; ++k) {
        if (rows[k] >= P.h) return fail(c, CLOTHO_ERR_ARG, "add_mutations: row out of range");
        const double x = positions[k], kk = x * 4294967296.0;
        if (!(x >= 0.0 && x < 1.0) || kk != std::floor(kk))
            return fail(c, CLOTHO_ERR_ARG, "add_mutations: positions must be k*2^-32 in [0,1)");
        keys[k] = (uint32_t)kk;
        for (uint32_t t = 0; t < T; ++t) {
            const double w = effects[(size_t)k * T + t] * P.eff_scale;
            if (w != std::floor(w) || std::fabs(w) > kEffectLimit)
                return fail(c, CLOTHO_ERR_ARG,
                            "add_mutations: effects must be integer multiples of 2^q");
            eq[(size_t)k * T + t] = (int32_t)w;
        }
    }
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    // slots: the free list (fixed ascending, lost ascending), then appended
    const uint32_t F = st.free_count, A = st.all_size;
    const uint64_t A1 = (uint64_t)A + (n > F ? n - F : 0);
    if (A1 > P.s_cap)
        return fail(c, CLOTHO_ERR_CAPACITY, "add_mutations: allele capacity exceeded");
    std::vector<uint32_t> slots(n);
    const uint32_t reuse = std::min(n, F);
    if (reuse)
        HIP_TRY(c, hipMemcpy(slots.data(), c->b.free_list + st.free_base, 4ull * reuse,
                             hipMemcpyDeviceToHost));
    for (uint32_t k = reuse; k < n; ++k) slots[k] = A + (k - F);
    const uint32_t W0 = st.wc, W1 = (uint32_t)(((A1 + 511) / 512) * 8);
    if (W1 > W0) {  // wider rows: re-stride into the other buffer, which becomes current
        HIP_TRY(c, launch_restride(c, c->cur, W0, W1));
        const size_t nt = (size_t)P.h * std::max<uint32_t>(1, T);
        HIP_TRY(c, hipMemcpyAsync(c->b.traits[1 - c->cur], c->b.traits[c->cur], 8 * nt,
                                  hipMemcpyDeviceToDevice, c->stream));
        HIP_TRY(c, hipMemcpyAsync(c->b.fitness[1 - c->cur], c->b.fitness[c->cur], 8ull * P.n,
                                  hipMemcpyDeviceToDevice, c->stream));
        c->cur = 1 - c->cur;
    }
    uint32_t *d_rows = nullptr, *d_slots = nullptr, *d_keys = nullptr;
    int32_t* d_eff = nullptr;
    uint64_t* d_mask = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    // reused fixed slots: the allele leaves every row first (removeFixedAlleles)
    const uint32_t n_fixed_reused = std::min(reuse, st.n_fixed);
    if (n_fixed_reused) {
        std::vector<uint64_t> mask(W1, 0);
        for (uint32_t k = 0; k < n_fixed_reused; ++k) mask[slots[k] / 64] |= 1ull << (slots[k] % 64);
        step(hipMalloc((void**)&d_mask, 8ull * W1));
        if (e == hipSuccess)
            step(hipMemcpy(d_mask, mask.data(), 8ull * W1, hipMemcpyHostToDevice));
        if (e == hipSuccess) step(launch_clear_columns(c, c->cur, W1, d_mask));
    }
    step(hipMalloc((void**)&d_rows, 4ull * n));
    step(hipMalloc((void**)&d_slots, 4ull * n));
    step(hipMalloc((void**)&d_keys, 4ull * n));
    step(hipMalloc((void**)&d_eff, 4ull * eq.size()));
    if (e == hipSuccess) step(hipMemcpy(d_rows, rows, 4ull * n, hipMemcpyHostToDevice));
    if (e == hipSuccess) step(hipMemcpy(d_slots, slots.data(), 4ull * n, hipMemcpyHostToDevice));
    if (e == hipSuccess) step(hipMemcpy(d_keys, keys.data(), 4ull * n, hipMemcpyHostToDevice));
    if (e == hipSuccess) step(hipMemcpy(d_eff, eq.data(), 4ull * eq.size(), hipMemcpyHostToDevice));
    // the age of the generation that produced the population
    const uint32_t age = st.gen > 0 ? st.gen - 1 : 0;
    if (e == hipSuccess)
        step(launch_add_mutations(c, c->cur, W1, d_rows, d_slots, d_keys, d_eff, n, age));
    step(hipStreamSynchronize(c->stream));
    for (void* p : {(void*)d_rows, (void*)d_slots, (void*)d_keys, (void*)d_eff, (void*)d_mask})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(c, CLOTHO_ERR_HIP, hipGetErrorString(e));
    // traits and free space of the new population
    st.alleles = st.all_size = (uint32_t)A1;
    st.wc = st.wp = W1;
    st.m = n;
    rc = analyse_upload(c, st);
    if (rc) return rc;
    if (slots_out) std::copy(slots.begin(), slots.end(), slots_out);
    return CLOTHO_OK;
}

int clotho_evaluate_fitness(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, launch_fitness_current(c));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return CLOTHO_OK;
}

int clotho_crossover_rows(clotho_ctx* c, const uint64_t* top, const uint64_t* bottom,
                          uint32_t words, const uint32_t* n_events, const uint32_t* event_keys,
                          uint32_t event_stride, uint32_t gametes, uint64_t* out,
                          uint32_t out_words) {
    if (!c) return CLOTHO_ERR_ARG;
    if (gametes == 0) return CLOTHO_OK;
    if (!top || !bottom || !n_events || !out || out_words < words)
        return fail(c, CLOTHO_ERR_ARG, "crossover_rows: null buffer or out_words < words");
    HIP_TRY(c, hipSetDevice(c->device));
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    const uint32_t A = st.all_size;
    if ((size_t)words * 64 > (size_t)c->p.s_cap)
        return fail(c, CLOTHO_ERR_ARG, "crossover_rows: rows wider than the allele capacity");
    size_t n_ev = 0;
    for (uint32_t g = 0; g < gametes; ++g) {
        if (n_events[g] > event_stride)
            return fail(c, CLOTHO_ERR_ARG, "crossover_rows: more events than event_stride");
        n_ev = std::max(n_ev, (size_t)n_events[g]);
    }
    if (n_ev && !event_keys) return fail(c, CLOTHO_ERR_ARG, "crossover_rows: null event keys");
    // bits of slots >= A have no position
    for (uint32_t g = 0; g < gametes; ++g)
        for (uint32_t w = A / 64; w < words; ++w) {
            const uint32_t lo = w * 64;
            const uint64_t allowed =
                A > lo ? (A - lo >= 64 ? ~0ull : ((1ull << (A - lo)) - 1)) : 0;
            if ((top[(size_t)g * words + w] | bottom[(size_t)g * words + w]) & ~allowed)
                return fail(c, CLOTHO_ERR_ARG, "crossover_rows: bit set beyond the allele count");
        }
    // the classifier sorts its events (region_classifier's constructor)
    const uint32_t stride = std::max<uint32_t>(1, event_stride);
    std::vector<uint32_t> ev((size_t)gametes * stride, 0);
    for (uint32_t g = 0; g < gametes; ++g) {
        const uint32_t ne = n_events[g];
        if (ne) std::copy(event_keys + (size_t)g * event_stride,
                          event_keys + (size_t)g * event_stride + ne, ev.begin() + (size_t)g * stride);
        std::sort(ev.begin() + (size_t)g * stride, ev.begin() + (size_t)g * stride + ne);
    }
    const size_t rw = (size_t)gametes * words, ow = (size_t)gametes * out_words;
    uint64_t *d_top = nullptr, *d_bot = nullptr, *d_out = nullptr;
    uint32_t *d_nev = nullptr, *d_ev = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    step(hipMalloc((void**)&d_top, 8 * std::max<size_t>(rw, 1)));
    step(hipMalloc((void**)&d_bot, 8 * std::max<size_t>(rw, 1)));
    step(hipMalloc((void**)&d_out, 8 * ow));
    step(hipMalloc((void**)&d_nev, 4ull * gametes));
    step(hipMalloc((void**)&d_ev, 4 * ev.size()));
    if (e == hipSuccess && rw) step(hipMemcpy(d_top, top, 8 * rw, hipMemcpyHostToDevice));
    if (e == hipSuccess && rw) step(hipMemcpy(d_bot, bottom, 8 * rw, hipMemcpyHostToDevice));
    if (e == hipSuccess) step(hipMemcpy(d_nev, n_events, 4ull * gametes, hipMemcpyHostToDevice));
    if (e == hipSuccess) step(hipMemcpy(d_ev, ev.data(), 4 * ev.size(), hipMemcpyHostToDevice));
    if (e == hipSuccess)
        step(launch_crossover_rows(c, d_top, d_bot, words, d_nev, d_ev, stride, gametes, d_out,
                                   out_words));
    step(hipStreamSynchronize(c->stream));
    if (e == hipSuccess) step(hipMemcpy(out, d_out, 8 * ow, hipMemcpyDeviceToHost));
    for (void* p : {(void*)d_top, (void*)d_bot, (void*)d_out, (void*)d_nev, (void*)d_ev})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(c, CLOTHO_ERR_HIP, hipGetErrorString(e));
    return CLOTHO_OK;
}

int clotho_classify_free_space(const uint64_t* fixed_and, const uint64_t* variable_or,
                               uint32_t words, uint32_t max_alleles, uint32_t* fixed,
                               uint32_t* lost, uint32_t* counts, int device) {
    if (!fixed_and || !variable_or || !fixed || !lost || !counts || words == 0)
        return CLOTHO_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return CLOTHO_ERR_HIP;
    const size_t cap = std::min<size_t>((size_t)words * 64, max_alleles);
    uint64_t *d_fix = nullptr, *d_var = nullptr;
    uint32_t *d_fixed = nullptr, *d_lost = nullptr, *d_counts = nullptr;
    hipStream_t st = nullptr;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) {
        if (e == hipSuccess) e = r;
    };
    step(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    step(hipMalloc((void**)&d_fix, 8ull * words));
    step(hipMalloc((void**)&d_var, 8ull * words));
    step(hipMalloc((void**)&d_fixed, 4 * std::max<size_t>(cap, 1)));
    step(hipMalloc((void**)&d_lost, 4 * std::max<size_t>(cap, 1)));
    step(hipMalloc((void**)&d_counts, 8));
    if (e == hipSuccess) step(hipMemcpy(d_fix, fixed_and, 8ull * words, hipMemcpyHostToDevice));
    if (e == hipSuccess) step(hipMemcpy(d_var, variable_or, 8ull * words, hipMemcpyHostToDevice));
    if (e == hipSuccess)
        step(launch_classify_columns(d_fix, d_var, words, max_alleles, d_fixed, d_lost, d_counts,
                                     st));
    if (e == hipSuccess) step(hipStreamSynchronize(st));
    uint32_t n[2] = {0, 0};
    if (e == hipSuccess) step(hipMemcpy(n, d_counts, 8, hipMemcpyDeviceToHost));
    if (e == hipSuccess && n[0]) step(hipMemcpy(fixed, d_fixed, 4ull * n[0], hipMemcpyDeviceToHost));
    if (e == hipSuccess && n[1]) step(hipMemcpy(lost, d_lost, 4ull * n[1], hipMemcpyDeviceToHost));
    for (void* p : {(void*)d_fix, (void*)d_var, (void*)d_fixed, (void*)d_lost, (void*)d_counts})
        if (p) (void)hipFree(p);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) return CLOTHO_ERR_HIP;
    counts[0] = n[0];
    counts[1] = n[1];
    return CLOTHO_OK;
}

int clotho_device_bytes(const clotho_ctx* c, uint64_t* bytes) {
    if (!c || !bytes) return CLOTHO_ERR_ARG;
    *bytes = c->device_bytes;
    return CLOTHO_OK;
}

int clotho_compact(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, launch_compact(c, c->cur));
    c->cur = 1 - c->cur;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return CLOTHO_OK;
}
int clotho_sync(clotho_ctx* c) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return CLOTHO_OK;
}

void clotho_set_timing(clotho_ctx* c, int enabled) {
    if (c) c->timing.enabled = enabled != 0;
}

int clotho_timings(clotho_ctx* c, double* ms, int cap) {
    if (!c || !ms) return CLOTHO_ERR_ARG;
    for (int i = 0; i < cap && i < 6; ++i) ms[i] = c->timing.ms[i];
    return c->timing.generations;
}

int clotho_run(clotho_ctx* c, uint32_t n, int use_graph) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    Timing& tm = c->timing;
    for (double& v : tm.ms) v = 0.0;
    tm.generations = 0;
    std::vector<hipEvent_t> marks;
    // a generation that changes the population size runs with eager launches
    if (c->next_n && c->next_n != c->p.n) use_graph = 0;
    if (use_graph) {
        end_generation(c);  // steady sizes: parent and offspring of equal size
        int rc = capture_graphs(c);
        if (rc) return rc;
        EventSet events;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (tm.enabled) {
            HIP_TRY(c, events.add(&e0));
            HIP_TRY(c, events.add(&e1));
            HIP_TRY(c, hipEventRecord(e0, c->stream));
        }
        for (uint32_t g = 0; g < n; ++g) {
            HIP_TRY(c, hipGraphLaunch(c->graphs.exec[c->cur], c->stream));
            c->cur = after_generation(c, c->cur);
        }
        if (tm.enabled) {
            HIP_TRY(c, hipEventRecord(e1, c->stream));
            HIP_TRY(c, hipEventSynchronize(e1));
            float ms = 0.f;
            HIP_TRY(c, hipEventElapsedTime(&ms, e0, e1));
            tm.ms[0] = ms;
            tm.generations = (int)n;
        }
    } else {
        EventSet events;
        std::vector<std::vector<hipEvent_t>> all;
        for (uint32_t g = 0; g < n; ++g) {
            std::vector<hipEvent_t>* mk = nullptr;
            if (tm.enabled) {
                all.emplace_back(6);
                for (hipEvent_t& e : all.back()) HIP_TRY(c, events.add(&e));
                mk = &all.back();
            }
            begin_generation(c);
            int rc = enqueue_generation(c, mk);
            if (rc) return rc;
            c->cur = after_generation(c, c->cur);
            end_generation(c);
        }
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (tm.enabled) {
            for (auto& mk : all) {
                float t = 0.f;
                HIP_TRY(c, hipEventElapsedTime(&t, mk[0], mk[5]));
                tm.ms[0] += t;
                for (int q = 0; q < 5; ++q) {
                    HIP_TRY(c, hipEventElapsedTime(&t, mk[q], mk[q + 1]));
                    tm.ms[1 + q] += t;
                }
            }
            tm.generations = (int)n;
        }
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    if (st.error)
        return fail(c, CLOTHO_ERR_CAPACITY,
                    "allele capacity exceeded at generation " + std::to_string(st.error_gen));
    c->has_population = true;
    return CLOTHO_OK;
}

int clotho_generation(clotho_ctx* c) { return clotho_run(c, 1, 0); }

// Launch-shape autotuner: times the recombination kernel for candidate
// row-chunk counts and keeps the fastest.  It replays the last generation's
// gametes (parents, events, mutation counts; zeros before the first
// generation) against the resident population and writes only buffers the
// next generation rewrites: the spare row buffer, the column partials and the
// per-word order statistics.  Visible state is unchanged.
int clotho_autotune(clotho_ctx* c, uint32_t* chosen_chunks) {
    if (!c) return CLOTHO_ERR_ARG;
    HIP_TRY(c, hipSetDevice(c->device));
    GenState st;
    int rc = read_state(c, &st);
    if (rc) return rc;
    if (st.wp == 0 || st.wc == 0) {  // empty parents: nothing to time
        if (chosen_chunks) *chosen_chunks = c->p.chunks;
        return CLOTHO_OK;
    }
    const uint32_t base = c->p.chunks;
    std::vector<uint32_t> cand;
    for (uint32_t f : {1u, 2u, 4u, 8u}) {
        cand.push_back(std::max<uint32_t>(1, base / f));
        cand.push_back(std::min<uint32_t>(c->p.max_chunks, base * f));
    }
    // grids that fill whole waves of resident blocks (two per CU)
    const uint32_t resident = 2u * (uint32_t)std::max(1, c->prop.multiProcessorCount);
    for (uint32_t k = 1; k <= 16; ++k) {
        const uint32_t ch = (k * resident) / std::max<uint32_t>(1, c->p.tiles);
        if (ch >= 1 && ch <= c->p.max_chunks) cand.push_back(ch);
    }
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    // Each timed recombination is followed by the traits pass over its output,
    // as in a generation, so the caches hold what they hold in the pipeline;
    // two passes over the candidates, each candidate keeping its faster pass.
    // Launches longer than 2 ms are timed once, in one pass.
    constexpr int kReps = 3;
    EventSet events;
    hipEvent_t ev[2 * kReps] = {};
    for (hipEvent_t& e : ev) HIP_TRY(c, events.add(&e));
    int reps = kReps, passes = 2;
    HIP_TRY(c, launch_word_sort(c));  // order statistics of the resident columns
    {
        HIP_TRY(c, launch_recombine(c));
        HIP_TRY(c, hipEventRecord(ev[0], c->stream));
        HIP_TRY(c, launch_recombine(c));
        HIP_TRY(c, hipEventRecord(ev[1], c->stream));
        HIP_TRY(c, hipEventSynchronize(ev[1]));
        float ms = 0.f;
        HIP_TRY(c, hipEventElapsedTime(&ms, ev[0], ev[1]));
        if (ms > 2.0f) reps = passes = 1;
    }
    std::vector<float> best_ms(cand.size(), 1e30f);
    for (int pass = 0; pass < passes; ++pass) {
        for (size_t i = 0; i < cand.size(); ++i) {
            c->p.chunks = cand[i];
            HIP_TRY(c, launch_recombine(c));  // warm-up
            HIP_TRY(c, launch_traits(c));
            for (int r = 0; r < reps; ++r) {
                HIP_TRY(c, hipEventRecord(ev[2 * r], c->stream));
                HIP_TRY(c, launch_recombine(c));
                HIP_TRY(c, hipEventRecord(ev[2 * r + 1], c->stream));
                HIP_TRY(c, launch_traits(c));
            }
            HIP_TRY(c, hipEventSynchronize(ev[2 * reps - 1]));
            float tot = 0.f;
            for (int r = 0; r < reps; ++r) {
                float ms = 0.f;
                HIP_TRY(c, hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]));
                tot += ms;
            }
            best_ms[i] = std::min(best_ms[i], tot);
        }
    }
    float best = 1e30f;
    uint32_t best_chunks = base;
    for (size_t i = 0; i < cand.size(); ++i) {
        if (best_ms[i] < best) {
            best = best_ms[i];
            best_chunks = cand[i];
        }
    }
    c->p.chunks = best_chunks;

    // The traits pass: every kernel shape over the current population (the
    // pass writes the traits it already holds), fastest of two passes.
    auto time_it = [&](auto&& launch, float* out_ms) -> int {
        HIP_TRY(c, launch());  // warm-up
        HIP_TRY(c, hipEventRecord(ev[0], c->stream));
        for (int r = 0; r < kReps; ++r) HIP_TRY(c, launch());
        HIP_TRY(c, hipEventRecord(ev[1], c->stream));
        HIP_TRY(c, hipEventSynchronize(ev[1]));
        HIP_TRY(c, hipEventElapsedTime(out_ms, ev[0], ev[1]));
        return CLOTHO_OK;
    };
    if (c->p.traits) {
        const uint32_t nv = trait_variants(c);
        std::vector<float> tbest(nv + 1, 1e30f);
        for (int pass = 0; pass < 2; ++pass)
            for (uint32_t v = 1; v <= nv; ++v) {
                c->trait_variant = v;
                float ms = 0.f;
                rc = time_it([&] { return launch_traits_current(c); }, &ms);
                if (rc) return rc;
                tbest[v] = std::min(tbest[v], ms);
            }
        uint32_t bv = 1;
        for (uint32_t v = 2; v <= nv; ++v)
            if (tbest[v] < tbest[bv]) bv = v;
        c->trait_variant = bv;
    }
    // The column compaction of contexts that compact after every generation:
    // rows per wave step and row chunks, the row pass timed from the current
    // population into the spare buffer (its plan rebuilt first; every buffer
    // the plan writes is rebuilt by the next generation before it is read).
    // The population was compacted at the end of the last generation, so the
    // timed pass removes nothing: the tile width, whose merit depends on
    // where columns are removed, is left as set.
    if (c->cfg.compact_every) {
        HIP_TRY(c, launch_compact_plan(c));
        float bms = 1e30f;
        uint32_t brows = c->compact_rows, bchunks = c->compact_chunks;
        c->compact_chunks = 0;
        const uint32_t dflt = compact_chunks_of(c);
        for (uint32_t rows : {4u, 8u, 16u})
            for (uint32_t ch : {std::max<uint32_t>(1, dflt / 2), dflt, dflt * 2}) {
                c->compact_rows = rows;
                c->compact_chunks = ch;
                float ms = 0.f;
                rc = time_it([&] { return launch_compact_rows(c, c->cur); }, &ms);
                if (rc) return rc;
                if (ms < bms) {
                    bms = ms;
                    brows = rows;
                    bchunks = ch;
                }
            }
        c->compact_rows = brows;
        c->compact_chunks = bchunks;
    }
    destroy_graphs(c);  // captured launches bake the shape in
    if (chosen_chunks) *chosen_chunks = best_chunks;
    return CLOTHO_OK;
}

int clotho_set_launch_shape(clotho_ctx* c, const clotho_launch_shape* in) {
    if (!c || !in) return CLOTHO_ERR_ARG;
    const uint32_t nv = c->p.traits ? trait_variants(c) : 1u;
    if (in->recombine_chunks < 1 || in->recombine_chunks > c->p.max_chunks ||
        in->traits_variant < 1 || in->traits_variant > nv ||
        (in->compact_rows != 4 && in->compact_rows != 8 && in->compact_rows != 16) ||
        (in->compact_tile_words != 0 && in->compact_tile_words != 48 &&
         in->compact_tile_words != 56 && in->compact_tile_words != 60 &&
         in->compact_tile_words != 64))
        return fail(c, CLOTHO_ERR_ARG,
                    "set_launch_shape: chunks 1.." + std::to_string(c->p.max_chunks) +
                        ", traits shape 1.." + std::to_string(nv) +
                        ", compaction rows 4/8/16, compaction tile words 48/56/60/64");
    c->p.chunks = in->recombine_chunks;
    c->trait_variant = in->traits_variant;
    c->compact_rows = in->compact_rows;
    c->compact_chunks = in->compact_chunks;  // 0: the default
    c->compact_tile_words = in->compact_tile_words ? in->compact_tile_words : 64u;
    destroy_graphs(c);
    return CLOTHO_OK;
}

int clotho_get_launch_shape(clotho_ctx* c, clotho_launch_shape* out) {
    if (!c || !out) return CLOTHO_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->recombine_chunks = c->p.chunks;
    out->traits_variant = c->trait_variant ? c->trait_variant : 1u;
    out->traits_variants = c->p.traits ? trait_variants(c) : 0u;
    out->compact_rows = c->compact_rows;
    out->compact_chunks = compact_chunks_of(c);
    out->compact_tile_words = c->compact_tile_words;
    return CLOTHO_OK;
}

}  // extern "C"
