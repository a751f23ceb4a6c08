// clotho_amd/engine.hpp — C++17 host adapters over the C ABI (clotho_amd.h)
// that keep the reference's names and call shapes for the offspring path, so
// a caller written against Engine<..., parallel_pipeline> switches by type.
//
// Reference families mirrored (DS/ = include/clotho/data_spaces/,
// QS/ = examples/QTLSim/):
//   property_tree          boost::property_tree::ptree keys used by the path
//                          (src/clotho/utility/clotho_strings.cpp:16-47), in config.hpp
//   population_growth_toolkit, ipopulation_growth
//                          src/clotho/genetics/population_growth_toolkit.cpp:36-62,
//                          in population_growth.hpp
//   ifitness, ifitness_generator, fitness_toolkit
//                          include/clotho/fitness/ifitness.hpp:21-38,
//                          include/clotho/fitness/ifitness_generator.hpp:22-44,
//                          src/clotho/fitness/fitness_toolkit.cpp:50-78
//   mutation_allocator     DS/mutation/mutation_allocator.hpp:25-55
//   common_crossover       DS/crossover/common_crossover_row_block_alignment.hpp:29-100
//   free_space_accumulator_mt + free_space_details accessors
//                          DS/free_space/free_space_accumulator.hpp:53-197,
//                          DS/free_space/free_space_details.hpp:193-277
//   GeneralFitness         DS/fitness/general_fitness.hpp:30-284
//   Engine<..., mi355x_pipeline>
//                          QS/engine_parallel_pipeline.hpp:59-483
// and the generic families of the reference's base engine
// (QS/qtlsim_engine.hpp:123-160):
//   Crossover<RNG,GS>::update, crossover_method<RNG,A,S>
//                          DS/crossover/crossover.hpp:27-80,
//                          DS/crossover/crossover_method_def.hpp:20-21
//   MutationGenerator<RNG,GS>::operator()(child, N)
//                          DS/mutation/mutation_generator.hpp:25-59
//   TraitWeightAccumulator, phenotype_evaluator, Fitness<phenotype_evaluator<...>>::update
//                          DS/phenotype_evaluator/trait_accumulator.hpp:67-84,
//                          DS/phenotype_evaluator/phenotype_evaluator.hpp:52-76,
//                          DS/fitness/phenotype_evaluator_fitness.hpp:25-72
//   FreeSpaceAnalyzer<GS>::update, free_space_evaluator<Space,Size>
//                          DS/free_space/free_space.hpp:29-78,
//                          DS/free_space/free_space_evaluator_def.hpp:20-21,
//                          DS/free_space/row_grouped_free_evaluator.hpp:121-199
//
// The per-gamete and per-thread granularity of the pipeline engine collapses
// into whole-generation device phases: each adapter delegates to the phase
// entry point that covers its work.  The generic families keep their own
// granularity (mate pairs, explicit mutation counts, single gametes) through
// the C ABI's host-directed entry points.  Errors from the C ABI throw
// std::runtime_error (the reference asserts).
#ifndef CLOTHO_AMD_ENGINE_HPP_
#define CLOTHO_AMD_ENGINE_HPP_

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <functional>
#include <map>
#include <memory>
#include <ostream>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../clotho_amd.h"
#include "config.hpp"
#include "population_growth.hpp"

namespace clotho_amd {

inline void check(int rc, const clotho_ctx* ctx, const char* what) {
    if (rc != CLOTHO_OK) {
        std::ostringstream os;
        os << what << " failed (" << rc << ")";
        if (ctx) os << ": " << clotho_last_error(ctx);
        throw std::runtime_error(os.str());
    }
}

// ---------------------------------------------------------------------------
// fitness plug-ins: ifitness / ifitness_generator / fitness_toolkit
// ---------------------------------------------------------------------------
struct ifitness {
    typedef double result_type;
    virtual double operator()(const std::vector<double>& pheno) = 0;
    virtual double operator()(double* first, double* last) = 0;
    virtual const std::string name() const = 0;
    virtual void log(std::ostream& out) const = 0;
    virtual ~ifitness() {}
};

struct ifitness_generator {
    virtual std::shared_ptr<ifitness_generator> create(property_tree& config) const = 0;
    virtual std::shared_ptr<ifitness> generate(size_t N) = 0;
    virtual const std::string name() const = 0;
    virtual void log(std::ostream& out) const = 0;
    // the device metric this generator runs as
    virtual uint32_t device_kind() const = 0;
    virtual double device_scale() const { return 1.0; }
    virtual ~ifitness_generator() {}
};

namespace detail {
// Host-side evaluation of a device metric (for callers that score a single
// phenotype); the device evaluates the same formula per individual.
class metric : public ifitness {
public:
    metric(std::string n, std::function<double(const double*, const double*)> f)
        : m_name(std::move(n)), m_f(std::move(f)) {}
    double operator()(const std::vector<double>& p) override {
        return m_f(p.data(), p.data() + p.size());
    }
    double operator()(double* first, double* last) override { return m_f(first, last); }
    const std::string name() const override { return m_name; }
    void log(std::ostream& out) const override { out << "{" << m_name << "}"; }

private:
    std::string m_name;
    std::function<double(const double*, const double*)> m_f;
};

class generator : public ifitness_generator {
public:
    typedef std::function<std::shared_ptr<ifitness>(const property_tree&, size_t)> factory;
    generator(std::string n, uint32_t kind, factory f, property_tree cfg = property_tree())
        : m_name(std::move(n)), m_kind(kind), m_factory(std::move(f)), m_cfg(std::move(cfg)) {}
    std::shared_ptr<ifitness_generator> create(property_tree& config) const override {
        return std::make_shared<generator>(m_name, m_kind, m_factory, config);
    }
    std::shared_ptr<ifitness> generate(size_t N) override { return m_factory(m_cfg, N); }
    const std::string name() const override { return m_name; }
    void log(std::ostream& out) const override { out << "{" << m_name << "_gen}"; }
    uint32_t device_kind() const override { return m_kind; }
    double device_scale() const override { return m_cfg.get("fitness.params.scale", 1.0); }

private:
    std::string m_name;
    uint32_t m_kind;
    factory m_factory;
    property_tree m_cfg;
};
}  // namespace detail

// String-keyed registry of fitness generators, selected by "fitness.name".
class fitness_toolkit {
public:
    static fitness_toolkit* getInstance() {
        static fitness_toolkit inst;
        return &inst;
    }
    void register_tool(std::shared_ptr<ifitness_generator> gen) { m_tools[gen->name()] = gen; }
    // The generator named by fitness.name; an absent block is written back
    // empty (src/clotho/fitness/fitness_toolkit.cpp:50-69).
    std::shared_ptr<ifitness_generator> get_tool(property_tree& config) const {
        if (!config.has_child("fitness")) {
            config.put("fitness.name", std::string(""));
            config.put("fitness.params", std::string(""));
            return nullptr;
        }
        const std::string n = config.get("fitness.name", std::string(""));
        auto it = m_tools.find(n);
        if (it == m_tools.end()) return nullptr;
        return it->second->create(config);
    }
    // fitness.toolkit.<name>.{name, params, depends}: every generator with the
    // parameters it reads and the keys of other blocks it depends on
    // (fitness_toolkit::tool_configurations, src/clotho/fitness/fitness_toolkit.cpp:20-48)
    void tool_configurations(property_tree& config) const {
        for (const auto& kv : m_tools) {
            const std::string p = "fitness.toolkit." + kv.first;
            config.put(p + ".name", kv.first);
            const uint32_t k = kv.second->device_kind();
            if (k == CLOTHO_FIT_QUADRATIC || k == CLOTHO_FIT_GAUSSIAN_FIXED ||
                k == CLOTHO_FIT_MV_STABILIZING)
                config.put(p + ".params.scale", 1.0);
            if (k == CLOTHO_FIT_GAUSSIAN || k == CLOTHO_FIT_QUADRATIC)
                config.put(p + ".depends.mutation.mu", 1e-4);
        }
    }
    std::vector<std::string> tool_names() const {
        std::vector<std::string> v;
        for (auto& kv : m_tools) v.push_back(kv.first);
        return v;
    }

private:
    fitness_toolkit() {
        using detail::generator;
        using detail::metric;
        // "gaussian": sigma = sqrt(4 N mu), (1/(sigma sqrt(2 pi))) exp(-x^2/(2 sigma^2)),
        // first trait (src/clotho/fitness/normal_fitness_generator.cpp:36-52)
        register_tool(std::make_shared<generator>(
            "gaussian", CLOTHO_FIT_GAUSSIAN, [](const property_tree& c, size_t N) {
                const double s = std::sqrt(4.0 * (double)N * c.get("mutation.mu", 1e-4));
                return std::make_shared<metric>("gaussian", [s](const double* a, const double* b) {
                    const double x = a == b ? 0.0 : *a;
                    return (1.0 / s) * 0.398942280401432677939946059934 *
                           std::exp(-(x * x) / (2.0 * s * s));
                });
            }));
        // "quadratic": s = scale sqrt(4 N mu), max(0, 1 - (x/s)^2)
        // (src/clotho/fitness/quadratic_fitness_generator.cpp:42-54)
        register_tool(std::make_shared<generator>(
            "quadratic", CLOTHO_FIT_QUADRATIC, [](const property_tree& c, size_t N) {
                const double s = c.get("fitness.params.scale", 1.0) *
                                 std::sqrt(4.0 * (double)N * c.get("mutation.mu", 1e-4));
                return std::make_shared<metric>("quadratic", [s](const double* a, const double* b) {
                    const double x = a == b ? 0.0 : *a;
                    const double r = (x / s) * (x / s);
                    return r >= 1.0 ? 0.0 : 1.0 - r;
                });
            }));
        register_tool(std::make_shared<generator>(
            "constant", CLOTHO_FIT_CONSTANT, [](const property_tree&, size_t) {
                return std::make_shared<metric>("constant",
                                                [](const double*, const double*) { return 1.0; });
            }));
        // fixed-sigma Gaussian exp(-x^2/(2 scale^2)) and multivariate stabilising
        // exp(-|x|^2/(2 scale^2)): new generators behind the same interface
        register_tool(std::make_shared<generator>(
            "gaussian_fixed", CLOTHO_FIT_GAUSSIAN_FIXED, [](const property_tree& c, size_t) {
                const double s = c.get("fitness.params.scale", 1.0);
                return std::make_shared<metric>(
                    "gaussian_fixed", [s](const double* a, const double* b) {
                        const double x = a == b ? 0.0 : *a;
                        return std::exp(-(x * x) / (2.0 * s * s));
                    });
            }));
        register_tool(std::make_shared<generator>(
            "mv_stabilizing", CLOTHO_FIT_MV_STABILIZING, [](const property_tree& c, size_t) {
                const double s = c.get("fitness.params.scale", 1.0);
                return std::make_shared<metric>(
                    "mv_stabilizing", [s](const double* a, const double* b) {
                        double ss = 0.0;
                        for (const double* p = a; p != b; ++p) ss += *p * *p;
                        return std::exp(-ss / (2.0 * s * s));
                    });
            }));
    }
    std::map<std::string, std::shared_ptr<ifitness_generator>> m_tools;
};

// ---------------------------------------------------------------------------
// device population: owns the C ABI context
// ---------------------------------------------------------------------------
// Founder size and individual capacity from the population-growth model:
// the reference starts from pN = growth(0, 0) (QS/engine_parallel_pipeline.hpp:150-153).
// The capacity covers the model's sizes over generations.size generations
// (or population_growth.capacity when given).  Without a model the size is
// population_growth.size, constant.
struct population_sizes {
    uint32_t founders = 0, capacity = 0;
};

inline population_sizes population_size(property_tree& config) {
    population_sizes out;
    auto growth = population_growth_toolkit::getInstance()->get_tool(config);
    if (growth) {
        const uint32_t horizon = (uint32_t)config.get("generations.size", 10000.0);
        const std::vector<uint32_t> sizes = growth_schedule(*growth, horizon);
        out.founders = sizes[0];
        for (uint32_t v : sizes) out.capacity = std::max(out.capacity, v);
    } else {
        out.founders = (uint32_t)config.get("population_growth.size", 0.0);
        out.capacity = out.founders;
    }
    if (config.has("population_growth.capacity"))
        out.capacity = std::max(out.founders, (uint32_t)config.get("population_growth.capacity", 0.0));
    if (out.founders == 0)
        throw std::runtime_error(
            "population_growth must give a positive founder size (e.g. linear with B >= 1)");
    return out;
}

inline int device_of(const property_tree& config) { return (int)config.get("device", 0.0); }

inline clotho_config make_config(property_tree& config) {
    clotho_config c{};
    c.seed = (uint64_t)config.get("random_number.seed", 0.0);
    c.replicate = (uint32_t)config.get("replicate", 0.0);
    const population_sizes ps = population_size(config);
    c.n_individuals = ps.founders;
    c.max_individuals = ps.capacity;
    c.trait_count = (uint32_t)config.get("traits.N", 1.0);
    c.mu = config.get("mutation.mu", 1e-4);
    c.rho = config.get("recombination.rho", 1e-4);
    c.bias = config.get("recombination.sequence_bias.p", 0.5);
    c.neutral_p = config.get("neutral.p", 0.5);
    c.trait_mean = config.get("traits.mean", 0.0);
    c.trait_sigma = config.get("traits.sigma", 1.0);
    c.env_sigma = config.get("environment.sigma", 0.0);
    c.effect_quantum_log2 = (int32_t)config.get("traits.quantum_log2", -24.0);
    c.max_alleles = (uint32_t)config.get("alleles.capacity", 1048576.0);
    // generations.intermediates (QS/engine_parallel_pipeline_batched.hpp:155-166), written back
    c.intermediates = (uint32_t)config.get("generations.intermediates", 0.0);
    if (!config.has("generations.intermediates")) config.put("generations.intermediates", 0.0);
    // no reference key: generations draw no mutations, MutationGenerator adds them
    c.flags = config.get("mutation.host_directed", 0.0) != 0.0 ? CLOTHO_FLAG_HOST_MUTATIONS : 0u;
    auto gen = fitness_toolkit::getInstance()->get_tool(config);
    c.fitness_kind = gen ? gen->device_kind() : CLOTHO_FIT_CONSTANT;
    c.fitness_scale = gen ? gen->device_scale() : 1.0;
    return c;
}

class device_population {
public:
    explicit device_population(const clotho_config& cfg, int device = 0) : m_cfg(cfg) {
        check(clotho_create(&cfg, device, &m_ctx), nullptr, "clotho_create");
    }
    ~device_population() { clotho_destroy(m_ctx); }
    device_population(const device_population&) = delete;
    device_population& operator=(const device_population&) = delete;

    clotho_ctx* ctx() const { return m_ctx; }
    const clotho_config& config() const { return m_cfg; }

    clotho_shape shape() const {
        clotho_shape s{};
        check(clotho_get_shape(m_ctx, &s), m_ctx, "clotho_get_shape");
        return s;
    }
    // population_space accessors
    unsigned int getIndividualCount() const { return shape().n_individuals; }
    unsigned int haploid_genome_count() const { return shape().haplotypes; }
    unsigned int getMaxBlocks() const { return shape().words; }
    unsigned int getMaxAlleles() const { return shape().alleles; }
    std::vector<uint64_t> rows() const {
        const clotho_shape s = shape();
        std::vector<uint64_t> r((size_t)s.haplotypes * s.words);
        check(clotho_download_population(m_ctx, r.data(), r.size()), m_ctx, "download");
        return r;
    }
    std::vector<double> genome_traits() const {
        std::vector<double> t((size_t)shape().haplotypes * m_cfg.trait_count);
        check(clotho_download_traits(m_ctx, t.data(), t.size()), m_ctx, "download_traits");
        return t;
    }
    std::vector<double> fitness() const {
        std::vector<double> f(getIndividualCount());
        check(clotho_download_fitness(m_ctx, f.data(), f.size()), m_ctx, "download_fitness");
        return f;
    }
    // genetic_space accessors of the generic families
    size_t sequence_count() const { return haploid_genome_count(); }
    size_t individual_count() const { return getIndividualCount(); }
    size_t allele_count() const { return getMaxAlleles(); }
    unsigned int trait_count() const { return m_cfg.trait_count; }

private:
    clotho_config m_cfg;
    clotho_ctx* m_ctx = nullptr;
};

// AlleleSpace<double, unsigned>-compatible view (positions, ages) plus effects.
struct allele_space_view {
    std::vector<double> positions, effects;
    std::vector<uint32_t> ages;
    size_t size() const { return positions.size(); }
    double getAllelePosition(size_t i) const { return positions[i]; }
    uint32_t getAlleleAge(size_t i) const { return ages[i]; }
};

// ---------------------------------------------------------------------------
// operator families
// ---------------------------------------------------------------------------

// mutation_allocator<RNG, IntType>: allocate(N) draws M ~ Poisson(mu * N) from
// the caller's engine, as the reference does, for callers that size their own
// mutation batch (MutationGenerator).  A pipeline context instead draws an
// independent Poisson(mu) count per gamete on the device (the same law for
// the total); drawn(pop) is the M of that population's last generation.
template <class RNG, class IntType = size_t>
class mutation_allocator {
public:
    mutation_allocator(RNG* rng, property_tree& config)
        : m_rng(rng), m_mu(config.get("mutation.mu", 1e-4)) {}
    IntType allocate(size_t N) {
        const double lambda = m_mu * (double)N;
        if (!(lambda > 0.0)) return 0;
        std::poisson_distribution<long long> dist(lambda);
        return (IntType)dist(*m_rng);
    }
    static IntType drawn(const device_population& pop) {
        return (IntType)pop.shape().new_mutations;
    }

private:
    RNG* m_rng;
    double m_mu;
};

namespace detail {
// Single gametes queued for one device call (clotho_crossover_rows): each
// keeps its input rows, event keys and output location until flush().
class gamete_queue {
public:
    void push(const uint64_t* top, const uint64_t* bottom, uint32_t words,
              std::vector<uint32_t> keys, uint64_t* out, uint32_t out_words) {
        if (!m_items.empty() && (words != m_words || out_words != m_out_words))
            throw std::runtime_error("gamete_queue: rows of one batch share their widths");
        m_words = words;
        m_out_words = out_words;
        m_items.push_back(item{top, bottom, std::move(keys), out});
    }
    size_t size() const { return m_items.size(); }
    void flush(clotho_ctx* ctx) {
        if (m_items.empty()) return;
        const uint32_t G = (uint32_t)m_items.size(), W = m_words, OW = m_out_words;
        size_t stride = 1;
        for (const item& it : m_items) stride = std::max(stride, it.keys.size());
        std::vector<uint64_t> top((size_t)G * W), bot((size_t)G * W), out((size_t)G * OW);
        std::vector<uint32_t> nev(G), keys((size_t)G * stride, 0);
        for (uint32_t g = 0; g < G; ++g) {
            const item& it = m_items[g];
            std::copy(it.top, it.top + W, top.begin() + (size_t)g * W);
            std::copy(it.bottom, it.bottom + W, bot.begin() + (size_t)g * W);
            nev[g] = (uint32_t)it.keys.size();
            std::copy(it.keys.begin(), it.keys.end(), keys.begin() + (size_t)g * stride);
        }
        std::vector<item> items;
        items.swap(m_items);
        check(clotho_crossover_rows(ctx, top.data(), bot.data(), W, nev.data(), keys.data(),
                                    (uint32_t)stride, G, out.data(), OW),
              ctx, "crossover_rows");
        for (uint32_t g = 0; g < G; ++g)
            std::copy(out.begin() + (size_t)g * OW, out.begin() + (size_t)(g + 1) * OW,
                      items[g].out);
    }

private:
    struct item {
        const uint64_t *top, *bottom;
        std::vector<uint32_t> keys;
        uint64_t* out;
    };
    std::vector<item> m_items;
    uint32_t m_words = 0, m_out_words = 0;
};
}  // namespace detail

// crossover_method<RNG, AlleleSpace, SequenceSpace>: one gamete of one parent
// genome at a time, the generic recombination method
// (DS/crossover/crossover_method_def.hpp:20-21; the row-grouped instance,
// DS/crossover/row_grouped_crossover_method.hpp:108-372, draws its events with
// small_crossover_event_generator, DS/generators/small_crossover_event_generator.hpp:99-154).
// Events come from the caller's engine: E ~ Poisson(rho), E positions uniform
// on [0, 1) as k 2^-32, a strand swap ~ Bernoulli(bias); the rows are
// recombined on the device against the allele positions of the population
// given to initAlleles, with the pipeline's tie rule (events <= position
// count).  A genome is the individual's two consecutive rows.  Gametes run
// one device call each, or one call per batch between begin_batch() and
// flush().
template <class RNG, class AlleleSpace, class SequenceSpace>
class crossover_method {
public:
    typedef uint64_t block_type;
    crossover_method(RNG* rng, property_tree& config)
        : m_rng(rng),
          m_rho(config.get("recombination.rho", 1e-4)),
          m_bias(config.get("recombination.sequence_bias.p", 0.5)) {}
    crossover_method(RNG* rng, double rho, double bias) : m_rng(rng), m_rho(rho), m_bias(bias) {}

    void initAlleles(SequenceSpace& pop) { m_pop = &pop; }

    // one gamete of the genome whose first row starts at `genome` (rows of
    // p_step words): the recombined row is written to c (c_step words)
    void crossover(const block_type* genome, size_t p_step, block_type* c, size_t c_step) {
        const bool swap = std::bernoulli_distribution(m_bias)(*m_rng);
        const block_type* first = genome;
        const block_type* second = genome + p_step;
        recombine(swap ? second : first, swap ? first : second, (uint32_t)p_step, c,
                  (uint32_t)c_step);
    }
    // the recombination of two given strands, top first (no strand draw)
    void recombine(const block_type* top, const block_type* bottom, uint32_t words,
                   block_type* out, uint32_t out_words) {
        if (!m_pop) throw std::runtime_error("crossover_method: initAlleles first");
        const uint32_t E = m_rho > 0.0 ? (uint32_t)std::poisson_distribution<int>(m_rho)(*m_rng) : 0u;
        std::uniform_int_distribution<uint32_t> key(0u, 0xFFFFFFFFu);
        std::vector<uint32_t> keys(E);
        for (uint32_t& k : keys) k = key(*m_rng);
        m_events.push_back(keys);
        m_queue.push(top, bottom, words, std::move(keys), out, out_words);
        if (!m_batch) flush();
    }
    void begin_batch() { m_batch = true; }
    void flush() {
        if (m_pop) m_queue.flush(m_pop->ctx());
        m_batch = false;
    }
    // event positions (k 2^-32 as uint32) of every gamete since the last
    // clear_events(), in call order
    const std::vector<std::vector<uint32_t>>& events() const { return m_events; }
    void clear_events() { m_events.clear(); }
    size_t pending() const { return m_queue.size(); }
    double rho() const { return m_rho; }
    double bias() const { return m_bias; }

private:
    RNG* m_rng;
    double m_rho, m_bias;
    SequenceSpace* m_pop = nullptr;
    bool m_batch = false;
    detail::gamete_queue m_queue;
    std::vector<std::vector<uint32_t>> m_events;
};

// common_crossover<RNG, population, alleles>: operator()(Population&) is the
// recombination + mutation scatter of every gamete of the generation (the draw
// phase supplies parents, events and strand swaps); the reference's
// per-gamete operator()(p0b, p0e, p1b, p1e, ob, oe)
// (DS/crossover/common_crossover_row_block_alignment.hpp:53-93) draws E ~
// Poisson(rho), E positions and the strand swap from the caller's engine and
// recombines the given rows on the device against the allele positions of the
// population given to bind(); gametes queue into one device call between
// begin_batch() and flush(), else each runs at once.
template <class RNG, class Population, class Alleles>
class common_crossover {
public:
    typedef uint64_t block_type;
    common_crossover(RNG* rng, Alleles* alleles, double recomb_rate, double bias_rate)
        : m_rng(rng), m_alleles(alleles), m_method(rng, recomb_rate, bias_rate) {}
    void operator()(Population& pop) {
        check(clotho_phase_recombine_mutate(pop.ctx()), pop.ctx(), "phase_recombine_mutate");
    }
    void bind(Population& pop) { m_method.initAlleles(pop); }
    void operator()(const block_type* p0_start, const block_type* p0_end,
                    const block_type* p1_start, const block_type* p1_end, block_type* offspring,
                    block_type* offspring_end) {
        if (p1_end - p1_start != p0_end - p0_start || offspring_end - offspring < p0_end - p0_start)
            throw std::runtime_error("common_crossover: parent rows of one width, offspring wider");
        const bool swap = std::bernoulli_distribution(m_method.bias())(*m_rng);
        m_swaps.push_back(swap);
        m_method.recombine(swap ? p1_start : p0_start, swap ? p0_start : p1_start,
                           (uint32_t)(p0_end - p0_start), offspring,
                           (uint32_t)(offspring_end - offspring));
    }
    void begin_batch() { m_method.begin_batch(); }
    void flush() { m_method.flush(); }
    // strand swaps and event positions of every gamete since clear_events()
    const std::vector<bool>& swaps() const { return m_swaps; }
    const std::vector<std::vector<uint32_t>>& events() const { return m_method.events(); }
    void clear_events() {
        m_swaps.clear();
        m_method.clear_events();
    }

private:
    RNG* m_rng;
    Alleles* m_alleles;
    crossover_method<RNG, Alleles, Population> m_method;
    std::vector<bool> m_swaps;
};

// free_space_details<Size> accessors over a fixed-then-lost index list
// (DS/free_space/free_space_details.hpp:193-277).
template <class SizeType = unsigned int>
class free_space_details {
public:
    typedef SizeType size_type;
    typedef const size_type* iterator;
    size_type fixed_size() const { return m_fixed; }
    size_type lost_size() const { return m_lost; }
    size_type free_size() const { return m_fixed + m_lost; }
    size_type variable_count() const { return m_size - free_size(); }
    size_type getFree(unsigned int i) const { return m_indices[i]; }
    size_type getFixed(unsigned int i) const { return m_indices[i]; }
    size_type getLost(unsigned int i) const { return m_indices[m_fixed + i]; }
    iterator fixed_begin() const { return m_indices.data(); }
    iterator fixed_end() const { return m_indices.data() + m_fixed; }
    iterator lost_begin() const { return m_indices.data() + m_fixed; }
    iterator lost_end() const { return m_indices.data() + m_fixed + m_lost; }
    iterator free_begin() const { return m_indices.data(); }
    iterator free_end() const { return m_indices.data() + m_fixed + m_lost; }

    // the device analysis of the population's current rows
    void refresh(const device_population& pop) {
        const clotho_shape s = pop.shape();
        std::vector<uint32_t> fx(std::max<uint32_t>(s.fixed_count, 1)),
            ls(std::max<uint32_t>(s.lost_count, 1));
        check(clotho_download_free_space(pop.ctx(), fx.data(), ls.data(),
                                         std::max(fx.size(), ls.size())),
              pop.ctx(), "download_free_space");
        assign(fx.data(), s.fixed_count, ls.data(), s.lost_count, s.alleles);
    }

protected:
    void assign(const uint32_t* fx, uint32_t nf, const uint32_t* ls, uint32_t nl, uint32_t size) {
        m_indices.assign(fx, fx + nf);
        m_indices.insert(m_indices.end(), ls, ls + nl);
        m_fixed = nf;
        m_lost = nl;
        m_size = size;
    }
    std::vector<size_type> m_indices;
    size_type m_fixed = 0, m_lost = 0, m_size = 0;
};

// free_space_accumulator_mt<Block, Size> with the free_space_details
// accessors.  finalize(pop) reduces the device column partials the
// recombination phase accumulated.  For rows produced on the host (the
// per-gamete common_crossover), each thread ORs / ANDs its rows into its own
// buffer (getThreadBuffer(i).update(row, len), as the reference's
// thread_buffer does, DS/free_space/free_space_accumulator.hpp:53-192) and
// finalize(max_alleles) reduces the buffers and classifies the columns on the
// device (clotho_classify_free_space).
template <class BlockType = uint64_t, class SizeType = unsigned int>
class free_space_accumulator_mt : public free_space_details<SizeType> {
public:
    typedef SizeType size_type;
    typedef BlockType block_type;

    struct thread_buffer {
        free_space_accumulator_mt* m_parent = nullptr;
        block_type* m_thread_var_buffer = nullptr;
        block_type* m_thread_fix_buffer = nullptr;
        thread_buffer() = default;
        thread_buffer(free_space_accumulator_mt* par, block_type* var, block_type* fix)
            : m_parent(par), m_thread_var_buffer(var), m_thread_fix_buffer(fix) {}
        void update(const block_type* buf, unsigned int len) {
            m_parent->updateVariableBuffer(m_thread_var_buffer, buf, len);
            m_parent->updateFixedBuffer(m_thread_fix_buffer, buf, len);
        }
    };
    typedef thread_buffer buffer_type;

    explicit free_space_accumulator_mt(unsigned int thread_count)
        : m_thread_count(std::max(1u, thread_count)) {}
    // block_count words per row; variable buffers start at 0, fixed at ~0
    void resetBuffers(unsigned int block_count) {
        m_blocks_per_row = block_count;
        m_var.assign((size_t)m_thread_count * block_count, 0);
        m_fix.assign((size_t)m_thread_count * block_count, ~(block_type)0);
    }
    buffer_type getThreadBuffer(unsigned int i) {
        if (i >= m_thread_count) throw std::runtime_error("getThreadBuffer: thread out of range");
        return buffer_type(this, m_var.data() + (size_t)i * m_blocks_per_row,
                           m_fix.data() + (size_t)i * m_blocks_per_row);
    }
    void updateVariableBuffer(block_type* dest, const block_type* src, unsigned int n) {
        for (unsigned int k = 0; k < n; ++k) dest[k] |= src[k];
    }
    // a row shorter than the buffer leaves the trailing fixed words clear
    void updateFixedBuffer(block_type* dest, const block_type* src, unsigned int n) {
        for (unsigned int k = 0; k < n; ++k) dest[k] &= src[k];
        for (unsigned int k = n; k < m_blocks_per_row; ++k) dest[k] = 0;
    }
    // the thread buffers reduced, then fixed (AND) and lost (~OR) below
    // max_alleles, classified on the device
    void finalize(unsigned int max_alleles, int device = 0) {
        const unsigned int W = m_blocks_per_row;
        if (W == 0) {
            this->assign(nullptr, 0, nullptr, 0, max_alleles);
            return;
        }
        std::vector<block_type> var(m_var.begin(), m_var.begin() + W),
            fix(m_fix.begin(), m_fix.begin() + W);
        for (unsigned int t = 1; t < m_thread_count; ++t)
            for (unsigned int k = 0; k < W; ++k) {
                var[k] |= m_var[(size_t)t * W + k];
                fix[k] &= m_fix[(size_t)t * W + k];
            }
        const size_t cap = std::max<size_t>(1, std::min<size_t>((size_t)W * 64, max_alleles));
        std::vector<uint32_t> fx(cap), ls(cap), n(2, 0);
        check(clotho_classify_free_space(fix.data(), var.data(), W, max_alleles, fx.data(),
                                         ls.data(), n.data(), device),
              nullptr, "classify_free_space");
        this->assign(fx.data(), n[0], ls.data(), n[1], max_alleles);
    }
    // device column OR/AND partials were accumulated by the recombination
    // phase; finalize reduces them and compacts the fixed and lost lists.
    void finalize(device_population& pop) {
        check(clotho_phase_free_space(pop.ctx()), pop.ctx(), "phase_free_space");
        this->refresh(pop);
    }

private:
    unsigned int m_thread_count, m_blocks_per_row = 0;
    std::vector<block_type> m_var, m_fix;
};

// GeneralFitness: phenotype (traits) + fitness of the child population.
class GeneralFitness {
public:
    typedef double fitness_type;
    typedef std::vector<double>::const_iterator const_iterator;
    explicit GeneralFitness(property_tree& config)
        : m_gen(fitness_toolkit::getInstance()->get_tool(config)) {}
    void operator()(device_population& pop) {
        check(clotho_phase_traits(pop.ctx()), pop.ctx(), "phase_traits");
        check(clotho_phase_fitness(pop.ctx()), pop.ctx(), "phase_fitness");
        m_cached = false;
    }
    const std::vector<double>& values(device_population& pop) {
        if (!m_cached) {
            m_fit.assign(pop.getIndividualCount(), 1.0);
            check(clotho_download_fitness(pop.ctx(), m_fit.data(), m_fit.size()), pop.ctx(),
                  "download_fitness");
            m_cached = true;
        }
        return m_fit;
    }
    std::shared_ptr<ifitness_generator> generator() const { return m_gen; }

private:
    std::shared_ptr<ifitness_generator> m_gen;
    std::vector<double> m_fit;
    bool m_cached = false;
};

// ---------------------------------------------------------------------------
// generic families of the reference's base engine (QS/qtlsim_engine.hpp:123-160):
// the genetic space is a device_population, whose context holds the parent
// population (current) and produces the offspring population in place.
// ---------------------------------------------------------------------------

// Crossover<RNG, GS>::update(parents, mate_pairs, offspring)
// (DS/crossover/crossover.hpp:49-72): offspring i takes a gamete of each
// individual of mate pair i.  Runs one device generation with those pairs
// in place of the fitness-proportional draw (clotho_set_parents); events and
// strand swaps are drawn on the device as in every generation.  parents and
// offspring name one context.  A context configured with host-directed
// mutations (mutation.host_directed) adds none of its own, so the
// MutationGenerator that follows places the generation's mutations.
template <class RNG, class GeneticSpaceType>
class Crossover {
public:
    typedef GeneticSpaceType genetic_space_type;
    typedef uint32_t individual_id_type;
    typedef std::vector<std::pair<individual_id_type, individual_id_type>> mate_pair_type;

    Crossover(RNG* rng, property_tree& /*config*/) : m_rng(rng) {}

    void update(genetic_space_type* parental, mate_pair_type& parents,
                genetic_space_type* offspring) {
        if (parental != offspring)
            throw std::runtime_error("Crossover::update: one context holds parents and offspring");
        clotho_ctx* c = parental->ctx();
        std::vector<uint32_t> flat(2 * parents.size());
        for (size_t i = 0; i < parents.size(); ++i) {
            flat[2 * i] = parents[i].first;
            flat[2 * i + 1] = parents[i].second;
        }
        check(clotho_set_population_size(c, (uint32_t)parents.size()), c, "set_population_size");
        check(clotho_set_parents(c, flat.data(), (uint32_t)flat.size()), c, "set_parents");
        check(clotho_run(c, 1, 0), c, "clotho_run");
    }

private:
    RNG* m_rng;
};

// MutationGenerator<RNG, GS>::operator()(child, N)
// (DS/mutation/mutation_generator.hpp:41-52): N new alleles, each at a row
// uniform on [0, sequence_count), in the next free slot, with an allele from
// AlleleGenerator (position uniform on [0, 1) as k 2^-32; per trait neutral
// with probability neutral.p, else Normal(traits.mean, traits.sigma), rounded
// to the effect quantum 2^traits.quantum_log2 as every effect on the device
// is; DS/phenotype_evaluator/trait_space_generator.hpp:118-129).  Draws come
// from the caller's engine in the reference's order (row, position, traits);
// the device places them (clotho_add_mutations) and recomputes traits and
// free space.
template <class RNG, class GeneticSpaceType>
class MutationGenerator {
public:
    typedef GeneticSpaceType genetic_space_type;
    MutationGenerator(RNG* rng, property_tree& config)
        : m_rng(rng),
          m_neutral(config.get("neutral.p", 0.5)),
          m_mean(config.get("traits.mean", 0.0)),
          m_sigma(config.get("traits.sigma", 1.0)),
          m_q((int)config.get("traits.quantum_log2", -24.0)) {}

    void operator()(genetic_space_type* child, size_t N) {
        const uint32_t H = (uint32_t)child->sequence_count(), T = child->trait_count();
        if (N == 0) return;
        if (H == 0) throw std::runtime_error("MutationGenerator: empty population");
        std::uniform_int_distribution<uint32_t> row(0u, H - 1), key(0u, 0xFFFFFFFFu);
        std::bernoulli_distribution neutral(m_neutral);
        std::normal_distribution<double> weight(m_mean, m_sigma);
        const double scale = std::ldexp(1.0, -m_q), quantum = std::ldexp(1.0, m_q);
        m_rows.resize(N);
        m_positions.resize(N);
        m_effects.assign(N * T, 0.0);
        for (size_t k = 0; k < N; ++k) {
            m_rows[k] = row(*m_rng);
            m_positions[k] = (double)key(*m_rng) * 0x1p-32;
            for (uint32_t t = 0; t < T; ++t) {
                double w = 0.0;
                if (!neutral(*m_rng)) {
                    w = std::nearbyint(weight(*m_rng) * scale);
                    w = std::min(std::max(w, -1073741823.0), 1073741823.0) * quantum;
                }
                m_effects[k * T + t] = w;
            }
        }
        m_slots.assign(N, 0);
        check(clotho_add_mutations(child->ctx(), m_rows.data(), m_positions.data(),
                                   m_effects.data(), (uint32_t)N, m_slots.data()),
              child->ctx(), "add_mutations");
    }
    // the draws and slots of the last call
    const std::vector<uint32_t>& rows() const { return m_rows; }
    const std::vector<double>& positions() const { return m_positions; }
    const std::vector<double>& effects() const { return m_effects; }
    const std::vector<uint32_t>& slots() const { return m_slots; }

private:
    RNG* m_rng;
    double m_neutral, m_mean, m_sigma;
    int m_q;
    std::vector<uint32_t> m_rows, m_slots;
    std::vector<double> m_positions, m_effects;
};

// TraitWeightAccumulator<GS> (DS/phenotype_evaluator/trait_accumulator.hpp:67-84):
// per-genome trait weights, here the device's exact sums over each row.
template <class GeneticSpaceType>
class TraitWeightAccumulator {
public:
    typedef GeneticSpaceType genetic_space_type;
    typedef const double* trait_iterator;
    void update(genetic_space_type& gs) {
        m_T = gs.trait_count();
        m_weights = gs.genome_traits();
    }
    size_t trait_count() const { return m_T; }
    size_t genome_count() const { return m_T ? m_weights.size() / m_T : 0; }
    trait_iterator begin_trait_weight(size_t genome) const { return m_weights.data() + genome * m_T; }
    trait_iterator end_trait_weight(size_t genome) const {
        return m_weights.data() + (genome + 1) * m_T;
    }

private:
    std::vector<double> m_weights;
    size_t m_T = 0;
};

// phenotype_evaluator<linear_combination<TraitWeightAccumulator<GS>, double>>
// (DS/phenotype_evaluator/phenotype_evaluator.hpp:52-76,
// DS/phenotype_evaluator/linear_combination_method.hpp:50-63): individual i's
// phenotype is the sum of its two genomes' trait weights.
template <class TraitAccumType>
class phenotype_evaluator {
public:
    typedef TraitAccumType trait_accumulator_type;
    typedef typename TraitAccumType::genetic_space_type genetic_space_type;
    typedef const double* phenotype_type;
    void update(genetic_space_type* pop, trait_accumulator_type& accum) {
        m_T = accum.trait_count();
        m_N = pop->individual_count();
        m_pheno.assign(m_N * m_T, 0.0);
        for (size_t i = 0; i < m_N; ++i)
            for (size_t t = 0; t < m_T; ++t)
                m_pheno[i * m_T + t] =
                    accum.begin_trait_weight(2 * i)[t] + accum.begin_trait_weight(2 * i + 1)[t];
    }
    size_t phenotype_count() const { return m_N; }
    phenotype_type begin_phenotype(size_t i) const { return m_pheno.data() + i * m_T; }
    phenotype_type end_phenotype(size_t i) const { return m_pheno.data() + (i + 1) * m_T; }

private:
    std::vector<double> m_pheno;
    size_t m_N = 0, m_T = 0;
};

template <class Evaluator>
class Fitness;

// Fitness<phenotype_evaluator<...>>::update(pop, eval)
// (DS/fitness/phenotype_evaluator_fitness.hpp:51-69): the fitness of every
// individual with the generator fitness.name selects, evaluated on the device
// from the population's traits (clotho_evaluate_fitness: the same metric over
// the same phenotype, plus the environmental noise the context adds).
template <class TraitAccumType>
class Fitness<phenotype_evaluator<TraitAccumType>> {
public:
    typedef phenotype_evaluator<TraitAccumType> evaluator_type;
    typedef typename evaluator_type::genetic_space_type genetic_space_type;
    explicit Fitness(property_tree& config)
        : m_fit_gen(fitness_toolkit::getInstance()->get_tool(config)) {
        if (!m_fit_gen) fitness_toolkit::getInstance()->tool_configurations(config);
    }
    void update(genetic_space_type* pop, evaluator_type& eval) {
        const size_t N = eval.phenotype_count();
        if (N == 0) return;
        if (N != pop->individual_count())
            throw std::runtime_error("Fitness::update: phenotypes of another population");
        check(clotho_evaluate_fitness(pop->ctx()), pop->ctx(), "evaluate_fitness");
    }
    std::shared_ptr<ifitness_generator> generator() const { return m_fit_gen; }

private:
    std::shared_ptr<ifitness_generator> m_fit_gen;
};

// FreeSpaceAnalyzer<GS, Size>::update(gs) (DS/free_space/free_space.hpp:58-73):
// the fixed and lost slots of the genetic space's current rows, through the
// free_space_details accessors.  Every operation that changes the rows (a
// generation, an upload, clotho_add_mutations) leaves the device analysis of
// the new rows in the context; update reads it.
template <class GeneticSpaceType, class SizeType = unsigned int>
class FreeSpaceAnalyzer : public free_space_details<SizeType> {
public:
    typedef GeneticSpaceType genetic_space_type;
    FreeSpaceAnalyzer() = default;
    explicit FreeSpaceAnalyzer(genetic_space_type& gs) { update(gs); }
    void update(genetic_space_type& gs) { this->refresh(gs); }
};

// free_space_evaluator<Space, Size>::operator()(space, res, fixed_offset,
// lost_offset, free_count, M) (DS/free_space/free_space_evaluator_def.hpp:20-21;
// layout of the row-grouped instance, DS/free_space/row_grouped_free_evaluator.hpp:140-196):
// word by word, the word's fixed slots then its lost slots below M are
// appended at res[free_count++], and each also at res[fixed_offset++] or
// res[lost_offset++].  Slots between the allele count and M are lost (no row
// carries them).  Space is a device_population.
template <class SpaceType, class SizeType = unsigned int>
class free_space_evaluator {
public:
    typedef SizeType size_type;
    typedef size_type* result_type;
    void operator()(SpaceType& ss, result_type res, size_type& fixed_offset,
                    size_type& lost_offset, size_type& free_count, size_type M) {
        free_space_details<size_type> d;
        d.refresh(ss);
        const uint32_t A = ss.shape().alleles;
        std::vector<uint32_t> fx(d.fixed_begin(), d.fixed_end()), ls(d.lost_begin(), d.lost_end());
        for (uint32_t j = A; j < M; ++j) ls.push_back(j);
        size_t a = 0, b = 0;
        while (a < fx.size() || b < ls.size()) {
            const uint32_t wa = a < fx.size() ? fx[a] / 64 : ~0u, wb = b < ls.size() ? ls[b] / 64 : ~0u;
            const uint32_t w = std::min(wa, wb);
            for (; a < fx.size() && fx[a] / 64 == w; ++a)
                if (fx[a] < M) res[free_count++] = res[fixed_offset++] = fx[a];
            for (; b < ls.size() && ls[b] / 64 == w; ++b)
                if (ls[b] < M) res[free_count++] = res[lost_offset++] = ls[b];
        }
    }
};

// ---------------------------------------------------------------------------
// Engine<RNG, Real, Block, Size, mi355x_pipeline>
// ---------------------------------------------------------------------------
// Per-period population statistics: allele_frequency + margin_details
// (include/clotho/data_spaces/analysis/allele_frequency/population_space_row_allele_frequency.hpp:31-60,
//  include/clotho/data_spaces/analysis/allele_frequency/margin_details.hpp:50-104)
// over a row sample drawn like random_sample_generator
// (include/clotho/data_spaces/generators/random_sample_generator.hpp:133-160).
// The margins come from one device pass (clotho_allele_counts); the summaries
// are computed here.  Differences from the reference: the median is exact
// (the reference reports Boost's P-square estimate), the variance is the
// population variance (as Boost's), and frequency_distribution has n + 1
// entries (the reference sizes it n, so a column carried by every sampled row
// would index past its end).
class allele_frequency {
public:
    // sample_size >= H: every haplotype; otherwise a sample drawn from Philox
    // stream 8 keyed by (seed, generation, replicate)
    void evaluate(device_population& pop, uint64_t seed, uint32_t generation,
                  uint32_t sample_size) {
        const clotho_shape s = pop.shape();
        std::vector<uint32_t> sample(std::min(sample_size, s.haplotypes));
        check(clotho_sample_rows(seed, generation, pop.config().replicate, s.haplotypes,
                                 sample_size, sample.data()),
              pop.ctx(), "sample_rows");
        run(pop, s, sample.data(), (uint32_t)sample.size());
    }
    void evaluate(device_population& pop) {
        const clotho_shape s = pop.shape();
        run(pop, s, nullptr, s.haplotypes);
    }

    const std::vector<uint32_t>& column_margin() const { return m_cols; }
    const std::vector<uint32_t>& row_margin() const { return m_rows; }
    double kernel_ms() const { return m_ms; }

    // frequency_distribution[v] = number of slots carried by v sampled rows
    std::vector<uint32_t> frequency_distribution() const {
        std::vector<uint32_t> f(m_n + 1, 0);
        for (uint32_t v : m_cols) ++f[v];
        return f;
    }

    struct summary {
        double min = 0, max = 0, mean = 0, median = 0, variance = 0;
        size_t total = 0;
    };
    static summary summarize(std::vector<uint32_t> v) {
        summary r;
        r.total = v.size();
        if (v.empty()) return r;
        std::sort(v.begin(), v.end());
        double sum = 0, sq = 0;
        for (uint32_t x : v) {
            sum += x;
            sq += (double)x * x;
        }
        r.min = v.front();
        r.max = v.back();
        r.mean = sum / v.size();
        r.variance = sq / v.size() - r.mean * r.mean;
        const size_t h = v.size() / 2;
        r.median = (v.size() % 2) ? v[h] : 0.5 * ((double)v[h - 1] + v[h]);
        return r;
    }

    // the reference's stats.* keys (recordResults) as JSON; the full
    // distribution arrays are available through the accessors above
    std::string stats_json() const {
        const summary c = summarize(m_cols), r = summarize(m_rows);
        char buf[1024];
        std::snprintf(buf, sizeof(buf),
                      "{\"sequences_per_allele\": {\"min\": %.17g, \"max\": %.17g, "
                      "\"mean\": %.17g, \"median\": %.17g, \"variance\": %.17g, \"total\": %zu}, "
                      "\"alleles_per_sequence\": {\"min\": %.17g, \"max\": %.17g, "
                      "\"mean\": %.17g, \"median\": %.17g, \"variance\": %.17g, \"total\": %zu}}",
                      c.min, c.max, c.mean, c.median, c.variance, c.total, r.min, r.max, r.mean,
                      r.median, r.variance, r.total);
        return buf;
    }

private:
    void run(device_population& pop, const clotho_shape& s, const uint32_t* rows, uint32_t n) {
        m_n = n;
        m_cols.assign(s.alleles, 0);
        m_rows.assign(n, 0);
        check(clotho_allele_counts(pop.ctx(), rows, n, m_cols.data(), m_cols.size(),
                                   m_rows.data(), &m_ms),
              pop.ctx(), "allele_counts");
    }
    std::vector<uint32_t> m_cols, m_rows;
    uint32_t m_n = 0;
    double m_ms = 0;
};

// pairwise_difference<association_matrix> (include/clotho/data_spaces/analysis/
// pairwise_difference/association_matrix_pairwise_difference.hpp:40-104): the
// sample goes through a bitset (unique rows, increasing order), every pair
// a < b contributes popcount(row a xor row b), and recordResults reports min,
// max, mean, median and variance (exact median and population variance, as
// allele_frequency::summarize).
class pairwise_difference {
public:
    template <class Iterator>
    void evaluate(device_population& pop, Iterator first, Iterator last) {
        const uint32_t H = pop.shape().haplotypes;
        std::vector<bool> in(H, false);
        for (; first != last; ++first)
            if ((size_t)*first < H) in[*first] = true;
        std::vector<uint32_t> rows;
        for (uint32_t i = 0; i < H; ++i)
            if (in[i]) rows.push_back(i);
        run(pop, rows);
    }
    void evaluate(device_population& pop) {
        std::vector<uint32_t> rows(pop.shape().haplotypes);
        for (uint32_t i = 0; i < rows.size(); ++i) rows[i] = i;
        run(pop, rows);
    }
    const std::vector<uint32_t>& differences() const { return m_diff; }
    double kernel_ms() const { return m_ms; }
    allele_frequency::summary results() const { return allele_frequency::summarize(m_diff); }
    std::string stats_json() const {
        const allele_frequency::summary r = results();
        char buf[256];
        std::snprintf(buf, sizeof(buf),
                      "{\"min\": %.17g, \"max\": %.17g, \"mean\": %.17g, \"median\": %.17g, "
                      "\"variance\": %.17g}",
                      r.min, r.max, r.mean, r.median, r.variance);
        return buf;
    }

private:
    void run(device_population& pop, const std::vector<uint32_t>& rows) {
        const size_t n = rows.size();
        m_diff.assign(n > 1 ? n * (n - 1) / 2 : 0, 0);
        check(clotho_pairwise_differences(pop.ctx(), rows.data(), (uint32_t)n, m_diff.data(),
                                          m_diff.size(), &m_ms),
              pop.ctx(), "pairwise_differences");
    }
    std::vector<uint32_t> m_diff;
    double m_ms = 0;
};

// ---------------------------------------------------------------------------
struct mi355x_pipeline {};

template <class RNG, class RealType, class BlockType, class SizeType, class Mode>
class Engine;

template <class RNG, class RealType, class BlockType, class SizeType>
class Engine<RNG, RealType, BlockType, SizeType, mi355x_pipeline> {
public:
    typedef device_population sequence_space_type;
    typedef allele_space_view allele_type;
    typedef free_space_accumulator_mt<BlockType, SizeType> free_space_type;
    typedef common_crossover<RNG, device_population, allele_type> crossover_type;

    Engine(RNG* rng, property_tree& config)
        : m_rng(rng),
          m_pop(make_config(config), device_of(config)),
          m_fit(config),
          m_free(1),
          m_xover(rng, &m_alleles, config.get("recombination.rho", 1e-4),
                  config.get("recombination.sequence_bias.p", 0.5)),
          m_growth(population_growth_toolkit::getInstance()->get_tool(config)),
          m_generation(1) {}

    // pN of the coming generation: growth(parent N, generation)
    // (QS/engine_parallel_pipeline.hpp:181-184); the parent size without a model
    uint32_t next_population_size() {
        const uint32_t parent = m_pop.getIndividualCount();
        return m_growth ? (*m_growth)(parent, (uint32_t)m_generation) : parent;
    }

    // one generation: draw, recombine+mutate, traits+fitness, free space,
    // commit; with k intermediate generations one batch of k + 1 generations
    // (the batched engine's simulate, QS/engine_parallel_pipeline_batched.hpp:198-288)
    void simulate() {
        clotho_ctx* c = m_pop.ctx();
        check(clotho_set_population_size(c, next_population_size()), c, "set_population_size");
        if (m_pop.config().intermediates) {
            check(clotho_run(c, m_pop.config().intermediates + 1, 0), c, "clotho_run");
            ++m_generation;
            return;
        }
        check(clotho_phase_draw(c), c, "phase_draw");
        m_xover(m_pop);
        m_fit(m_pop);
        check(clotho_phase_free_space(c), c, "phase_free_space");
        check(clotho_phase_commit(c), c, "phase_commit");
        ++m_generation;
    }
    // many generations (batches of k + 1 with k intermediates) without host
    // round trips (hipGraph replay when asked); a growth model splits them
    // into runs of equal population size
    void simulate(unsigned int generations, bool graph) {
        clotho_ctx* c = m_pop.ctx();
        uint32_t parent = m_pop.getIndividualCount();
        while (generations) {
            const uint32_t n =
                m_growth ? (*m_growth)(parent, (uint32_t)m_generation) : parent;
            unsigned int run = 1;
            while (m_growth && run < generations &&
                   (*m_growth)(n, (uint32_t)(m_generation + run)) == n)
                ++run;
            if (!m_growth) run = generations;
            check(clotho_set_population_size(c, n), c, "set_population_size");
            check(clotho_run(c, run * (m_pop.config().intermediates + 1), graph ? 1 : 0), c,
                  "clotho_run");
            m_generation += run;
            generations -= run;
            parent = n;
        }
    }
    // pick the recombination launch shape for the resident population
    unsigned int autotune() {
        uint32_t ch = 0;
        check(clotho_autotune(m_pop.ctx(), &ch), m_pop.ctx(), "clotho_autotune");
        return ch;
    }
    size_t getGeneration() const { return m_generation; }
    sequence_space_type* getChildPopulation() { return &m_pop; }
    allele_type* getAlleleSpace() {
        const clotho_shape s = m_pop.shape();
        m_alleles.positions.assign(s.alleles, 0.0);
        m_alleles.ages.assign(s.alleles, 0);
        m_alleles.effects.assign((size_t)s.alleles * m_pop.config().trait_count, 0.0);
        check(clotho_download_alleles(m_pop.ctx(), m_alleles.positions.data(),
                                      m_alleles.effects.data(), m_alleles.ages.data(), s.alleles),
              m_pop.ctx(), "download_alleles");
        return &m_alleles;
    }
    free_space_type* getFreeSpace() {
        m_free.refresh(m_pop);
        return &m_free;
    }
    const std::vector<double>& getFitness() { return m_fit.values(m_pop); }
    // performance.* of the reference, as per-phase milliseconds of the last
    // simulate(n, false) run with timing enabled
    std::map<std::string, double> getPerformanceResults() {
        double ms[6] = {0, 0, 0, 0, 0, 0};
        const int n = clotho_timings(m_pop.ctx(), ms, 6);
        std::map<std::string, double> out;
        const char* names[6] = {"generation", "draw", "crossover", "phenotype", "fixed",
                                "fitness"};
        for (int i = 0; i < 6; ++i) out[std::string("performance.") + names[i]] = n ? ms[i] / n : 0;
        return out;
    }

private:
    RNG* m_rng;
    device_population m_pop;
    allele_type m_alleles;
    GeneralFitness m_fit;
    free_space_type m_free;
    crossover_type m_xover;
    std::shared_ptr<ipopulation_growth> m_growth;
    size_t m_generation;
};

}  // namespace clotho_amd

#endif  // CLOTHO_AMD_ENGINE_HPP_
