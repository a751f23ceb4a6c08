// generic_demo.cpp — drives the generic operator families through the C++
// adapters in the order of the reference's base engine
// (examples/QTLSim/qtlsim_engine.hpp:123-160): per generation
//   M = mutation_allocator::allocate(2N)
//   mate pairs (random mating from the caller's engine)
//   Crossover::update(pop, pairs, pop)
//   MutationGenerator(pop, M)
//   TraitWeightAccumulator::update, phenotype_evaluator::update, Fitness::update
//   FreeSpaceAnalyzer::update
// then single gametes through common_crossover's per-gamete operator (half
// one call each, half as one batch) and their free space through
// free_space_accumulator_mt's thread buffers, and free_space_evaluator over
// the final population.  Every draw and result goes to binary files in OUTDIR
// so that a checker can replay the draws and compare.
//
//   generic_demo OUTDIR GENERATIONS SEED [key=value ...]
//   generic_demo --allocator-sample K MEAN_N [key=value ...]   (no GPU)
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "clotho_amd/engine.hpp"

namespace {

typedef clotho_amd::device_population space_type;

struct writer {
    std::FILE* f;
    explicit writer(const std::string& path) : f(std::fopen(path.c_str(), "wb")) {
        if (!f) throw std::runtime_error("cannot write " + path);
    }
    ~writer() { std::fclose(f); }
    template <class T>
    void vec(const std::vector<T>& v) {
        const uint64_t n = v.size();
        std::fwrite(&n, 8, 1, f);
        if (n) std::fwrite(v.data(), sizeof(T), n, f);
    }
    template <class T>
    void raw(const T* p, size_t n) {
        const uint64_t m = n;
        std::fwrite(&m, 8, 1, f);
        if (n) std::fwrite(p, sizeof(T), n, f);
    }
};

void parse(int argc, char** argv, int first, clotho_amd::property_tree& config) {
    for (int i = first; i < argc; ++i) {
        const std::string a = argv[i];
        const size_t eq = a.find('=');
        if (eq == std::string::npos) throw std::runtime_error("expected key=value: " + a);
        config.put(a.substr(0, eq), a.substr(eq + 1));
    }
}

}  // namespace

int main(int argc, char** argv) {
    try {
        clotho_amd::property_tree config;
        if (argc >= 4 && std::string(argv[1]) == "--allocator-sample") {
            parse(argc, argv, 4, config);
            std::mt19937 rng(7);
            clotho_amd::mutation_allocator<std::mt19937, size_t> alloc(&rng, config);
            const long k = std::stol(argv[2]);
            const size_t n = (size_t)std::stoul(argv[3]);
            for (long i = 0; i < k; ++i) std::printf("%zu\n", alloc.allocate(n));
            return 0;
        }
        if (argc < 4) {
            std::fprintf(stderr, "usage: generic_demo OUTDIR GENERATIONS SEED [key=value ...]\n");
            return 2;
        }
        const std::string out = argv[1];
        const unsigned generations = (unsigned)std::stoul(argv[2]);
        std::mt19937 rng((uint32_t)std::stoul(argv[3]));
        parse(argc, argv, 4, config);
        config.put("mutation.host_directed", 1.0);

        const clotho_config cfg = clotho_amd::make_config(config);
        space_type pop(cfg, clotho_amd::device_of(config));
        {
            writer w(out + "/config.bin");
            w.raw(&cfg, 1);
        }
        clotho_amd::mutation_allocator<std::mt19937, size_t> alloc(&rng, config);
        clotho_amd::Crossover<std::mt19937, space_type> cross(&rng, config);
        clotho_amd::MutationGenerator<std::mt19937, space_type> mutate(&rng, config);
        typedef clotho_amd::TraitWeightAccumulator<space_type> accum_type;
        typedef clotho_amd::phenotype_evaluator<accum_type> pheno_type;
        accum_type accum;
        pheno_type pheno;
        clotho_amd::Fitness<pheno_type> fit(config);
        clotho_amd::FreeSpaceAnalyzer<space_type> free_space;

        const uint32_t N = pop.getIndividualCount();
        std::uniform_int_distribution<uint32_t> mate(0u, N - 1);
        for (unsigned g = 0; g < generations; ++g) {
            const size_t M = alloc.allocate(2 * (size_t)N);
            clotho_amd::Crossover<std::mt19937, space_type>::mate_pair_type pairs(N);
            std::vector<uint32_t> flat(2 * N);
            for (uint32_t i = 0; i < N; ++i) {
                pairs[i].first = flat[2 * i] = mate(rng);
                pairs[i].second = flat[2 * i + 1] = mate(rng);
            }
            cross.update(&pop, pairs, &pop);
            mutate(&pop, M);
            accum.update(pop);
            pheno.update(&pop, accum);
            fit.update(&pop, pheno);
            free_space.update(pop);
            writer w(out + "/gen" + std::to_string(g) + ".bin");
            w.vec(flat);
            w.vec(mutate.rows());
            w.vec(mutate.positions());
            w.vec(mutate.effects());
            w.vec(mutate.slots());
            w.vec(std::vector<uint32_t>(free_space.fixed_begin(), free_space.fixed_end()));
            w.vec(std::vector<uint32_t>(free_space.lost_begin(), free_space.lost_end()));
        }

        // final population, phenotypes and the evaluator's layout
        const clotho_shape s = pop.shape();
        const std::vector<uint64_t> rows = pop.rows();
        std::vector<double> ph;
        for (size_t i = 0; i < pheno.phenotype_count(); ++i)
            ph.insert(ph.end(), pheno.begin_phenotype(i), pheno.end_phenotype(i));
        const uint32_t M_eval = s.alleles + 70;
        std::vector<uint32_t> res(3 * (size_t)M_eval, 0xFFFFFFFFu);
        uint32_t fo = M_eval, lo = 2 * M_eval, fr = 0;
        clotho_amd::free_space_evaluator<space_type> eval;
        eval(pop, res.data(), fo, lo, fr, M_eval);
        {
            writer w(out + "/final.bin");
            w.vec(rows);
            w.vec(pop.genome_traits());
            w.vec(pop.fitness());
            w.vec(ph);
            w.vec(res);
            w.vec(std::vector<uint32_t>{fo, lo, fr, M_eval, s.words, s.alleles});
        }

        // single gametes: pairs of rows of the final population, the first
        // half one device call each, the second half as one batch
        typedef clotho_amd::allele_space_view alleles_type;
        alleles_type alleles;
        clotho_amd::common_crossover<std::mt19937, space_type, alleles_type> xo(
            &rng, &alleles, config.get("recombination.rho", 1e-4),
            config.get("recombination.sequence_bias.p", 0.5));
        xo.bind(pop);
        const uint32_t W = s.words, H = s.haplotypes, G = 16, OW = W + 8;
        std::uniform_int_distribution<uint32_t> pick(0u, H - 1);
        std::vector<uint32_t> picked(2 * G);
        std::vector<uint64_t> offspring((size_t)G * OW, ~0ull);
        for (uint32_t g = 0; g < G; ++g) {
            if (g == G / 2) xo.begin_batch();
            const uint32_t a = pick(rng), b = pick(rng);
            picked[2 * g] = a;
            picked[2 * g + 1] = b;
            const uint64_t* p0 = rows.data() + (size_t)a * W;
            const uint64_t* p1 = rows.data() + (size_t)b * W;
            uint64_t* o = offspring.data() + (size_t)g * OW;
            xo(p0, p0 + W, p1, p1 + W, o, o + OW);
        }
        xo.flush();
        // their free space, two threads of eight rows each
        clotho_amd::free_space_accumulator_mt<uint64_t, uint32_t> fsa(2);
        fsa.resetBuffers(OW);
        for (uint32_t g = 0; g < G; ++g)
            fsa.getThreadBuffer(g / (G / 2)).update(offspring.data() + (size_t)g * OW, OW);
        fsa.finalize(s.alleles);
        {
            writer w(out + "/gametes.bin");
            w.vec(picked);
            std::vector<uint32_t> swaps(xo.swaps().begin(), xo.swaps().end());
            w.vec(swaps);
            std::vector<uint32_t> nev, keys;
            for (const auto& e : xo.events()) {
                nev.push_back((uint32_t)e.size());
                keys.insert(keys.end(), e.begin(), e.end());
            }
            w.vec(nev);
            w.vec(keys);
            w.vec(offspring);
            w.vec(std::vector<uint32_t>(fsa.fixed_begin(), fsa.fixed_end()));
            w.vec(std::vector<uint32_t>(fsa.lost_begin(), fsa.lost_end()));
        }
        std::printf("{\"generations\": %u, \"alleles\": %u, \"words\": %u}\n", generations,
                    s.alleles, s.words);
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "generic_demo: %s\n", e.what());
        return 1;
    }
}
