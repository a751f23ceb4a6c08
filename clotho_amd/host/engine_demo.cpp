// engine_demo.cpp — drives Engine<..., mi355x_pipeline> through the C++
// adapters the way examples/QTLSim/qtlsim_main.cpp:89-196 drives the
// reference engine: read the configuration keys, construct the engine, call
// simulate() per generation, report free/variable/fixed counts and the
// per-phase performance map.  Prints one JSON object.
//
//   engine_demo [--registry] [key=value ...] [--generations G] [--graph] [--stats K]
//
// --registry lists the fitness_toolkit generators and the parsed
// configuration without touching a GPU.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "clotho_amd/engine.hpp"

int main(int argc, char** argv) {
    clotho_amd::property_tree config;
    // reference defaults (src/clotho/utility/clotho_strings.cpp keys)
    config.put("random_number.seed", 42.0);
    config.put("mutation.mu", 1.0);
    config.put("recombination.rho", 1.0);
    config.put("recombination.sequence_bias.p", 0.5);
    config.put("neutral.p", 1.0);
    config.put("traits.N", 1.0);
    config.put("fitness.name", std::string("constant"));
    config.put("population_growth.name", std::string("linear"));
    config.put("population_growth.params.A", 0.0);
    config.put("population_growth.params.B", 1000.0);
    config.put("alleles.capacity", 32768.0);
    unsigned generations = 10;
    bool graph = false, registry = false;
    long stats_sample = -1;  // --stats K: allele-frequency statistics over K rows (0 = all)
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--registry") {
            registry = true;
        } else if (a == "--graph") {
            graph = true;
        } else if (a == "--stats" && i + 1 < argc) {
            stats_sample = std::stol(argv[++i]);
        } else if (a == "--generations" && i + 1 < argc) {
            generations = (unsigned)std::stoul(argv[++i]);
        } else if (a.find('=') != std::string::npos) {
            config.put(a.substr(0, a.find('=')), a.substr(a.find('=') + 1));
        }
    }
    if (registry) {
        std::printf("{\"fitness_tools\": [");
        const auto names = clotho_amd::fitness_toolkit::getInstance()->tool_names();
        for (size_t i = 0; i < names.size(); ++i)
            std::printf("%s\"%s\"", i ? ", " : "", names[i].c_str());
        const clotho_config c = clotho_amd::make_config(config);
        std::printf("], \"n\": %u, \"mu\": %.17g, \"rho\": %.17g, \"fitness_kind\": %u, "
                    "\"seed\": %llu}\n",
                    c.n_individuals, c.mu, c.rho, c.fitness_kind, (unsigned long long)c.seed);
        return 0;
    }
    std::mt19937 rng(0);
    typedef clotho_amd::Engine<std::mt19937, double, uint64_t, unsigned int,
                               clotho_amd::mi355x_pipeline>
        engine_type;
    engine_type engine(&rng, config);
    if (graph) {
        engine.simulate(generations, true);
    } else {
        for (unsigned g = 0; g < generations; ++g) engine.simulate();
    }
    const clotho_shape s = engine.getChildPopulation()->shape();
    const auto rows = engine.getChildPopulation()->rows();
    uint64_t h = 1469598103934665603ull;  // FNV-1a over the row words
    for (uint64_t w : rows) {
        for (int b = 0; b < 8; ++b) {
            h ^= (w >> (8 * b)) & 0xFF;
            h *= 1099511628211ull;
        }
    }
    auto* fs = engine.getFreeSpace();
    std::string stats = "null";
    if (stats_sample >= 0) {
        clotho_amd::allele_frequency af;
        if (stats_sample == 0)
            af.evaluate(*engine.getChildPopulation());
        else
            af.evaluate(*engine.getChildPopulation(), (uint64_t)config.get("random_number.seed", 0.0),
                        s.generation, (uint32_t)stats_sample);
        const auto fd = af.frequency_distribution();
        uint64_t fh = 1469598103934665603ull;  // FNV-1a over the spectrum
        for (uint32_t v : fd)
            for (int b = 0; b < 4; ++b) {
                fh ^= (v >> (8 * b)) & 0xFF;
                fh *= 1099511628211ull;
            }
        char buf[64];
        std::snprintf(buf, sizeof(buf), "\"%016llx\"", (unsigned long long)fh);
        stats = "{\"stats\": " + af.stats_json() + ", \"frequency_distribution_fnv1a\": " + buf +
                "}";
    }
    std::printf("{\"generation\": %u, \"words\": %u, \"alleles\": %u, \"free\": %u, "
                "\"fixed\": %u, \"variable\": %u, \"rows_fnv1a\": \"%016llx\", "
                "\"allele_frequency\": %s}\n",
                s.generation, s.words, s.alleles, (unsigned)fs->free_size(),
                (unsigned)fs->fixed_size(), (unsigned)fs->variable_count(),
                (unsigned long long)h, stats.c_str());
    return 0;
}
