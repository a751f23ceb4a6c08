// engine_demo.cpp — drives Engine<..., mi355x_pipeline> through the C++
// adapters the way examples/QTLSim/qtlsim_main.cpp:89-196 drives the
// reference engine: read the configuration keys, construct the engine, call
// simulate() per generation, report free/variable/fixed counts and the
// per-phase performance map.  Prints one JSON object.
//
//   engine_demo [-c|--config FILE.json] [key=value ...] [--generations G] [--graph]
//               [--stats K] [-p|--prefix|--log-prefix P] [--registry [--schedule K]]
//               [--print-config]
//
// -c reads a JSON configuration the way the reference driver does
// (QS/qtlsim_common.cpp:17-23, "configuration" block optional); key=value
// pairs given after it override single keys (a value starting with '[' or
// '{' is read as JSON).  --registry lists the fitness and population-growth
// toolkits and the parsed sizes (--schedule K adds the population sizes of
// the founders and generations 1..K) without touching a GPU; --print-config
// prints the resolved configuration with every default and the toolkit
// defaults, as the reference's driver does when run without a configuration
// (QS/main.cpp:113-127).
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>

#include "clotho_amd/engine.hpp"
#include "clotho_amd/log_writer.hpp"

// The flat configuration as the reference writes it: nested objects, with
// "<key>.<i>" elements gathered into arrays in index order.
inline clotho_amd::json_tree config_json(const clotho_amd::property_tree& config,
                                         const std::string& root) {
    auto split = [](const std::string& k) {
        std::vector<std::string> parts;
        size_t b = 0;
        while (true) {
            const size_t e = k.find('.', b);
            parts.push_back(k.substr(b, e == std::string::npos ? e : e - b));
            if (e == std::string::npos) return parts;
            b = e + 1;
        }
    };
    auto numeric = [](const std::string& p) {
        return !p.empty() && p.find_first_not_of("0123456789") == std::string::npos;
    };
    std::vector<std::vector<std::string>> keys;
    std::vector<std::string> values;
    for (const auto& kv : config.items()) {
        keys.push_back(split(kv.first));
        values.push_back(kv.second);
    }
    std::vector<size_t> order(keys.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const auto &x = keys[a], &y = keys[b];
        for (size_t i = 0; i < std::min(x.size(), y.size()); ++i) {
            if (x[i] == y[i]) continue;
            if (numeric(x[i]) && numeric(y[i])) return std::stoul(x[i]) < std::stoul(y[i]);
            return x[i] < y[i];
        }
        return x.size() < y.size();
    });
    clotho_amd::json_tree out;
    for (size_t i : order) {
        clotho_amd::json_tree* t = root.empty() ? &out : &out.child(root);
        for (const std::string& p : keys[i]) t = numeric(p) ? &t->element(std::stoul(p)) : &t->child(p);
        t->set_value(values[i]);
    }
    return out;
}

// Sample size of log.sampling_size entry i (qtl_logging_parameter,
// examples/qtl/qtl_logging_parameter.hpp:38-62: an object's "size", default
// 100, or a bare number); -1 past the last entry.  "log.sampling.<i>.size" is
// accepted as well.
inline double sample_size(const clotho_amd::property_tree& config, unsigned i) {
    const std::string k = "log.sampling_size." + std::to_string(i);
    if (config.has_child(k)) return config.has(k) ? config.get(k, 0.0) : config.get(k + ".size", 100.0);
    return config.get("log.sampling." + std::to_string(i) + ".size", -1.0);
}

// The reference driver's generation loop with logging (QS/qtlsim_main.cpp:
// 122-192): .config once, a .status record every log.period generations
// (allele_frequency of all rows, sample.<i> for log.sampling.<i>.size, and
// population_data with the free space, allele space and fitness), and a
// .performance record with the per-generation simulate times, their start
// and stop stamps, the statistics times and the free/variable/fixed counts.
template <class Engine>
void run_logged(Engine& engine, clotho_amd::property_tree& config, unsigned generations,
                const std::string& prefix) {
    using clock = std::chrono::steady_clock;
    auto ns = [](clock::time_point t) {
        return (long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
                   t.time_since_epoch())
            .count();
    };
    clotho_amd::log_writer(prefix, ".config").write(config_json(config, "configuration"));
    clotho_amd::log_writer status(prefix, ".status");
    const unsigned period0 = (unsigned)config.get("log.period", (double)generations);
    unsigned period = std::min(period0 ? period0 : generations, generations);
    std::vector<long long> sim, starts, stops, stats;
    std::vector<unsigned> free_c, var_c, fixed_c;
    const auto run0 = clock::now();
    for (unsigned left = generations; left-- > 0;) {
        const auto t0 = clock::now();
        engine.simulate();
        const auto t1 = clock::now();
        sim.push_back(std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count());
        starts.push_back(ns(t0));
        stops.push_back(ns(t1));
        auto* fs = engine.getFreeSpace();
        free_c.push_back((unsigned)fs->free_size());
        var_c.push_back((unsigned)fs->variable_count());
        fixed_c.push_back((unsigned)fs->fixed_size());
        if (!--period) {
            const auto s0 = clock::now();
            clotho_amd::json_tree log;
            auto* pop = engine.getChildPopulation();
            const clotho_shape sh = pop->shape();
            auto record = [&](clotho_amd::allele_frequency& af, const std::string& at) {
                const auto& cm = af.column_margin();
                log.at(at).add_value_array("distribution", cm.begin(), cm.end());
                const auto fd = af.frequency_distribution();
                log.at(at).add_value_array("frequency_distribution", fd.begin(), fd.end());
                const auto c = clotho_amd::allele_frequency::summarize(cm);
                const auto r = clotho_amd::allele_frequency::summarize(af.row_margin());
                const char* names[2] = {"sequences_per_allele", "alleles_per_sequence"};
                const clotho_amd::allele_frequency::summary* sm[2] = {&c, &r};
                for (int q = 0; q < 2; ++q) {
                    const std::string b = at + ".stats." + names[q] + ".";
                    log.put(b + "min", sm[q]->min);
                    log.put(b + "max", sm[q]->max);
                    log.put(b + "mean", sm[q]->mean);
                    log.put(b + "median", sm[q]->median);
                    log.put(b + "variance", sm[q]->variance);
                    log.put(b + "total", sm[q]->total);
                }
            };
            clotho_amd::allele_frequency all;
            all.evaluate(*pop);
            record(all, "allele_frequency");
            for (unsigned i = 0;; ++i) {
                const double k = sample_size(config, i);
                if (k < 0) break;
                if (k == 0) continue;
                clotho_amd::allele_frequency af;
                const uint64_t seed = (uint64_t)config.get("random_number.seed", 0.0);
                af.evaluate(*pop, seed, sh.generation, (uint32_t)k);
                record(af, "sample." + std::to_string(i) + ".allele_frequency");
                // sample_log_params::pairwise (examples/qtl/qtl_logging_parameter.hpp:31-62):
                // the pairwise differences of the same sample
                if (config.get("log.sampling_size." + std::to_string(i) + ".pairwise", 0.0) != 0.0) {
                    std::vector<uint32_t> rows(std::min<uint32_t>((uint32_t)k, sh.haplotypes));
                    clotho_amd::check(clotho_sample_rows(seed, sh.generation, pop->config().replicate,
                                                         sh.haplotypes, (uint32_t)k, rows.data()),
                                      pop->ctx(), "sample_rows");
                    clotho_amd::pairwise_difference pd;
                    pd.evaluate(*pop, rows.begin(), rows.end());
                    const auto r = pd.results();
                    const std::string b = "sample." + std::to_string(i) + ".pairwise_difference.";
                    log.put(b + "min", r.min);
                    log.put(b + "max", r.max);
                    log.put(b + "mean", r.mean);
                    log.put(b + "median", r.median);
                    log.put(b + "variance", r.variance);
                }
            }
            log.at("population_data.free_space")
                .add_value_array("free", fs->free_begin(), fs->free_end());
            log.put("population_data.free_space.free_size", fs->free_size());
            log.put("population_data.free_space.variable_size", fs->variable_count());
            auto* alleles = engine.getAlleleSpace();
            clotho_amd::json_tree& as = log.at("population_data.allele_space");
            for (size_t j = 0; j < alleles->positions.size(); ++j) {
                clotho_amd::json_tree a;
                a.put("position", alleles->positions[j]);
                a.put("age", alleles->ages[j]);
                as.push_back(a);
            }
            const auto& fit = engine.getFitness();
            log.at("population_data.child").add_value_array("fitness", fit.begin(), fit.end());
            status.write(log);
            stats.push_back(std::chrono::duration_cast<std::chrono::nanoseconds>(clock::now() - s0)
                                .count());
            period = std::min(period0 ? period0 : generations, left);
        }
    }
    clotho_amd::json_tree perf;
    perf.put("performance.runtime",
             std::chrono::duration_cast<std::chrono::nanoseconds>(clock::now() - run0).count());
    perf.at("performance").add_value_array("simulate", sim.begin(), sim.end());
    perf.at("performance.generations").add_value_array("start", starts.begin(), starts.end());
    perf.at("performance.generations").add_value_array("stop", stops.begin(), stops.end());
    perf.at("performance").add_value_array("stats", stats.begin(), stats.end());
    perf.at("memory").add_value_array("free_count", free_c.begin(), free_c.end());
    perf.at("memory").add_value_array("variable_count", var_c.begin(), var_c.end());
    perf.at("memory").add_value_array("fixed_count", fixed_c.begin(), fixed_c.end());
    clotho_amd::log_writer(prefix, ".performance").write(perf);
}

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
    bool graph = false, registry = false, print_config = false;
    long schedule = -1;  // --schedule K with --registry
    long stats_sample = -1;  // --stats K: allele-frequency statistics over K rows (0 = all)
    std::string log_prefix;  // --log-prefix P: the reference's .config/.status/.performance files
    bool logging = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--registry") {
            registry = true;
        } else if (a == "--print-config") {
            print_config = true;
        } else if (a == "--schedule" && i + 1 < argc) {
            schedule = std::stol(argv[++i]);
        } else if (a == "--graph") {
            graph = true;
        } else if ((a == "-c" || a == "--config") && i + 1 < argc) {
            clotho_amd::read_json(argv[++i], config);
            if (config.has("generations.size"))
                generations = (unsigned)config.get("generations.size", (double)generations);
        } else if ((a == "--log-prefix" || a == "-p" || a == "--prefix") && i + 1 < argc) {
            log_prefix = argv[++i];
            logging = true;
        } else if (a == "--stats" && i + 1 < argc) {
            stats_sample = std::stol(argv[++i]);
        } else if (a == "--generations" && i + 1 < argc) {
            generations = (unsigned)std::stoul(argv[++i]);
        } else if (a.find('=') != std::string::npos) {
            const std::string key = a.substr(0, a.find('=')), value = a.substr(a.find('=') + 1);
            if (!value.empty() && (value[0] == '[' || value[0] == '{')) {
                config.erase(key);
                clotho_amd::read_json_text(value, config, key);
            } else {
                config.put(key, value);
            }
            if (key == "generations.size" || key == "configuration.generations.size")
                generations = (unsigned)config.get("generations.size", (double)generations);
        }
    }
    config.put("generations.size", (double)generations);
    if (print_config) {
        (void)clotho_amd::make_config(config);  // completes every parameter block
        clotho_amd::fitness_toolkit::getInstance()->tool_configurations(config);
        clotho_amd::population_growth_toolkit::getInstance()->tool_configurations(config);
        clotho_amd::log_writer("", "").write(config_json(config, "configuration"));
        return 0;
    }
    if (registry) {
        std::printf("{\"fitness_tools\": [");
        const auto names = clotho_amd::fitness_toolkit::getInstance()->tool_names();
        for (size_t i = 0; i < names.size(); ++i)
            std::printf("%s\"%s\"", i ? ", " : "", names[i].c_str());
        std::printf("], \"population_growth_tools\": [");
        const auto gnames = clotho_amd::population_growth_toolkit::getInstance()->names();
        for (size_t i = 0; i < gnames.size(); ++i)
            std::printf("%s\"%s\"", i ? ", " : "", gnames[i].c_str());
        const clotho_config c = clotho_amd::make_config(config);
        std::printf("], \"n\": %u, \"max_individuals\": %u, \"mu\": %.17g, \"rho\": %.17g, "
                    "\"fitness_kind\": %u, \"seed\": %llu",
                    c.n_individuals, c.max_individuals, c.mu, c.rho, c.fitness_kind,
                    (unsigned long long)c.seed);
        if (schedule >= 0) {
            auto growth = clotho_amd::population_growth_toolkit::getInstance()->get_tool(config);
            std::printf(", \"schedule\": [");
            if (growth) {
                const auto sizes = clotho_amd::growth_schedule(*growth, (uint32_t)schedule);
                for (size_t i = 0; i < sizes.size(); ++i)
                    std::printf("%s%u", i ? ", " : "", sizes[i]);
            }
            std::printf("]");
        }
        std::printf("}\n");
        return 0;
    }
    std::mt19937 rng(0);
    typedef clotho_amd::Engine<std::mt19937, double, uint64_t, unsigned int,
                               clotho_amd::mi355x_pipeline>
        engine_type;
    engine_type engine(&rng, config);
    if (logging) {
        run_logged(engine, config, generations, log_prefix);
    } else if (graph) {
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
