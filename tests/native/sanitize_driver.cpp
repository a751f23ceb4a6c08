// sanitize_driver.cpp — host code under AddressSanitizer and
// UndefinedBehaviorSanitizer: the oracle's whole-generation restatement
// (threads, compaction, growth, intermediate generations, host-directed
// mutations, statistics) and the configuration, growth and log-record
// headers of the C++ front end.  Built and run by tests/test_sanitizers.py;
// any sanitizer report aborts with a non-zero status.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../oracle/oracle.h"
#include "clotho_amd/config.hpp"
#include "clotho_amd/log_writer.hpp"
#include "clotho_amd/population_growth.hpp"

namespace {

clotho_config base(uint32_t n, double mu, double rho) {
    clotho_config c;
    std::memset(&c, 0, sizeof(c));
    c.seed = 42;
    c.n_individuals = n;
    c.trait_count = 2;
    c.fitness_kind = CLOTHO_FIT_MV_STABILIZING;
    c.mu = mu;
    c.rho = rho;
    c.bias = 0.5;
    c.neutral_p = 0.3;
    c.trait_sigma = 0.2;
    c.fitness_scale = 2.0;
    c.env_sigma = 0.5;
    c.effect_quantum_log2 = -24;
    c.max_alleles = 1 << 14;
    return c;
}

int run(const clotho_config& cfg, int gens, int threads, uint32_t grow_to) {
    or_sim* s = or_sim_create(&cfg);
    if (!s) return 1;
    or_sim_set_threads(s, threads);
    for (int g = 0; g < gens; ++g) {
        if (grow_to && g == gens / 2) or_sim_set_population_size(s, grow_to);
        if (or_sim_step(s) != 0) return 2;
    }
    clotho_shape sh;
    or_sim_shape(s, &sh);
    std::vector<uint64_t> rows((size_t)sh.haplotypes * sh.words + 1);
    or_sim_rows(s, rows.data());
    std::vector<uint32_t> cols(sh.alleles + 1), rc(sh.haplotypes + 1), sample(sh.haplotypes);
    for (uint32_t i = 0; i < sh.haplotypes; ++i) sample[i] = i;
    or_allele_counts(rows.data(), sh.words, sh.alleles, sample.data(), sh.haplotypes, cols.data(),
                     rc.data());
    const uint32_t k = sh.haplotypes < 40 ? sh.haplotypes : 40;
    std::vector<uint32_t> pw(k * (k - 1) / 2 + 1);
    or_pairwise_differences(rows.data(), sh.words, sample.data(), k, pw.data());
    if (!cfg.intermediates) {  // host-directed mutations, then fitness
        std::vector<uint32_t> mrows = {0, 1, 2, sh.haplotypes - 1};
        std::vector<double> pos = {0.125, 0.25, 0.375, 0.5};
        std::vector<double> eff(4 * cfg.trait_count, 0.5);
        std::vector<uint32_t> slots(4);
        if (or_sim_add_mutations(s, mrows.data(), pos.data(), eff.data(), 4, slots.data()) != 0)
            return 3;
        or_sim_evaluate_fitness(s);
    }
    or_sim_destroy(s);
    return 0;
}

}  // namespace

int main() {
    int rc = 0;
    rc |= run(base(150, 1.5, 2.0), 12, 1, 0);
    rc |= run(base(150, 1.5, 2.0), 6, 4, 260);  // threads and growth
    clotho_config c = base(120, 2.0, 3.0);
    c.compact_every = 1;
    rc |= run(c, 8, 2, 0);
    c = base(120, 1.0, 1.0);
    c.intermediates = 2;
    rc |= run(c, 9, 1, 0);
    c = base(80, 1.0, 1.0);
    c.flags = CLOTHO_FLAG_HOST_MUTATIONS;
    c.fitness_kind = CLOTHO_FIT_GAUSSIAN;
    rc |= run(c, 4, 1, 0);

    // configuration, growth and log records
    clotho_amd::property_tree cfg;
    clotho_amd::read_json_text(
        "{\"configuration\": {\"population_growth\": {\"name\": \"predefined\", "
        "\"params\": {\"sizes\": [10, 20, 35]}}, \"mutation\": {\"mu\": 0.5}, "
        "\"log\": {\"sampling_size\": [{\"size\": 5, \"pairwise\": true}, 7]}}}",
        cfg, "");
    auto growth = clotho_amd::population_growth_toolkit::getInstance()->get_tool(cfg);
    if (!growth) return 4;
    const std::vector<uint32_t> sizes = clotho_amd::growth_schedule(*growth, 7);
    clotho_amd::population_growth_toolkit::getInstance()->tool_configurations(cfg);
    clotho_amd::json_tree t;
    for (const auto& kv : cfg.items()) t.put(kv.first, kv.second);
    for (size_t i = 0; i < sizes.size(); ++i) t.at("sizes").element(i).put("n", (double)sizes[i]);
    clotho_amd::log_writer w("", "");
    w.write(t);
    std::printf("sanitized run: %s\n", rc ? "FAILED" : "ok");
    return rc;
}
