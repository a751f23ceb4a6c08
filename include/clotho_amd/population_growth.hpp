// clotho_amd/population_growth.hpp — population-growth models selected by
// "population_growth.name" with parameters under "population_growth.params",
// as the reference's population_growth_toolkit registry
// (src/clotho/genetics/population_growth_toolkit.cpp:36-62):
//
//   linear      pN' = A pN + B, converted to an unsigned count
//               (src/clotho/genetics/linear_population_growth.cpp:23-25;
//               defaults A = 1, B = 0, linear_population_growth_generator.cpp:18-74)
//   predefined  pN' = sizes[generation mod |sizes|]
//               (src/clotho/genetics/predefined_population_growth.cpp:25-27;
//               default sizes [10000], predefined_population_growth_generator.cpp:20-87)
//
// The engine calls the model as pN = growth(0, 0) for the founders and
// pN = growth(parent N, generation) before every generation
// (QS/engine_parallel_pipeline.hpp:150-153,181-184).  Generators write their
// defaults back into the configuration, as the reference's parseConfig does.
#ifndef CLOTHO_AMD_POPULATION_GROWTH_HPP_
#define CLOTHO_AMD_POPULATION_GROWTH_HPP_

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "config.hpp"

namespace clotho_amd {

struct ipopulation_growth {
    virtual ~ipopulation_growth() = default;
    virtual uint32_t operator()(uint32_t psize, uint32_t generation) = 0;
    virtual std::string name() const = 0;
    virtual void log(std::ostream& out) const = 0;
};

class linear_population_growth : public ipopulation_growth {
public:
    linear_population_growth(double a, double b) : m_A(a), m_B(b) {}
    // double -> unsigned conversion truncates toward zero; values below zero
    // (undefined in the reference) give 0 here
    uint32_t operator()(uint32_t psize, uint32_t) override {
        const double v = m_A * (double)psize + m_B;
        return v > 0.0 ? (uint32_t)v : 0u;
    }
    std::string name() const override { return "linear"; }
    void log(std::ostream& out) const override { out << "{linear," << m_A << "," << m_B << "}"; }

private:
    double m_A, m_B;
};

class predefined_population_growth : public ipopulation_growth {
public:
    explicit predefined_population_growth(std::vector<uint32_t> sizes) : m_sizes(std::move(sizes)) {
        if (m_sizes.empty()) m_sizes.push_back(10000);
    }
    uint32_t operator()(uint32_t, uint32_t generation) override {
        return m_sizes[generation % m_sizes.size()];
    }
    std::string name() const override { return "predefined"; }
    void log(std::ostream& out) const override {
        out << "{predefined";
        for (uint32_t v : m_sizes) out << "," << v;
        out << "}";
    }

private:
    std::vector<uint32_t> m_sizes;
};

// ipopulation_growth_generator: reads (and completes) the parameter block at
// `prefix` and builds the model.
struct ipopulation_growth_generator {
    virtual ~ipopulation_growth_generator() = default;
    virtual std::string name() const = 0;
    virtual std::shared_ptr<ipopulation_growth> create(property_tree& config,
                                                       const std::string& prefix) const = 0;
};

struct linear_population_growth_generator : ipopulation_growth_generator {
    std::string name() const override { return "linear"; }
    std::shared_ptr<ipopulation_growth> create(property_tree& c, const std::string& p) const override {
        if (!c.has(p + ".A")) c.put(p + ".A", 1.0);
        if (!c.has(p + ".B")) c.put(p + ".B", 0.0);
        return std::make_shared<linear_population_growth>(c.get(p + ".A", 1.0), c.get(p + ".B", 0.0));
    }
};

struct predefined_population_growth_generator : ipopulation_growth_generator {
    std::string name() const override { return "predefined"; }
    std::shared_ptr<ipopulation_growth> create(property_tree& c, const std::string& p) const override {
        std::vector<uint32_t> sizes;
        for (double v : c.get_array(p + ".sizes")) sizes.push_back(v > 0.0 ? (uint32_t)v : 0u);
        if (sizes.empty()) {
            sizes.push_back(10000);
            c.put_array(p + ".sizes", {10000.0});
        }
        return std::make_shared<predefined_population_growth>(sizes);
    }
};

class population_growth_toolkit {
public:
    static population_growth_toolkit* getInstance() {
        static population_growth_toolkit inst;
        return &inst;
    }
    // The model named by population_growth.name, or nullptr when the block or
    // the name is missing or unknown (the reference then writes an empty
    // block back and the engine keeps the parent size).
    std::shared_ptr<ipopulation_growth> get_tool(property_tree& config) const {
        if (!config.has_child("population_growth")) {
            config.put("population_growth.name", std::string(""));
            config.put("population_growth.params", std::string(""));
            return nullptr;
        }
        const std::string name = config.get("population_growth.name", std::string(""));
        auto it = m_tools.find(name);
        if (name.empty() || it == m_tools.end()) return nullptr;
        return it->second->create(config, "population_growth.params");
    }
    // population_growth.toolkit.<name>.{name, params} with every default
    void tool_configurations(property_tree& config) const {
        for (const auto& kv : m_tools) {
            const std::string p = "population_growth.toolkit." + kv.first;
            config.put(p + ".name", kv.first);
            property_tree defaults;
            kv.second->create(defaults, "params");
            for (const auto& d : defaults.items()) config.put(p + "." + d.first, d.second);
        }
    }
    std::vector<std::string> names() const {
        std::vector<std::string> out;
        for (const auto& kv : m_tools) out.push_back(kv.first);
        return out;
    }

private:
    population_growth_toolkit() {
        add(std::make_shared<linear_population_growth_generator>());
        add(std::make_shared<predefined_population_growth_generator>());
    }
    void add(std::shared_ptr<ipopulation_growth_generator> g) { m_tools[g->name()] = g; }
    std::map<std::string, std::shared_ptr<ipopulation_growth_generator>> m_tools;
};

// Population sizes of the founders and of generations 1..horizon under a
// model: sizes[0] = growth(0, 0), sizes[g] = growth(sizes[g-1], g).
inline std::vector<uint32_t> growth_schedule(ipopulation_growth& growth, uint32_t horizon) {
    std::vector<uint32_t> sizes(horizon + 1);
    sizes[0] = growth(0, 0);
    for (uint32_t g = 1; g <= horizon; ++g) sizes[g] = growth(sizes[g - 1], g);
    return sizes;
}

}  // namespace clotho_amd

#endif  // CLOTHO_AMD_POPULATION_GROWTH_HPP_
