// clotho_amd/log_writer.hpp — the reference's JSON log files, header-only.
//
// The reference writes Boost property trees with write_json
// (examples/QTLSim/log_writer.cpp:16-31): one file per record named
// <prefix>.<index><suffix>.json (<suffix> = ".config", ".status",
// ".performance"), or stdout when the prefix is empty; every leaf is written
// as a JSON string and arrays are children with empty keys
// (clotho::utility::add_value_array).  json_tree reproduces that shape so
// the reference's log_parser (src/tools/log_parser/log_parser.cpp) reads
// these files like its own.
#pragma once

#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace clotho_amd {

class json_tree {
public:
    // child at a dotted path, created on demand
    json_tree& at(const std::string& path) {
        json_tree* t = this;
        size_t b = 0;
        while (true) {
            const size_t e = path.find('.', b);
            const std::string key = path.substr(b, e == std::string::npos ? e : e - b);
            json_tree* next = nullptr;
            for (auto& kv : t->m_children)
                if (kv.first == key) next = &kv.second;
            if (!next) {
                t->m_children.emplace_back(key, json_tree());
                next = &t->m_children.back().second;
            }
            t = next;
            if (e == std::string::npos) return *t;
            b = e + 1;
        }
    }
    void put(const std::string& path, const std::string& v) { at(path).m_value = v; }
    template <class T>
    void put(const std::string& path, T v) {
        put(path, number(v));
    }
    void put_child(const std::string& path, const json_tree& t) { at(path) = t; }
    // element i of this node as an array (""-keyed children), created on demand
    json_tree& element(size_t i) {
        while (m_children.size() <= i) m_children.emplace_back("", json_tree());
        return m_children[i].second;
    }
    // key of this node's child, created on demand (no dotted-path splitting)
    json_tree& child(const std::string& key) {
        for (auto& kv : m_children)
            if (kv.first == key) return kv.second;
        m_children.emplace_back(key, json_tree());
        return m_children.back().second;
    }
    void set_value(const std::string& v) { m_value = v; }
    void push_back(const json_tree& t) { m_children.emplace_back("", t); }
    template <class T>
    void push_value(T v) {
        json_tree leaf;
        leaf.m_value = number(v);
        push_back(leaf);
    }
    template <class It>
    void add_value_array(const std::string& path, It first, It last) {
        json_tree& t = at(path);
        for (; first != last; ++first) t.push_value(*first);
    }

    void write(std::ostream& os, int indent = 0) const {
        const std::string pad(indent, ' ');
        if (m_children.empty()) {
            os << '"' << escape(m_value) << '"';
            return;
        }
        const bool array = m_children.front().first.empty();
        os << (array ? "[\n" : "{\n");
        for (size_t i = 0; i < m_children.size(); ++i) {
            os << pad << "    ";
            if (!array) os << '"' << escape(m_children[i].first) << "\": ";
            m_children[i].second.write(os, indent + 4);
            os << (i + 1 < m_children.size() ? ",\n" : "\n");
        }
        os << pad << (array ? "]" : "}");
    }

private:
    template <class T>
    static std::string number(T v) {
        std::ostringstream s;
        s.precision(17);
        s << v;
        return s.str();
    }
    static std::string escape(const std::string& s) {
        std::string o;
        for (char c : s) {
            if (c == '"' || c == '\\') o += '\\';
            o += c;
        }
        return o;
    }
    std::string m_value;
    std::vector<std::pair<std::string, json_tree>> m_children;
};

// JSONLogWriter / OutputLogWriter (examples/QTLSim/log_writer.cpp:16-31)
class log_writer {
public:
    log_writer(std::string prefix, std::string suffix)
        : m_prefix(std::move(prefix)), m_suffix(std::move(suffix)) {}
    void write(const json_tree& t) {
        if (m_prefix.empty()) {
            t.write(std::cout);
            std::cout << "\n";
            return;
        }
        std::ostringstream name;
        name << m_prefix << "." << m_index++ << m_suffix << ".json";
        std::ofstream f(name.str());
        t.write(f);
        f << "\n";
    }

private:
    std::string m_prefix, m_suffix;
    unsigned m_index = 0;
};

}  // namespace clotho_amd
