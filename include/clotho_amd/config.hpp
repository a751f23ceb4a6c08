// clotho_amd/config.hpp — the reference's run configuration: a JSON document
// whose "configuration" block holds the parameter groups
// (src/clotho/utility/clotho_strings.cpp:16-47), read by the driver with
// boost::property_tree::read_json (QS/qtlsim_common.cpp:17-23) and queried by
// dotted key with a default that is written back
// (include/clotho/mutation/mutation_rate_parameter.hpp:30-40).
//
// property_tree keeps the document flat: nested objects become dotted keys
// ("mutation.mu"), element i of an array becomes key "<array>.<i>", and every
// leaf is kept as its text (Boost's write_json quotes numbers, so "100" and
// 100 read alike).  The "configuration." prefix is optional on every key.
#ifndef CLOTHO_AMD_CONFIG_HPP_
#define CLOTHO_AMD_CONFIG_HPP_

#include <cctype>
#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace clotho_amd {

class property_tree {
public:
    void put(const std::string& key, const std::string& value) { m_kv[strip(key)] = value; }
    void put(const std::string& key, double value) {
        std::ostringstream os;
        os.precision(17);
        os << value;
        m_kv[strip(key)] = os.str();
    }
    bool has(const std::string& key) const { return m_kv.count(strip(key)) != 0; }
    // true when the key or any key below it is present (get_child_optional)
    bool has_child(const std::string& key) const {
        const std::string k = strip(key);
        if (m_kv.count(k)) return true;
        auto it = m_kv.lower_bound(k + ".");
        return it != m_kv.end() && it->first.compare(0, k.size() + 1, k + ".") == 0;
    }
    std::string get(const std::string& key, const std::string& dflt) const {
        auto it = m_kv.find(strip(key));
        return it == m_kv.end() ? dflt : it->second;
    }
    double get(const std::string& key, double dflt) const {
        auto it = m_kv.find(strip(key));
        if (it == m_kv.end()) return dflt;
        if (it->second == "true") return 1.0;
        if (it->second == "false") return 0.0;
        return std::stod(it->second);
    }
    // elements "<key>.0", "<key>.1", ... up to the first missing index
    size_t array_size(const std::string& key) const {
        size_t n = 0;
        while (has_child(key + "." + std::to_string(n))) ++n;
        return n;
    }
    std::vector<double> get_array(const std::string& key) const {
        std::vector<double> v(array_size(key));
        for (size_t i = 0; i < v.size(); ++i) v[i] = get(key + "." + std::to_string(i), 0.0);
        return v;
    }
    void put_array(const std::string& key, const std::vector<double>& v) {
        erase(key);
        for (size_t i = 0; i < v.size(); ++i) put(key + "." + std::to_string(i), v[i]);
    }
    // remove the key and everything below it
    void erase(const std::string& key) {
        const std::string k = strip(key);
        m_kv.erase(k);
        auto it = m_kv.lower_bound(k + ".");
        while (it != m_kv.end() && it->first.compare(0, k.size() + 1, k + ".") == 0)
            it = m_kv.erase(it);
    }
    const std::map<std::string, std::string>& items() const { return m_kv; }

private:
    static std::string strip(const std::string& k) {
        const std::string p = "configuration.";
        return k.compare(0, p.size(), p) == 0 ? k.substr(p.size()) : k;
    }
    std::map<std::string, std::string> m_kv;
};

// ---------------------------------------------------------------------------
// JSON reader (RFC 8259 values; \u escapes outside ASCII become '?')
// ---------------------------------------------------------------------------
namespace json_detail {

class reader {
public:
    reader(const std::string& text, property_tree& out) : s(text), t(out) {}
    void parse(const std::string& root) {
        value(root);
        ws();
        if (i != s.size()) fail("trailing characters");
    }

private:
    const std::string& s;
    property_tree& t;
    size_t i = 0;

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("JSON: " + what + " at offset " + std::to_string(i));
    }
    void ws() {
        while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    }
    bool eat(char c) {
        ws();
        if (i < s.size() && s[i] == c) {
            ++i;
            return true;
        }
        return false;
    }
    static std::string join(const std::string& a, const std::string& b) {
        return a.empty() ? b : a + "." + b;
    }
    std::string string_lit() {
        if (!eat('"')) fail("expected a string");
        std::string out;
        while (i < s.size() && s[i] != '"') {
            char c = s[i++];
            if (c == '\\') {
                if (i >= s.size()) fail("bad escape");
                const char e = s[i++];
                switch (e) {
                    case '"': out += '"'; break;
                    case '\\': out += '\\'; break;
                    case '/': out += '/'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'n': out += '\n'; break;
                    case 'r': out += '\r'; break;
                    case 't': out += '\t'; break;
                    case 'u': {
                        if (i + 4 > s.size()) fail("bad \\u escape");
                        const unsigned v = (unsigned)std::stoul(s.substr(i, 4), nullptr, 16);
                        i += 4;
                        out += v < 0x80 ? (char)v : '?';
                        break;
                    }
                    default: fail("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i >= s.size()) fail("unterminated string");
        ++i;
        return out;
    }
    void value(const std::string& key) {
        ws();
        if (i >= s.size()) fail("unexpected end");
        const char c = s[i];
        if (c == '{') {
            ++i;
            if (eat('}')) return;
            do {
                const std::string k = string_lit();
                if (!eat(':')) fail("expected ':'");
                value(join(key, k));
            } while (eat(','));
            if (!eat('}')) fail("expected '}'");
        } else if (c == '[') {
            ++i;
            if (eat(']')) return;
            size_t n = 0;
            do {
                value(join(key, std::to_string(n++)));
            } while (eat(','));
            if (!eat(']')) fail("expected ']'");
        } else if (c == '"') {
            t.put(key, string_lit());
        } else {
            const size_t b = i;
            while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '-' ||
                                    s[i] == '+' || s[i] == '.'))
                ++i;
            const std::string lit = s.substr(b, i - b);
            if (lit.empty()) fail("unexpected character");
            if (lit == "null") return;  // absent
            if (lit != "true" && lit != "false") {
                size_t used = 0;
                try {
                    (void)std::stod(lit, &used);
                } catch (const std::exception&) {
                    used = 0;
                }
                if (used != lit.size()) fail("bad literal '" + lit + "'");
            }
            t.put(key, lit);
        }
    }
};

}  // namespace json_detail

// Merge a JSON document into tree (keys below `root`, "" = top level).
inline void read_json_text(const std::string& text, property_tree& tree,
                           const std::string& root = "") {
    json_detail::reader(text, tree).parse(root);
}

inline void read_json(const std::string& path, property_tree& tree) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot read configuration file " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    read_json_text(ss.str(), tree);
}

}  // namespace clotho_amd

#endif  // CLOTHO_AMD_CONFIG_HPP_
