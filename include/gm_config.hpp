// gm_config.hpp - the reference's configuration file (config.toml) for the
// MI355X solver: the Config keys, defaults and validation of
// include/Config.hpp:94-156 of rdellamonica/boson-cloud-binary, read by a
// small TOML reader of this header (the reference links toml++ for this).
//
// The reader takes the subset of TOML a config.toml uses: [table] headers,
// bare keys, basic "strings", booleans, integers and floats (with '_'
// separators, exponents, inf and nan), and '#' comments. Anything else is a
// parse error (std::runtime_error). As with toml++'s value_or, a key of the
// wrong type falls back to its default; an integer is taken for a real key,
// and a whole-number float for an integer key.
#pragma once

#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

namespace gm {

struct TomlValue {
    enum Kind { NONE, BOOL, INT, FLOAT, STRING } kind = NONE;
    bool b = false;
    int64_t i = 0;
    double f = 0.0;
    std::string s;
};

class TomlTables {
public:
    static TomlTables parse(const std::string &text)
    {
        TomlTables t;
        std::istringstream in(text);
        std::string line, table;
        int lineno = 0;
        while (std::getline(in, line)) {
            ++lineno;
            const std::string ln = trim(strip_comment(line, lineno));
            if (ln.empty()) continue;
            if (ln.front() == '[') {
                if (ln.size() < 3 || ln.back() != ']' || ln[1] == '[') fail(lineno, "malformed table header");
                table = trim(ln.substr(1, ln.size() - 2));
                if (!bare_key(table)) fail(lineno, "malformed table name");
                if (t.tables_.count(table) && t.defined_.count(table)) fail(lineno, "table defined twice");
                t.tables_[table];
                t.defined_[table] = true;
                continue;
            }
            const size_t eq = ln.find('=');
            if (eq == std::string::npos) fail(lineno, "expected key = value");
            const std::string key = trim(ln.substr(0, eq));
            if (!bare_key(key)) fail(lineno, "malformed key");
            auto &tab = t.tables_[table];
            if (tab.count(key)) fail(lineno, "key defined twice");
            tab[key] = value(trim(ln.substr(eq + 1)), lineno);
        }
        return t;
    }

    static TomlTables parse_file(const std::string &path)
    {
        std::ifstream f(path);
        if (!f) throw std::runtime_error("Cannot open configuration file " + path);
        std::ostringstream ss;
        ss << f.rdbuf();
        return parse(ss.str());
    }

    bool has_table(const std::string &table) const { return defined_.count(table) != 0; }

    const TomlValue *get(const std::string &table, const std::string &key) const
    {
        auto t = tables_.find(table);
        if (t == tables_.end()) return nullptr;
        auto v = t->second.find(key);
        return v == t->second.end() ? nullptr : &v->second;
    }

    double real_or(const std::string &table, const std::string &key, double def) const
    {
        const TomlValue *v = get(table, key);
        if (!v) return def;
        if (v->kind == TomlValue::FLOAT) return v->f;
        if (v->kind == TomlValue::INT) return static_cast<double>(v->i);
        return def;
    }
    int int_or(const std::string &table, const std::string &key, int def) const
    {
        const TomlValue *v = get(table, key);
        if (!v) return def;
        if (v->kind == TomlValue::INT) return static_cast<int>(v->i);
        if (v->kind == TomlValue::FLOAT && std::isfinite(v->f) && std::floor(v->f) == v->f) return static_cast<int>(v->f);
        return def;
    }
    bool bool_or(const std::string &table, const std::string &key, bool def) const
    {
        const TomlValue *v = get(table, key);
        return v && v->kind == TomlValue::BOOL ? v->b : def;
    }
    std::string string_or(const std::string &table, const std::string &key, const std::string &def) const
    {
        const TomlValue *v = get(table, key);
        return v && v->kind == TomlValue::STRING ? v->s : def;
    }

private:
    [[noreturn]] static void fail(int lineno, const std::string &what)
    {
        throw std::runtime_error("config.toml line " + std::to_string(lineno) + ": " + what);
    }
    static std::string trim(const std::string &s)
    {
        const size_t a = s.find_first_not_of(" \t\r");
        if (a == std::string::npos) return "";
        const size_t b = s.find_last_not_of(" \t\r");
        return s.substr(a, b - a + 1);
    }
    static bool bare_key(const std::string &k)
    {
        if (k.empty()) return false;
        for (char c : k)
            if (!(std::isalnum(static_cast<unsigned char>(c)) || c == '_' || c == '-')) return false;
        return true;
    }
    // drop a '#' comment that is not inside a basic string
    static std::string strip_comment(const std::string &s, int lineno)
    {
        bool in_str = false;
        for (size_t k = 0; k < s.size(); ++k) {
            const char c = s[k];
            if (in_str && c == '\\') {
                ++k;
                continue;
            }
            if (c == '"') in_str = !in_str;
            if (c == '#' && !in_str) return s.substr(0, k);
        }
        if (in_str) fail(lineno, "unterminated string");
        return s;
    }
    static TomlValue value(const std::string &v, int lineno)
    {
        TomlValue out;
        if (v.empty()) fail(lineno, "missing value");
        if (v.front() == '"') {
            if (v.size() < 2 || v.back() != '"') fail(lineno, "malformed string");
            out.kind = TomlValue::STRING;
            for (size_t k = 1; k + 1 < v.size(); ++k) {
                char c = v[k];
                if (c == '\\') {
                    if (k + 2 >= v.size()) fail(lineno, "malformed escape");
                    const char e = v[++k];
                    c = e == 'n' ? '\n' : e == 't' ? '\t' : e == '"' ? '"' : e == '\\' ? '\\' : 0;
                    if (!c) fail(lineno, "unsupported escape");
                } else if (c == '"') {
                    fail(lineno, "malformed string");
                }
                out.s.push_back(c);
            }
            return out;
        }
        if (v == "true" || v == "false") {
            out.kind = TomlValue::BOOL;
            out.b = v == "true";
            return out;
        }
        std::string num;
        for (size_t k = 0; k < v.size(); ++k) {
            if (v[k] == '_') {
                if (k == 0 || k + 1 == v.size() || !std::isdigit(static_cast<unsigned char>(v[k - 1])) ||
                    !std::isdigit(static_cast<unsigned char>(v[k + 1])))
                    fail(lineno, "malformed number");
                continue;
            }
            num.push_back(v[k]);
        }
        const std::string body = (num[0] == '+' || num[0] == '-') ? num.substr(1) : num;
        if (body == "inf" || body == "nan") {
            out.kind = TomlValue::FLOAT;
            out.f = body == "inf" ? INFINITY : NAN;
            if (num[0] == '-') out.f = -out.f;
            return out;
        }
        const bool is_float = num.find_first_of(".eE") != std::string::npos;
        for (char c : body)
            if (!(std::isdigit(static_cast<unsigned char>(c)) || c == '.' || c == 'e' || c == 'E' || c == '+' ||
                  c == '-'))
                fail(lineno, "unsupported value '" + v + "'");
        if (body.empty() || !std::isdigit(static_cast<unsigned char>(body[0]))) fail(lineno, "malformed number");
        char *end = nullptr;
        errno = 0;
        if (is_float) {
            out.kind = TomlValue::FLOAT;
            out.f = std::strtod(num.c_str(), &end);
        } else {
            out.kind = TomlValue::INT;
            out.i = std::strtoll(num.c_str(), &end, 10);
        }
        if (errno != 0 || end != num.c_str() + num.size()) fail(lineno, "malformed number '" + v + "'");
        return out;
    }

    std::map<std::string, std::map<std::string, TomlValue>> tables_;
    std::map<std::string, bool> defined_;
};

// The reference's Config (include/Config.hpp:17-88), same keys and defaults.
struct Config {
    int N = 2048;
    double L = 5000.0;
    double dt = 10.0;
    double num_periods = 0.0;
    int save_every = 50;
    double ramp_periods = 0.0;
    bool ramp_secondary_potential = true;
    double steps_per_period = 50.0;
    double power_average_orbits = 2.0;
    bool evolve_orbit = false;
    bool absorption = false;
    int l_max = 5;
    int num_steps = 0;
    double epsilon = 0.001;
    double epsilon_smear = 2.0;
    double r_ex = 1000.0;
    double q = 0.001;
    double R = 50.0;
    bool counter = false;
    double mu = 0.2;
    double a = 0.0;
    double M_c = 0.0;
    int n = 2, l = 1, m = 1;
    int threads = 0;
    std::string preconditioner = "radial";
    bool mask = false;
    double mask_L_factor = 1.0, mask_W = 0.1;
    int mask_n = 3;
    bool cap = false;
    double cap_L_factor = 1.0, cap_W = 0.1;
    int cap_n = 3;
};

// load_config (include/Config.hpp:94-156): the same keys, defaults and
// std::invalid_argument checks; a missing or malformed file throws
// std::runtime_error.
inline Config load_config_text(const std::string &text)
{
    const TomlTables t = TomlTables::parse(text);
    Config c;
    c.threads = t.int_or("parallel", "threads", 0);
    c.preconditioner = t.string_or("solver", "preconditioner", "radial");
    if (c.threads < 0) throw std::invalid_argument("parallel.threads must be nonnegative");
    if (c.preconditioner != "radial" && c.preconditioner != "ilut")
        throw std::invalid_argument("solver.preconditioner must be radial or ilut");
    c.N = t.int_or("grid", "N", 2048);
    c.L = t.real_or("grid", "L", 5000.0);
    c.dt = t.real_or("simulation", "dt", 10.0);
    c.num_periods = t.real_or("simulation", "num_periods", 0.0);
    c.num_steps = t.int_or("simulation", "num_steps", 0);
    c.save_every = t.int_or("simulation", "save_every", 50);
    c.l_max = t.int_or("simulation", "l_max", 5);
    c.r_ex = t.real_or("simulation", "r_ex", 1000.0);
    c.ramp_periods = t.real_or("simulation", "ramp_periods", 0.0);
    c.ramp_secondary_potential = t.bool_or("simulation", "ramp_secondary_potential", true);
    c.steps_per_period = t.real_or("simulation", "steps_per_period", 50.0);
    if (!(c.steps_per_period > 0.0)) throw std::invalid_argument("simulation.steps_per_period must be positive");
    c.power_average_orbits = t.real_or("simulation", "power_average_orbits", 2.0);
    if (!(c.power_average_orbits > 0.0)) throw std::invalid_argument("simulation.power_average_orbits must be positive");
    c.epsilon = t.real_or("simulation", "epsilon", 0.001);
    c.epsilon_smear = t.real_or("simulation", "epsilon_smear", 2.0);
    c.evolve_orbit = t.bool_or("simulation", "evolve_orbit", false);
    c.absorption = t.bool_or("simulation", "absorption", false);
    c.q = t.real_or("binary", "q", 0.001);
    c.R = t.real_or("binary", "R", 50.0);
    c.counter = t.bool_or("binary", "counter", false);
    c.mu = t.real_or("atom", "mu", 0.2);
    c.n = t.int_or("atom", "n", 2);
    c.l = t.int_or("atom", "l", 1);
    c.m = t.int_or("atom", "m", 1);
    c.a = t.real_or("atom", "a", 0.0);
    c.M_c = t.real_or("atom", "M_c", 0.0);
    if (t.has_table("mask")) {
        c.mask = true;
        c.mask_L_factor = t.real_or("mask", "L_factor", 1.0);
        c.mask_W = t.real_or("mask", "W", 0.1);
        c.mask_n = t.int_or("mask", "n", 3);
    }
    if (t.has_table("cap")) {
        c.cap = true;
        c.cap_L_factor = t.real_or("cap", "L_factor", 1.0);
        c.cap_W = t.real_or("cap", "W", 0.1);
        c.cap_n = t.int_or("cap", "n", 3);
    }
    return c;
}

inline Config load_config(const std::string &path)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Cannot open configuration file " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return load_config_text(ss.str());
}

}  // namespace gm
