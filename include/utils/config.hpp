// config.hpp -- "key = value" input files and command-line overrides.
//
// Interface of src/utils/config.hpp:19-70 (Config::load, parse_line, has, str, num, integer,
// list, all), with its error texts ("cannot open input file: ...", "missing required
// input: ..."). Semantics kept: text after '#' is a comment; a line without '=' is ignored;
// key and value are trimmed of blanks; a later assignment of a key replaces an earlier one;
// numbers are read from the start of the value (std::stod / std::stoi); a list is the run of
// numbers at the start of the value, separated by blanks or commas.
#pragma once
#include <cctype>
#include <cstdlib>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace bem {

class Config {
public:
    Config() = default;

    static Config load(const std::string& path) {
        std::ifstream file(path);
        if (!file) throw std::runtime_error("cannot open input file: " + path);
        Config cfg;
        for (std::string text; std::getline(file, text);) cfg.parse_line(text);
        return cfg;
    }

    // one "key = value  # comment" assignment (an input-file line or a command-line override)
    void parse_line(std::string text) {
        text = text.substr(0, text.find('#'));
        const size_t eq = text.find('=');
        if (eq == std::string::npos) return;
        std::string key = blank_trimmed(text.substr(0, eq));
        if (key.empty()) return;
        values_[key] = blank_trimmed(text.substr(eq + 1));
    }

    bool has(const std::string& key) const { return values_.find(key) != values_.end(); }

    std::string str(const std::string& key, const std::string& fallback = "") const {
        const auto it = values_.find(key);
        return it == values_.end() ? fallback : it->second;
    }
    double num(const std::string& key, double fallback) const {
        const auto it = values_.find(key);
        return it == values_.end() ? fallback : std::stod(it->second);
    }
    double num(const std::string& key) const {
        const auto it = values_.find(key);
        if (it == values_.end()) throw std::runtime_error("missing required input: " + key);
        return std::stod(it->second);
    }
    int integer(const std::string& key, int fallback) const {
        const auto it = values_.find(key);
        return it == values_.end() ? fallback : std::stoi(it->second);
    }
    std::vector<double> list(const std::string& key, std::vector<double> fallback = {}) const {
        const auto it = values_.find(key);
        if (it == values_.end()) return fallback;
        std::vector<double> out;
        const char* p = it->second.c_str();
        for (;;) {
            while (*p == ',' || std::isspace((unsigned char)*p)) ++p;
            if (*p == '\0') break;
            char* end = nullptr;
            const double v = std::strtod(p, &end);
            if (end == p) break;  // the numbers end at the first token that is not one
            out.push_back(v);
            p = end;
        }
        return out;
    }
    const std::map<std::string, std::string>& all() const { return values_; }

private:
    std::map<std::string, std::string> values_;

    static bool blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\n'; }
    static std::string blank_trimmed(const std::string& s) {
        size_t lo = 0, hi = s.size();
        while (lo < hi && blank(s[lo])) ++lo;
        while (hi > lo && blank(s[hi - 1])) --hi;
        return s.substr(lo, hi - lo);
    }
};

}  // namespace bem
