// Shared helpers of the reference-app runners (wob_general.cpp, wob_vpl.cpp): the flat JSON
// configuration reader, the reference grid and the .scalar / .vector field files.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cctype>
#include <fstream>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace wobapp {

// ---------------------------------------------------------------------------------------
// configuration: a flat JSON object of strings, numbers and booleans
// ---------------------------------------------------------------------------------------
struct Config {
    std::map<std::string, std::string> s;
    std::map<std::string, double> n;
    bool has(const std::string &k) const { return s.count(k) || n.count(k); }
    std::string str(const std::string &k) const {
        auto it = s.find(k);
        if (it == s.end()) throw std::runtime_error("missing string key: " + k);
        return it->second;
    }
    double num(const std::string &k) const {
        auto it = n.find(k);
        if (it == n.end()) throw std::runtime_error("missing number key: " + k);
        return it->second;
    }
    double num_or(const std::string &k, double d) const { return n.count(k) ? n.at(k) : d; }
};

inline Config parse_config(const std::string &text) {
    Config c;
    size_t i = 0;
    auto ws = [&]() { while (i < text.size() && std::isspace((unsigned char)text[i])) ++i; };
    auto expect = [&](char ch) {
        ws();
        if (i >= text.size() || text[i] != ch) throw std::runtime_error(std::string("config: expected '") + ch + "'");
        ++i;
    };
    auto string_lit = [&]() {
        expect('"');
        std::string out;
        while (i < text.size() && text[i] != '"') {
            if (text[i] == '\\' && i + 1 < text.size()) ++i;
            out += text[i++];
        }
        expect('"');
        return out;
    };
    expect('{');
    ws();
    if (i < text.size() && text[i] == '}') return c;
    while (true) {
        const std::string key = string_lit();
        expect(':');
        ws();
        if (i < text.size() && text[i] == '"') {
            c.s[key] = string_lit();
        } else {
            size_t j = i;
            while (j < text.size() && text[j] != ',' && text[j] != '}' && !std::isspace((unsigned char)text[j])) ++j;
            const std::string tok = text.substr(i, j - i);
            i = j;
            if (tok == "true" || tok == "false") c.n[key] = tok == "true" ? 1.0 : 0.0;
            else if (tok == "null") {}
            else c.n[key] = std::stod(tok);
        }
        ws();
        if (i < text.size() && text[i] == ',') { ++i; continue; }
        expect('}');
        break;
    }
    return c;
}

// utils.cuh:92-99 with domain size 1: cell centres of (-1/2, 1/2)^2, x fastest
inline void grid_point(unsigned idx, unsigned res, double *p2) {
    const double dx = 1.0 / res;
    p2[0] = ((double)(idx % res) + 0.5) * dx - 0.5;
    p2[1] = ((double)(idx / res) + 0.5) * dx - 0.5;
}

template <typename T> void save_field(const std::string &path, unsigned res, const std::vector<T> &v) {
    std::ofstream out(path, std::ios::binary);
    out.write((const char *)&res, sizeof(unsigned));
    out.write((const char *)v.data(), (std::streamsize)(v.size() * sizeof(T)));
}

#define HIP_OK(x)                                                                                \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

}  // namespace wobapp
