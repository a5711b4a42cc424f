// Shared helpers of the reference-app runners (wob_general.cpp, wob_vpl.cpp): the flat JSON
// configuration reader, the reference grid and the .scalar / .vector field files.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cctype>
#include <cmath>
#include <fstream>
#include <sstream>
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

// ---------------------------------------------------------------------------------------
// OBJ: "v x y z" and "f i j [k]" records (1-based, "i/t/n" accepted, negative = relative)
// ---------------------------------------------------------------------------------------
inline void read_obj(const std::string &path, std::vector<double> &V, std::vector<int> &F, int &verts_per_face) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open mesh " + path);
    std::string line;
    verts_per_face = 0;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        ss >> tag;
        if (tag == "v") {
            double x = 0, y = 0, z = 0;
            ss >> x >> y >> z;
            V.insert(V.end(), {x, y, z});
        } else if (tag == "f" || tag == "l") {
            std::vector<int> f;
            std::string tok;
            while (ss >> tok) {
                int k = std::stoi(tok.substr(0, tok.find('/')));
                k = k < 0 ? (int)(V.size() / 3) + k : k - 1;
                f.push_back(k);
            }
            if (verts_per_face == 0) verts_per_face = (int)f.size();
            if ((int)f.size() != verts_per_face) throw std::runtime_error("mixed face sizes in " + path);
            F.insert(F.end(), f.begin(), f.end());
        }
    }
}

// winding number of a closed 2D polyline (segments F[2i], F[2i+1] of V, 3 doubles per vertex)
// around x: the sum of the signed angles the segments subtend, over 2 pi
inline double winding_2d(const std::vector<double> &V, const std::vector<int> &F, const double *x) {
    double w = 0.0;
    for (size_t f = 0; f + 1 < F.size(); f += 2) {
        const double ax = V[3 * F[f]] - x[0], ay = V[3 * F[f] + 1] - x[1];
        const double bx = V[3 * F[f + 1]] - x[0], by = V[3 * F[f + 1] + 1] - x[1];
        w += std::atan2(ax * by - ay * bx, ax * bx + ay * by);
    }
    return w / (2.0 * M_PI);
}

}  // namespace wobapp
