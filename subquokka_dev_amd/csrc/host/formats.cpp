// formats.cpp -- circuit text grammar and seeded synthetic inputs.
//
// Grammar: the MEM branch of Simulator::setGate_MEM (simulator.cpp:542-597)
// with the argument readers setGate_1/_2/_vswap/_Phase/setUnitary
// (simulator.cpp:259-356) and check_targ_in_N (simulator.cpp:764-774).
// Subcircuit files: setupSubCircuits_MEM (simulator.cpp:619-636).
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "gate_lowering.hpp"
#include "rng.hpp"

namespace {

struct NameInfo { const char *name; int kind; int nq; int has_param; int n_reals; };

const NameInfo kNames[] = {
    {"H", SQ_GATE_H, 1, 0, 0},   {"X", SQ_GATE_X, 1, 0, 0},   {"Y", SQ_GATE_Y, 1, 0, 0},
    {"Z", SQ_GATE_Z, 1, 0, 0},   {"P", SQ_GATE_P, 1, 1, 0},   {"RX", SQ_GATE_RX, 1, 1, 0},
    {"RY", SQ_GATE_RY, 1, 1, 0}, {"RZ", SQ_GATE_RZ, 1, 1, 0}, {"U1", SQ_GATE_U1, 1, 0, 8},
    {"U2", SQ_GATE_U2, 2, 0, 32}, {"SWAP", SQ_GATE_SWAP, 2, 0, 0}, {"CP", SQ_GATE_CP, 2, 1, 0},
    {"RZZ", SQ_GATE_RZZ, 2, 1, 0}, {"VSWAP_1_1", SQ_GATE_VSWAP_1_1, 2, 0, 0},
    {"VSWAP_2_2", SQ_GATE_VSWAP_2_2, 4, 0, 0},
};

// Names the reference parses but whose MEM implementation is broken or
// disabled (circuit.cpp:4179-4181, simulator.cpp:577-578), or that need MPI ranks.
const char *kUnsupported[] = {"VSWAP_3_3", "VSWAP_4_4", "VSWAP_6_6", "U3",
                              "MPI_VSWAP_1_1", "MPI_VSWAP_2_2"};

bool read_int(const char *&p, int &v) {
    char *end;
    long x = std::strtol(p, &end, 10);
    if (end == p) return false;
    v = (int)x; p = end;
    return true;
}
bool read_double(const char *&p, double &v) {
    char *end;
    double x = std::strtod(p, &end);
    if (end == p) return false;
    v = x; p = end;
    return true;
}
bool is_blank(const std::string &s) { return s.find_first_not_of(" \t\r") == std::string::npos; }

using sq::SplitMix64;

void put(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void put(std::string &s, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    int n = std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    s.append(buf, (size_t)n);
}

int emit(const std::string &s, char *buf, uint64_t cap, uint64_t *len) {
    if (len) *len = s.size() + 1;
    if (!buf || cap < s.size() + 1) return SQ_ERR_ARG;
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return SQ_OK;
}

// CNOT c -> t as "U2 min max" with the 4x4 matrix in basis |first second>,
// first = most significant (the layout circuit.cpp:2811-2819 reverses).
void put_cnot(std::string &s, int c, int t) {
    static const int ct[16] = {1,0,0,0, 0,1,0,0, 0,0,0,1, 0,0,1,0};  // c < t: control is first
    static const int tc[16] = {1,0,0,0, 0,0,0,1, 0,0,1,0, 0,1,0,0};  // c > t: control is second
    const int *m = c < t ? ct : tc;
    put(s, "U2 %d %d", c < t ? c : t, c < t ? t : c);
    for (int i = 0; i < 16; i++) put(s, " %d 0", m[i]);
    s += '\n';
}

}  // namespace

extern "C" int sq_parse_gate(const char *line, int32_t total_qbit, sq_gate *out) {
    if (!line || !out) return SQ_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    char name[32];
    int consumed = 0;
    if (std::sscanf(line, " %31s%n", name, &consumed) != 1) return SQ_ERR_PARSE;
    const char *p = line + consumed;
    const NameInfo *info = nullptr;
    for (auto &ni : kNames) if (std::strcmp(ni.name, name) == 0) { info = &ni; break; }
    if (!info) {
        for (auto u : kUnsupported) if (std::strcmp(u, name) == 0) return SQ_ERR_UNSUPPORTED;
        return SQ_ERR_UNKNOWN_GATE;
    }
    out->kind = info->kind;
    out->nq = info->nq;
    for (int i = 0; i < info->nq; i++) if (!read_int(p, out->q[i])) return SQ_ERR_PARSE;
    for (int i = 0; i < info->nq; i++)
        if (out->q[i] < 0 || out->q[i] >= total_qbit) return SQ_ERR_QUBIT_INDEX;
    if (info->has_param && !read_double(p, out->param)) return SQ_ERR_PARSE;
    for (int i = 0; i < info->n_reals; i++) if (!read_double(p, out->mat[i])) return SQ_ERR_PARSE;
    sq::LoweredGate lg;
    return sq::lower_gate(*out, total_qbit, lg, nullptr);   // arity / repeat / U2-order checks
}

extern "C" int sq_parse_circuit(const char *text, int32_t total_qbit, int32_t is_subcircuit,
                                sq_gate *gates, uint64_t gate_cap, uint64_t *n_gates,
                                uint32_t *block_sizes, uint64_t block_cap, uint64_t *n_blocks,
                                int32_t *err_line) {
    if (!text) return SQ_ERR_ARG;
    std::vector<std::string> lines;
    {
        std::stringstream ss(text);
        std::string line;
        while (std::getline(ss, line)) lines.push_back(line);
    }
    uint64_t ng = 0, nb = 0;
    bool overflow = false;
    auto add_gate = [&](const std::string &l, int ln) -> int {
        sq_gate g;
        int rc = sq_parse_gate(l.c_str(), total_qbit, &g);
        if (rc) { if (err_line) *err_line = ln; return rc; }
        if (gates && ng < gate_cap) gates[ng] = g; else overflow = true;
        ng++;
        return SQ_OK;
    };
    for (size_t i = 0; i < lines.size(); i++) {
        if (is_blank(lines[i])) continue;
        if (!is_subcircuit) {
            int rc = add_gate(lines[i], (int)i + 1);
            if (rc) return rc;
            continue;
        }
        const char *p = lines[i].c_str();
        int cnt;
        if (!read_int(p, cnt) || cnt < 0) { if (err_line) *err_line = (int)i + 1; return SQ_ERR_PARSE; }
        for (int k = 0; k < cnt; k++) {
            if (++i >= lines.size()) { if (err_line) *err_line = (int)i; return SQ_ERR_PARSE; }
            int rc = add_gate(lines[i], (int)i + 1);
            if (rc) return rc;
        }
        if (block_sizes && nb < block_cap) block_sizes[nb] = (uint32_t)cnt; else if (block_sizes || block_cap) overflow = true;
        nb++;
    }
    if (n_gates) *n_gates = ng;
    if (n_blocks) *n_blocks = nb;
    return overflow ? SQ_ERR_ARG : SQ_OK;
}

extern "C" int sq_gen_state(int32_t n_qubits, uint64_t seed, double *out) {
    if (n_qubits < 0 || n_qubits > 34 || !out) return SQ_ERR_ARG;
    const uint64_t n = 1ULL << n_qubits;
    SplitMix64 rng{seed};
    const double two_pi = 6.283185307179586;
    for (uint64_t i = 0; i < n; i++) {
        // Box-Muller: both outputs used as (re, im); 1-u keeps the log argument in (0, 1].
        double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        out[2 * i] = r * std::cos(two_pi * u2);
        out[2 * i + 1] = r * std::sin(two_pi * u2);
    }
    // Pairwise (blocked) FP64 sum of |a|^2.
    std::vector<double> part;
    const uint64_t blk = 4096;
    for (uint64_t b = 0; b < n; b += blk) {
        double s = 0;
        for (uint64_t i = b; i < b + blk && i < n; i++) s += out[2 * i] * out[2 * i] + out[2 * i + 1] * out[2 * i + 1];
        part.push_back(s);
    }
    while (part.size() > 1) {
        std::vector<double> nx;
        for (size_t i = 0; i < part.size(); i += 2) nx.push_back(part[i] + (i + 1 < part.size() ? part[i + 1] : 0.0));
        part.swap(nx);
    }
    double inv = 1.0 / std::sqrt(part[0]);
    for (uint64_t i = 0; i < 2 * n; i++) out[i] *= inv;
    return SQ_OK;
}

extern "C" int sq_gen_circuit(int32_t kind, int32_t n, int32_t layers, uint64_t seed,
                              const double *theta, char *buf, uint64_t cap, uint64_t *len) {
    if (n < 1 || n > 64 || layers < 0) return SQ_ERR_ARG;
    std::string s;
    const double pi = 3.141592653589793;
    SplitMix64 rng{seed};
    if (kind == 0) {
        // QFT in python_script/cir_qft.py:27-35 order.
        for (int i = 0; i < n; i++) {
            put(s, "H %d\n", i);
            double angle = pi / 2;
            for (int j = i + 1; j < n; j++) {
                put(s, "CP %d %d %.17g\n", i, j, angle);
                angle /= 2;
            }
        }
        for (int i = 0; i < n / 2; i++) put(s, "SWAP %d %d\n", i, n - 1 - i);
    } else if (kind == 1) {
        // Random layers: U3(theta, phi, lambda) on every qubit as U1, then CZ on a
        // seeded perfect matching as CP min max pi.
        std::vector<int> perm(n);
        for (int l = 0; l < layers; l++) {
            for (int q = 0; q < n; q++) {
                double th = 2 * pi * rng.uniform(), ph = 2 * pi * rng.uniform(), la = 2 * pi * rng.uniform();
                double c = std::cos(th / 2), sn = std::sin(th / 2);
                double m[8] = {c, 0.0,
                               -std::cos(la) * sn, -std::sin(la) * sn,
                               std::cos(ph) * sn, std::sin(ph) * sn,
                               std::cos(ph + la) * c, std::sin(ph + la) * c};
                put(s, "U1 %d", q);
                for (double v : m) put(s, " %.17g", v);
                s += '\n';
            }
            for (int i = 0; i < n; i++) perm[i] = i;
            for (int i = n - 1; i > 0; i--) {   // Fisher-Yates
                int j = (int)(rng.uniform() * (double)(i + 1));
                std::swap(perm[i], perm[j]);
            }
            for (int m = 0; m + 1 < n; m += 2) {
                int a = std::min(perm[m], perm[m + 1]), b = std::max(perm[m], perm[m + 1]);
                put(s, "CP %d %d %.17g\n", a, b, pi);
            }
        }
    } else if (kind == 2) {
        // Hardware-efficient ansatz: RY on all, RZ on all, CNOT ring i -> (i+1) mod n.
        int p = 0;
        for (int l = 0; l < layers; l++) {
            for (int q = 0; q < n; q++, p++)
                put(s, "RY %d %.17g\n", q, theta ? theta[p] : -pi + 2 * pi * rng.uniform());
            for (int q = 0; q < n; q++, p++)
                put(s, "RZ %d %.17g\n", q, theta ? theta[p] : -pi + 2 * pi * rng.uniform());
            if (n > 1)
                for (int q = 0; q < n; q++) put_cnot(s, q, (q + 1) % n);
        }
    } else {
        return SQ_ERR_ARG;
    }
    return emit(s, buf, cap, len);
}
