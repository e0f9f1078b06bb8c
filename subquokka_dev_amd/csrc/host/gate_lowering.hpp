// gate_lowering.hpp -- turns one parsed gate (sq_gate) into the form the
// kernels consume: a dense 2^k x 2^k matrix, a diagonal, or a qubit-swap
// permutation, always in the internal basis where local index bit j belongs
// to the j-th smallest target.
//
// The constants are the ones the reference's MEM gate constructors compute
// (circuit.cpp:2550-3041, 3330-3347), including the truncated H literal
// 0.70710678118 (circuit.cpp:247) and the U2 coefficient bit reversal
// (circuit.cpp:2811-2819), so an engine run reproduces the reference's
// gate values, not textbook ones.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <vector>
#include <cstdint>
#include <string>

#include "subquokka_hip.h"

namespace sq {

constexpr double kHCoeff = 0.70710678118;  // circuit.cpp:247

enum class OpClass : int { Dense = 0, Diag = 1, Perm = 2 };

struct LoweredGate {
    OpClass cls;
    int k;                        // number of targets (1, 2 or 4 for Perm)
    int t[4];                     // sorted targets (Dense/Diag); Perm: pairs (t[0]<->t[1]), (t[2]<->t[3])
    int n_pairs;                  // Perm: number of swapped qubit pairs (1 or 2)
    std::complex<double> m[16];   // Dense: row-major 2^k x 2^k; Diag: m[0..2^k-1]
    uint32_t diag_active;         // Diag: bitmask of local indices whose entry != 1
    double pass_fraction;         // algorithmic-bytes factor f_g (fraction of state rewritten)
    const char *name;
};

inline int gate_arity(int kind) {
    switch (kind) {
    case SQ_GATE_H: case SQ_GATE_X: case SQ_GATE_Y: case SQ_GATE_Z: case SQ_GATE_P:
    case SQ_GATE_RX: case SQ_GATE_RY: case SQ_GATE_RZ: case SQ_GATE_U1:
        return 1;
    case SQ_GATE_U2: case SQ_GATE_SWAP: case SQ_GATE_CP: case SQ_GATE_RZZ:
    case SQ_GATE_VSWAP_1_1:
        return 2;
    case SQ_GATE_VSWAP_2_2:
        return 4;
    default:
        return -1;
    }
}

// Validates targets against n qubits and lowers.  Returns SQ_OK or an error
// with a message in *why.
inline int lower_gate(const sq_gate &g, int n_qubits, LoweredGate &out, std::string *why) {
    using C = std::complex<double>;
    const int ar = gate_arity(g.kind);
    if (ar < 0) { if (why) *why = "[setGate]: Not implemented yet."; return SQ_ERR_UNKNOWN_GATE; }
    if (g.nq != ar) { if (why) *why = "gate arity mismatch"; return SQ_ERR_ARG; }
    for (int i = 0; i < ar; i++)
        if (g.q[i] < 0 || g.q[i] >= n_qubits) { if (why) *why = "[setGate]: Qubit Index Error"; return SQ_ERR_QUBIT_INDEX; }
    for (int i = 0; i < ar; i++)
        for (int j = i + 1; j < ar; j++)
            if (g.q[i] == g.q[j]) { if (why) *why = "repeated target qubit"; return SQ_ERR_CONTRACT; }

    out = LoweredGate{};
    out.k = ar;
    for (int i = 0; i < ar; i++) out.t[i] = g.q[i];
    out.pass_fraction = 1.0;
    const double phi = g.param;
    const C one(1.0, 0.0), zero(0.0, 0.0);

    if (ar == 2 && g.kind != SQ_GATE_U2) std::sort(out.t, out.t + 2);  // symmetric gates
    if (g.kind == SQ_GATE_U2 && g.q[0] > g.q[1]) {
        if (why) *why = "U2 needs targ[0] < targ[1] (circuit.cpp:2809)";
        return SQ_ERR_CONTRACT;
    }

    switch (g.kind) {
    case SQ_GATE_H:
        out.cls = OpClass::Dense; out.name = "H";
        out.m[0] = kHCoeff; out.m[1] = kHCoeff; out.m[2] = kHCoeff; out.m[3] = -kHCoeff;
        break;
    case SQ_GATE_X:
        out.cls = OpClass::Dense; out.name = "X";
        out.m[0] = zero; out.m[1] = one; out.m[2] = one; out.m[3] = zero;
        break;
    case SQ_GATE_Y:  // YGATE: a0' = -i a1, a1' = i a0 (circuit.cpp:263-265)
        out.cls = OpClass::Dense; out.name = "Y";
        out.m[0] = zero; out.m[1] = C(0, -1); out.m[2] = C(0, 1); out.m[3] = zero;
        break;
    case SQ_GATE_Z:
        out.cls = OpClass::Diag; out.name = "Z";
        out.m[0] = one; out.m[1] = C(-1.0, 0.0); out.diag_active = 0x2; out.pass_fraction = 0.5;
        break;
    case SQ_GATE_P:  // circuit.cpp:2655
        out.cls = OpClass::Diag; out.name = "P";
        out.m[0] = one; out.m[1] = C(std::cos(phi), std::sin(phi)); out.diag_active = 0x2;
        out.pass_fraction = 0.5;
        break;
    case SQ_GATE_RX: {  // circuit.cpp:2708-2709
        out.cls = OpClass::Dense; out.name = "RX";
        C c(std::cos(phi / 2), 0), s(0, std::sin(-phi / 2));
        out.m[0] = c; out.m[1] = s; out.m[2] = s; out.m[3] = c;
        break; }
    case SQ_GATE_RY: {  // circuit.cpp:2736-2737, RYGATE 297-300
        out.cls = OpClass::Dense; out.name = "RY";
        double c = std::cos(phi / 2), s = std::sin(phi / 2);
        out.m[0] = c; out.m[1] = -s; out.m[2] = s; out.m[3] = c;
        break; }
    case SQ_GATE_RZ:  // circuit.cpp:2764-2765, RZGATE 307-309
        out.cls = OpClass::Diag; out.name = "RZ";
        out.m[0] = C(std::cos(-phi / 2), std::sin(-phi / 2));
        out.m[1] = C(std::cos(phi / 2), std::sin(phi / 2));
        out.diag_active = 0x3;
        break;
    case SQ_GATE_U1:  // circuit.cpp:274-277
        out.cls = OpClass::Dense; out.name = "U1";
        for (int i = 0; i < 4; i++) out.m[i] = C(g.mat[2 * i], g.mat[2 * i + 1]);
        break;
    case SQ_GATE_U2: {  // circuit.cpp:2811-2819
        out.cls = OpClass::Dense; out.name = "U2";
        for (int i = 0; i < 16; i++) {
            int b1 = i & 1, b2 = (i >> 1) & 1, b3 = (i >> 2) & 1, b4 = (i >> 3) & 1;
            int res = (b3 << 3) | (b4 << 2) | (b1 << 1) | b2;
            out.m[i] = C(g.mat[2 * res], g.mat[2 * res + 1]);
        }
        // An exactly diagonal matrix (a CZ, a controlled phase written as U2)
        // is a diagonal gate: the same products, only the amplitudes it changes
        // are visited, and register passes carry it on any qubits.
        bool diagonal = true;
        uint32_t active = 0;
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++)
                if (r != c && out.m[r * 4 + c] != zero) diagonal = false;
        for (int r = 0; r < 4; r++) if (out.m[r * 5] != one) active |= 1u << r;
        if (diagonal && active) {
            const C d[4] = {out.m[0], out.m[5], out.m[10], out.m[15]};
            out.cls = OpClass::Diag;
            for (int r = 0; r < 4; r++) out.m[r] = d[r];
            for (int i = 4; i < 16; i++) out.m[i] = zero;
            out.diag_active = active;
            out.pass_fraction = __builtin_popcount(active) / 4.0;
        }
        break; }
    case SQ_GATE_SWAP: case SQ_GATE_VSWAP_1_1:  // SWAPGate circuit.cpp:351-355
        out.cls = OpClass::Perm; out.name = g.kind == SQ_GATE_SWAP ? "SWAP" : "VSWAP_1_1";
        out.n_pairs = 1; out.pass_fraction = 0.5;
        break;
    case SQ_GATE_CP:  // circuit.cpp:2949, CPGATE 362-363
        out.cls = OpClass::Diag; out.name = "CP";
        out.m[0] = one; out.m[1] = one; out.m[2] = one;
        out.m[3] = C(std::cos(phi), std::sin(phi));
        out.diag_active = 0x8; out.pass_fraction = 0.25;
        break;
    case SQ_GATE_RZZ: {  // circuit.cpp:2990-2991, RZZGATE 365-369
        out.cls = OpClass::Diag; out.name = "RZZ";
        C p(std::cos(phi / 2), std::sin(phi / 2)), n(std::cos(-phi / 2), std::sin(-phi / 2));
        out.m[0] = n; out.m[1] = p; out.m[2] = p; out.m[3] = n;
        out.diag_active = 0xF;
        break; }
    case SQ_GATE_VSWAP_2_2:  // swap t0<->t2, t1<->t3 (circuit.cpp:3457-3556)
        out.cls = OpClass::Perm; out.name = "VSWAP_2_2";
        out.n_pairs = 2;
        out.t[0] = g.q[0]; out.t[1] = g.q[2]; out.t[2] = g.q[1]; out.t[3] = g.q[3];
        out.pass_fraction = 1.0;
        break;
    default:
        if (why) *why = "[setGate]: Not implemented yet.";
        return SQ_ERR_UNKNOWN_GATE;
    }
    return SQ_OK;
}

// The gate with its targets mapped through sigma (a permutation of qubits:
// the layout after VSWAP data moves, or a pass's virtual positions).  A two-qubit gate keeps its internal basis
// (local bit j <-> j-th smallest target), so when the mapping reverses the
// order of its targets, its matrix or diagonal is permuted to match.
inline void relabel(LoweredGate &g, const std::vector<int> &sigma) {
    if (g.cls == OpClass::Perm) {
        for (int i = 0; i < 2 * g.n_pairs; i++) g.t[i] = sigma[g.t[i]];
        return;
    }
    if (g.k == 1) {
        g.t[0] = sigma[g.t[0]];
        return;
    }
    const int a = sigma[g.t[0]], b = sigma[g.t[1]];
    if (a < b) {
        g.t[0] = a;
        g.t[1] = b;
        return;
    }
    g.t[0] = b;
    g.t[1] = a;
    auto sw = [](int l) { return ((l & 1) << 1) | ((l >> 1) & 1); };
    if (g.cls == OpClass::Dense) {
        std::complex<double> m[16];
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++) m[sw(r) * 4 + sw(c)] = g.m[r * 4 + c];
        std::copy(m, m + 16, g.m);
    } else {
        std::complex<double> d[4];
        uint32_t act = 0;
        for (int l = 0; l < 4; l++) {
            d[sw(l)] = g.m[l];
            act |= ((g.diag_active >> l) & 1u) << sw(l);
        }
        std::copy(d, d + 4, g.m);
        g.diag_active = act;
    }
}

}  // namespace sq
