// sq_pass.hip -- the plain sweep (MEM_Runner::run(vector<Gate*>&),
// MEM_Runner.cpp:101-134) as a schedule of register passes.
//
// The planner (plan_register_passes) cuts the gate list into steps: a merged
// step is a run of one-qubit dense and diagonal gates whose dense targets lie
// in the resident qubits -- lane bits 0-2 plus up to pass_qbits + 3 chosen
// qubits, pass_qbits of them in registers and three in the top lane
// positions -- and costs one read and one write of the state
// (kernels/pass.hpp).  The encoder writes each step's ops in the pass's
// virtual positions and folds a controlled phase into the dense gate that
// follows it on either qubit (complex<double>).  Every other gate
// (two-qubit dense, permutations) runs alone with its own kernel, and so does
// a merged step that would cost no more than one pass done gate by gate.
// Per-qubit gate order is kept, so the schedule computes the same map as the
// gates in input order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "host/planner.hpp"
#include "kernels/pass.hpp"

using sq::LoweredGate;
using sq::OpClass;
using namespace sqk;

namespace {

static_assert(kPassMaxKH <= 5, "PlainLaunch holds five register qubits");
static_assert(kPassLaneHi == 3, "PlainLaunch holds three lane positions");

template <typename T>
void encode_pass_op(const LoweredGate &g, const std::vector<int> &regs, std::vector<PassOp<T>> &ops) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    auto slot = [&](int t) {
        const auto it = std::find(regs.begin(), regs.end(), t);
        return it == regs.end() ? -1 : (int)(it - regs.begin());
    };
    PassOp<T> o{};
    int nc = 0;
    auto push = [&](std::complex<double> z) { o.coef[nc++] = {(T)z.real(), (T)z.imag()}; };
    o.b = -1;
    if (g.cls == OpClass::Dense) {
        for (int i = 0; i < 4; i++) push(g.m[i]);
        const int t = g.t[0];
        // imaginary parts that are exactly zero as stored (in T)
        bool all_real = true;
        for (int i = 0; i < 4; i++) all_real = all_real && o.coef[i].im == (T)0;
        if (t < VB) {
            o.kind = PO_DENSE_VEC;
            o.c = all_real ? 15 : 0;
        } else if (t < S) {
            o.kind = PO_DENSE_LANE;
            o.a = t - VB;
            o.c = all_real ? 15 : 0;
        } else {
            o.kind = PO_DENSE_REG;
            o.a = slot(t);
            o.c = all_real ? 15 : (o.coef[0].im == (T)0 ? 1 : 0);
        }
        ops.push_back(o);
        return;
    }
    // diagonal: entry l = bit(t0) | bit(t1) << 1, t0 < t1.  The op's entries are
    // the gate's entries in the order the kernel reads them; bit k of active
    // is set when entry k is not 1.
    auto entries = [&](int s0, int s1, int s2, int s3) {
        const int src[4] = {s0, s1, s2, s3};
        o.active = 0;
        for (int k = 0; k < 4; k++) {
            push(g.m[src[k]]);
            o.active |= ((g.diag_active >> src[k]) & 1u) << k;
        }
    };
    if (g.k == 1) {
        const int t = g.t[0], j = slot(t);
        if (j >= 0) {            // e[(register bit << 1) | other bit], no other bit
            o.kind = PO_DIAG_REG1;
            o.a = j;
            o.c = 1;             // no lane-dependent bit: factors per register half
            entries(0, 0, 1, 1);
        } else {
            o.kind = PO_DIAG_FREE;
            o.a = t;
            o.c = t >= S;        // a tile bit: the same for the whole wave
            entries(0, 1, 0, 1);
        }
        ops.push_back(o);
        return;
    }
    const int t0 = g.t[0], t1 = g.t[1], j0 = slot(t0), j1 = slot(t1);
    if (j0 >= 0 && j1 >= 0) {
        o.kind = PO_DIAG_REG2;
        o.a = j0;
        o.b = j1;
        entries(0, 1, 2, 3);
    } else if (j0 >= 0) {        // register bit is entry bit 0: e[(r << 1) | y] = d[r | y << 1]
        o.kind = PO_DIAG_REG1;
        o.a = j0;
        o.b = t1;
        o.c = t1 >= S;           // y a tile bit: one factor per register half for the wave
        entries(0, 2, 1, 3);
    } else if (j1 >= 0) {        // register bit is entry bit 1: e = d
        o.kind = PO_DIAG_REG1;
        o.a = j1;
        o.b = t0;
        o.c = t0 >= S;
        entries(0, 1, 2, 3);
    } else {
        o.kind = PO_DIAG_FREE;
        o.a = t0;
        o.b = t1;
        o.c = t0 >= S && t1 >= S;
        entries(0, 1, 2, 3);
    }
    ops.push_back(o);
}

// FP operations the kernel executes per amplitude for one op (an FMA counts
// two): dense pairs take 14 per amplitude in general, 12 with m0 real and 6
// with all four coefficients real; a diagonal factor costs 6 per amplitude it
// multiplies, and register halves, register entries and tile-uniform entries
// whose factor is 1 are skipped.  A controlled phase folded into a dense op
// adds nothing.
template <typename T> double executed_flops(const PassOp<T> &o) {
    const double quarter = __builtin_popcount(o.active & 15u) / 4.0;
    switch (o.kind) {
    case PO_DENSE_REG:
    case PO_DENSE_REG_C: return o.c == 15 ? 6.0 : (o.c == 1 ? 12.0 : 14.0);
    case PO_DENSE_LANE:
    case PO_DENSE_LANE_C:
    case PO_DENSE_VEC: return o.c == 15 ? 6.0 : 14.0;
    case PO_DIAG_FREE: return o.c ? 6.0 * quarter : 6.0;
    case PO_DIAG_REG1: return o.c ? 6.0 * quarter : 3.0 * (((o.active & 3u) != 0) + ((o.active & 12u) != 0));
    case PO_DIAG_REG2: return 6.0 * quarter;
    // the lane factor (or the scalar alone) and the register table: 6 each;
    // forming L x t per amplitude when both are present: 6 more
    case PO_PHASES: return o.b ? (o.a ? 12.0 : 6.0) : 6.0;
    // three shears: 6 FMAs per pair; a scaled lane rotation: one FMA per part
    case PO_ROT_REG: return 6.0;
    case PO_ROT_LANE: return o.a >= 4 ? 6.0 : 4.0;
    default: return 0.0;
    }
}

// A one-qubit dense gate as phase, real rotation, phase:
//   U = g diag(1, zl) R diag(1, zr),   R = [[c, -s], [s, c]],  c, s >= 0,
// i.e. U = g [[c, -s zr], [s zl, c zl zr]].  The pass applies R as three
// shears (6 FMAs per pair) or, on lane bits 0-3, scaled to one FMA per part,
// instead of 14-16 operations per pair for U; the phases of all qubits of a
// pass that fall due together go into one PO_PHASES op.  ok = false when
// the coefficients are already real, or when the product of the factors
// differs from U by more than 1e-15 (a matrix that is not unitary).
struct PhaseSplit {
    bool ok = false;
    std::complex<double> g{1.0, 0.0}, zl{1.0, 0.0}, zr{1.0, 0.0};
    double c = 1.0, s = 0.0;
};
PhaseSplit phase_split(const LoweredGate &u) {
    PhaseSplit p;
    if (u.cls != OpClass::Dense || u.k != 1) return p;
    bool real = true;
    for (int i = 0; i < 4; i++) real = real && u.m[i].imag() == 0.0;
    if (real) return p;
    const std::complex<double> u00 = u.m[0], u01 = u.m[1], u10 = u.m[2], u11 = u.m[3];
    const double c = std::abs(u00), s = std::abs(u10);
    std::complex<double> g, zl(1.0, 0.0), zr(1.0, 0.0);
    if (c > 0) g = u00 / c;
    else if (s > 0) g = u10 / s;   // zl = 1
    else return p;
    if (s > 0) {
        if (c > 0) zl = u10 * std::conj(g) / s;
        zr = -u01 * std::conj(g) / s;
    } else {
        zr = u11 * std::conj(g) / c;
    }
    const std::complex<double> w[4] = {g * c, -g * s * zr, g * s * zl, g * c * zl * zr};
    for (int i = 0; i < 4; i++)
        if (std::abs(w[i] - u.m[i]) > 1e-15) return p;
    p.ok = true;
    p.g = g; p.zl = zl; p.zr = zr; p.c = c; p.s = s;
    return p;
}

// The rotation [[c, -s], [s, c]] (c, s >= 0, c^2 + s^2 = 1 to rounding: the
// shears reproduce it only then) on virtual position q as an op:
// three shears on a register slot, inside the complex64 vector and on lane
// bits 4-5 (after the swap with register bit 0); on lane bits 0-3 the
// rotation divided by c (or by s when c is small against s), which *scale is
// set to (the caller applies it with the pass's scalar).
template <typename T>
PassOp<T> rotation_op(double c, double s, int q, const std::vector<int> &regs, double &scale) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    PassOp<T> o{};
    o.b = -1;
    scale = 1.0;
    const double k = -s / (1.0 + c);
    o.coef[0] = {(T)k, (T)s};
    if (q < VB) {
        o.kind = PO_ROT_REG;
        o.c = 1;
    } else if (q < S) {
        o.kind = PO_ROT_LANE;
        o.a = q - VB;
        if (o.a < 4) {
            // R / c (one FMA per part; its error stays eps (c|x| + s|p|)) while
            // s / c <= 8 (complex<float>: s / c <= 1), else R / s with the
            // partner's sign flipped on half the lanes; the scale per gate stays
            // above 1/8 (1/sqrt(2)), far from the range limits of the pass
            o.c = (sizeof(T) == 8 ? 8.0 : 1.0) * c >= s ? 0 : 1;
            scale = o.c == 0 ? c : s;
            o.coef[0] = {(T)(o.c == 0 ? s / c : c / s), (T)0};
        }
    } else {
        o.kind = PO_ROT_REG;
        o.a = (int)(std::find(regs.begin(), regs.end(), q) - regs.begin());
    }
    return o;
}

// CP(pi) / CZ as lowered: a two-qubit diagonal whose only entry other than 1
// is -1 (within one ulp: e^{i pi} evaluates to -1 + 1.2e-16 i; leaving out
// that term changes each amplitude by less than one ulp of its magnitude).
bool is_sign_cp(const LoweredGate &d) {
    return d.cls == OpClass::Diag && d.k == 2 && d.diag_active == 8u && d.m[3].real() == -1.0 &&
           std::abs(d.m[3].imag()) <= 0x1p-52;
}
template <typename T> PassOp<T> sign_op(const LoweredGate &d, const std::vector<int> &regs, int S) {
    PassOp<T> o{};
    o.kind = PO_SIGN;
    o.pad[0] = o.pad[1] = -1;
    int nt = 0;
    for (int x = 0; x < 2; x++) {
        const int t = d.t[x];
        const auto it = std::find(regs.begin(), regs.end(), t);
        if (t < S) o.a |= 1 << t;
        else if (it != regs.end()) o.b |= 1 << (int)(it - regs.begin());
        else o.pad[nt++] = t;
    }
    return o;
}

// Appends a PO_PHASES op (and its coefficient records): fac[p] is the factor
// diag(1, fac[p]) of virtual position p, for p < S and the register slots;
// gs a scalar.  A single register factor goes into a PO_DIAG_REG1 op of its
// own (half the registers), two or more into the 2^KH table.  Returns the FP
// operations per amplitude.
template <typename T>
double emit_phases(const std::vector<std::complex<double>> &fac, const std::vector<int> &regs, int S,
                   std::complex<double> gs, std::vector<PassOp<T>> &ops) {
    static_assert(sizeof(PassOp<T>) == 8 * sizeof(cx<T>), "a coefficient record holds 8 coefficients");
    const std::complex<double> one(1.0, 0.0);
    int lane_mask = 0;
    for (int p = 0; p < S; p++)
        if (fac[p] != one) lane_mask |= 1 << p;
    std::vector<int> js;
    for (int j = 0; j < (int)regs.size(); j++)
        if (fac[regs[j]] != one) js.push_back(j);
    double flops = 0;
    if (js.size() == 1) {   // e[(register bit << 1)]: 1 for bit 0, the factor for bit 1
        PassOp<T> o{};
        o.kind = PO_DIAG_REG1;
        o.a = js[0];
        o.b = -1;
        o.c = 1;
        const std::complex<double> z = fac[regs[js[0]]];
        o.coef[0] = o.coef[1] = {(T)1, (T)0};
        o.coef[2] = o.coef[3] = {(T)z.real(), (T)z.imag()};
        o.active = 12u;
        ops.push_back(o);
        flops += executed_flops(o);
    }
    const bool table = js.size() >= 2;
    if (!lane_mask && !table && gs == one) return flops;
    auto rec = [](const std::complex<double> *z, int n) {
        PassOp<T> r{};
        cx<T> w[8];
        for (int i = 0; i < 8; i++) w[i] = {(T)(i < n ? z[i].real() : 1.0), (T)(i < n ? z[i].imag() : 0.0)};
        std::memcpy(static_cast<void *>(&r), w, sizeof w);
        return r;
    };
    PassOp<T> o{};
    o.kind = PO_PHASES;
    o.a = lane_mask;
    o.b = table;
    const int R = 1 << regs.size();
    o.c = (lane_mask ? 1 : 0) + (table ? (R + 7) / 8 : 0);
    const std::complex<double> l0 = lane_mask || !table ? gs : one;   // else the scalar goes into the table
    o.coef[0] = {(T)l0.real(), (T)l0.imag()};
    ops.push_back(o);
    flops += executed_flops(o);
    if (lane_mask) {
        std::complex<double> z[8];
        for (int p = 0; p < 8; p++) z[p] = p < S ? fac[p] : one;
        ops.push_back(rec(z, 8));
    }
    if (table) {
        std::vector<std::complex<double>> t(R, lane_mask ? one : gs);
        for (int r = 0; r < R; r++)
            for (int j : js)
                if ((r >> j) & 1) t[r] *= fac[regs[j]];
        for (int k = 0; k < (R + 7) / 8; k++) ops.push_back(rec(t.data() + 8 * k, std::min(8, R - 8 * k)));
    }
    return flops;
}

template <typename T, int KH>
int launch_pass(sq_engine *e, const PassArgs<T> &a) {
    const uint64_t tiles = a.n_tiles * (a.s2 ? 2 : 1);
    const unsigned grid = e->grid_flat(tiles, kBlock / 64);
    if ((uint64_t)grid * (kBlock / 64) < tiles)   // one tile per wave, no grid-stride loop
        return e->fail(SQ_ERR_ARG, "register pass: " + std::to_string(tiles) + " wave tiles exceed one launch");
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_pass<T, KH>), grid, kBlock, 0, e->stream,
                       reinterpret_cast<cx<T> *>(e->d_state), a);
    return e->check_launch("k_pass");
}

template <typename T>
int build_plain_t(sq_engine *e, const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low,
                  PlainProgram &prog, void *dst, size_t dst_cap) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    const int n = e->n;
    const int kh = e->pass_qbits;
    prog = PlainProgram{};
    prog.n = n;
    prog.kh = kh;
    if (e->sweep_mode == 1 || n < S + kh) {
        prog.per_gate = true;
        return SQ_OK;
    }
    const size_t ng = gates.size();
    std::vector<char> mergeable(ng, 0), anywhere(ng, 0);
    for (size_t i = 0; i < ng; i++) {
        const LoweredGate &g = low[i];
        if (g.cls == OpClass::Diag && g.k <= 2) mergeable[i] = anywhere[i] = 1;
        else if (g.cls == OpClass::Dense && g.k == 1) mergeable[i] = 1;
    }
    // the top three lane positions may hold any qubit up to bit 27 of the
    // 16-byte slot index (a lane's offset stays a 32-bit one)
    const int moves = std::max(0, std::min(e->lane_moves, kPassLaneHi));
    // The schedule depends on the gates' qubits and classes only: a list
    // with the structure of a recent one (the same circuit again, or with new
    // angles) reuses its schedule; the op records are encoded every call.
    std::vector<int32_t> key = {n, S, kh, moves, std::min(n - 1, 27 + VB), e->plan_beam, (int32_t)ng};
    key.reserve(6 + 6 * ng);
    for (size_t i = 0; i < ng; i++) {
        key.push_back(mergeable[i] | (anywhere[i] << 1) | (gates[i].nq << 2));
        for (int j = 0; j < gates[i].nq && j < 4; j++) key.push_back(gates[i].q[j]);
    }
    const std::vector<sq::SweepStep> *steps_ptr = nullptr;
    for (size_t c = 0; c < e->plan_cache.size(); c++) {
        if (e->plan_cache[c].key != key) continue;
        if (c) std::rotate(e->plan_cache.begin(), e->plan_cache.begin() + c, e->plan_cache.begin() + c + 1);
        steps_ptr = &e->plan_cache[0].steps;
        break;
    }
    if (!steps_ptr) {
        sq_engine::PlanCacheEntry ent;
        ent.steps = sq::plan_register_passes(gates, mergeable, anywhere, n, S, kh, moves, std::min(n - 1, 27 + VB),
                                             e->plan_beam);
        ent.key = std::move(key);
        e->plan_cache.insert(e->plan_cache.begin(), std::move(ent));
        if (e->plan_cache.size() > 4) e->plan_cache.pop_back();
        steps_ptr = &e->plan_cache[0].steps;
    }
    const std::vector<sq::SweepStep> &steps = *steps_ptr;

    std::vector<PassOp<T>> ops;
    for (const auto &st : steps) {
        double frac = 0;
        for (int gi : st.gates) frac += low[gi].pass_fraction;
        if (!st.merged || frac <= 1.0) {
            for (int gi : st.gates) {
                PlainLaunch l;
                l.gate = gi;
                prog.launches.push_back(l);
            }
            continue;
        }
        // Virtual positions: lane position S-3+k holds qubit st.lanes[k], and
        // that qubit's own position holds the one it displaced.
        std::vector<int> vpos(n);
        for (int q = 0; q < n; q++) vpos[q] = q;
        PlainLaunch l;
        for (int k = 0; k < kPassLaneHi; k++) {
            const int u = S - kPassLaneHi + k;
            const int off = k - (kPassLaneHi - moves);   // st.lanes covers the top `moves` positions
            const int w = off >= 0 ? st.lanes[off] : u;
            if (w != u) std::swap(vpos[u], vpos[w]);
        }
        l.v2l.assign(n, 0);
        for (int q = 0; q < n; q++) l.v2l[vpos[q]] = q;
        // exactly kh register qubits: pad with qubits outside the lane positions
        std::vector<int> regs;
        for (int q : st.regs) regs.push_back(vpos[q]);
        for (int q = 0; q < n && (int)regs.size() < kh; q++)
            if (vpos[q] >= S && std::find(regs.begin(), regs.end(), vpos[q]) == regs.end()) regs.push_back(vpos[q]);
        std::sort(regs.begin(), regs.end());
        l.merged = true;
        l.op_begin = (int)ops.size();
        std::vector<LoweredGate> lg;
        lg.reserve(st.gates.size());
        for (int gi : st.gates) {
            lg.push_back(low[gi]);
            sq::relabel(lg.back(), vpos);
        }
        const std::complex<double> one(1.0, 0.0);
        std::vector<PhaseSplit> sp(lg.size());
        bool any_split = false;
        if (e->phase_split)
            for (size_t i = 0; i < lg.size(); i++) {
                sp[i] = phase_split(lg[i]);
                any_split = any_split || sp[i].ok;
            }
        // complex<double>: a controlled phase (diagonal, only |11> not 1) whose
        // next gate on either qubit is a dense gate on a register slot or a
        // lane bit goes into that gate's op, with the other qubit as the control
        // (not into a split gate: its rotation runs as shears)
        std::vector<int> fold_into(lg.size(), -1), folded_from(lg.size(), -1);
        if constexpr (sizeof(T) == 8) {
            auto in_regs = [&](int t) { return std::find(regs.begin(), regs.end(), t) != regs.end(); };
            for (size_t i = 0; i < lg.size(); i++) {
                const LoweredGate &d = lg[i];
                if (!(d.cls == OpClass::Diag && d.k == 2 && d.diag_active == 8u)) continue;
                for (size_t j = i + 1; j < lg.size(); j++) {
                    const LoweredGate &u = lg[j];
                    bool touches = false;
                    for (int x = 0; x < u.k; x++) touches = touches || u.t[x] == d.t[0] || u.t[x] == d.t[1];
                    if (!touches) continue;
                    if (u.cls == OpClass::Dense && u.k == 1 && folded_from[j] < 0 && !sp[j].ok &&
                        (in_regs(u.t[0]) || (u.t[0] >= VB && u.t[0] < S))) {
                        fold_into[i] = (int)j;
                        folded_from[j] = (int)i;
                    }
                    break;
                }
            }
        }
        // Phase split (phase_split above): each dense gate's rotation runs
        // as R; its phases wait per qubit in pend[] until a flush applies all
        // factors due -- placed just before the first gate that needs one, it
        // also takes, for every other resident qubit, the right factor of that
        // qubit's next gate and the one-qubit diagonals before it (all of them
        // diagonal, so they commute up to that gate).  The scalars collect in
        // gpass; the pass ends with a flush of what is left.
        auto resident = [&](int t) { return t < S || std::find(regs.begin(), regs.end(), t) != regs.end(); };
        std::vector<std::complex<double>> pend(n, one);
        std::vector<char> absorbed(lg.size(), 0), zr_done(lg.size(), 0);
        std::complex<double> gpass = one;
        auto diag1 = [&](size_t j) {
            const LoweredGate &d = lg[j];
            return any_split && d.cls == OpClass::Diag && d.k == 1 && resident(d.t[0]) && d.m[0] != 0.0;
        };
        auto flush = [&](size_t from, bool last) {
            std::vector<std::complex<double>> fac(n, one);
            auto take = [&](int q) {
                std::complex<double> f = pend[q];
                pend[q] = one;
                for (size_t j = from; j < lg.size(); j++) {
                    const LoweredGate &u = lg[j];
                    if (u.cls == OpClass::Diag) {
                        if (diag1(j) && u.t[0] == q && !absorbed[j]) {
                            f *= u.m[1] / u.m[0];
                            gpass *= u.m[0];
                            absorbed[j] = 1;
                        }
                        continue;
                    }
                    if (u.t[0] != q) continue;
                    if (sp[j].ok && !zr_done[j]) {
                        f *= sp[j].zr;
                        zr_done[j] = 1;
                    }
                    break;
                }
                fac[q] = f;
            };
            for (int p = 0; p < S; p++) take(p);
            for (int r : regs) take(r);
            l.flops_per_amp += emit_phases<T>(fac, regs, S, last ? gpass : one, ops);
        };
        for (size_t i = 0; i < lg.size(); i++) {
            if (fold_into[i] >= 0) continue;   // applied inside the dense gate's op: no work of its own
            if (diag1(i)) {                    // a factor of its qubit's pending phase
                if (!absorbed[i]) {
                    pend[lg[i].t[0]] *= lg[i].m[1] / lg[i].m[0];
                    gpass *= lg[i].m[0];
                    absorbed[i] = 1;
                }
                continue;
            }
            // Rotations as shears (or scaled) for both types; CP(pi) as a sign
            // flip for complex<double> only: complex<float> keeps the diagonal
            // op, which measured faster there (FP32 arithmetic is cheap next to
            // the bit operations)
            constexpr bool kShears = true, kSigns = sizeof(T) == 8;
            if (kSigns && any_split && is_sign_cp(lg[i])) {   // CP(pi): a sign flip
                ops.push_back(sign_op<T>(lg[i], regs, S));
                continue;
            }
            LoweredGate gi = lg[i];
            if (any_split && gi.cls == OpClass::Dense) {
                const int q = gi.t[0];
                std::complex<double> need = pend[q];
                if (sp[i].ok && !zr_done[i]) need *= sp[i].zr;
                if (need != one) flush(i, false);
                pend[q] = one;
                zr_done[i] = 1;
                if (sp[i].ok) {
                    pend[q] = sp[i].zl;
                    gpass *= sp[i].g;
                    const double c = sp[i].c, s = sp[i].s;
                    if (c == 1.0 && s == 0.0) continue;   // R = I
                    if (kShears && std::abs(c * c + s * s - 1.0) <= 1e-15) {   // a rotation: shears or scaled
                        double scale = 1.0;
                        ops.push_back(rotation_op<T>(c, s, q, regs, scale));
                        gpass *= scale;
                        l.flops_per_amp += executed_flops(ops.back());
                        continue;
                    }
                    // complex<float>, or a multiple of a rotation (a matrix that
                    // is not unitary): the real 2x2 product
                    gi.m[0] = c;
                    gi.m[1] = -s;
                    gi.m[2] = s;
                    gi.m[3] = c;
                }
            }
            encode_pass_op<T>(gi, regs, ops);
            if constexpr (sizeof(T) == 8) if (folded_from[i] >= 0) {
                const LoweredGate &d = lg[folded_from[i]];
                PassOp<T> &o = ops.back();
                const int ctl = gi.t[0] == d.t[0] ? d.t[1] : d.t[0];
                const auto it = std::find(regs.begin(), regs.end(), ctl);
                const bool lane_target = o.kind == PO_DENSE_LANE;
                o.kind = lane_target ? PO_DENSE_LANE_C : PO_DENSE_REG_C;
                if (it != regs.end()) { o.active = 0; o.b = (int)(it - regs.begin()); }
                else { o.active = ctl < S ? 1 : 2; o.b = ctl; }
                // A real gate stays real with a real phase.  A phase within one
                // ulp of the real axis counts as real: CP(pi) evaluates to
                // e^{i pi} = -1 + 1.2e-16 i, and leaving out that term changes
                // each amplitude by less than one ulp of its magnitude.
                const std::complex<double> ph = d.m[3];
                const bool ph_real = std::abs(ph.imag()) <= 0x1p-52 * std::abs(ph.real());
                const bool real = o.c == 15 && ph_real;
                o.c = real ? 15 : (!lane_target && o.coef[0].im == (T)0 ? 1 : 0);
                o.coef[kPassFoldPhase] = {(T)ph.real(), real ? (T)0 : (T)ph.imag()};
            }
            l.flops_per_amp += executed_flops(ops.back());
        }
        if (any_split) flush(lg.size(), true);
        l.n_ops = (int)ops.size() - l.op_begin;
        for (int j = 0; j < kh; j++) l.hpos[j] = regs[j];
        prog.launches.push_back(l);
    }

    // op census of this program (sq_pass_census)
    e->pass_census.fill(0);
    for (size_t oi = 0; oi < ops.size(); oi++) {
        const PassOp<T> &o = ops[oi];
        int idx = -1;
        switch (o.kind) {
        case PO_DENSE_REG: idx = 0; break;
        case PO_DENSE_LANE: idx = o.a < 4 ? 1 : 2; break;
        case PO_DENSE_VEC: idx = 3; break;
        case PO_DIAG_FREE: idx = 4; break;
        case PO_DIAG_REG1: idx = 5; break;
        case PO_DIAG_REG2: idx = 6; break;
        case PO_PHASES: idx = 7; oi += o.c; break;   // past its coefficient records
        case PO_ROT_REG: idx = o.c == 1 ? 3 : 0; break;
        case PO_ROT_LANE: idx = o.a < 4 ? 1 : 2; break;
        case PO_SIGN: idx = o.b == 0 ? 4 : (__builtin_popcount((unsigned)o.b) == 1 ? 5 : 6); break;
        case PO_DENSE_REG_C:
        case PO_DENSE_LANE_C: {
            const int t = o.kind == PO_DENSE_REG_C ? 0 : (o.a < 4 ? 1 : 2);
            const int c = o.active == 0 ? (o.b == 0 ? 0 : 1) : (o.active == 1 ? 2 : 3);
            idx = 8 + 4 * t + c;
            break;
        }
        default: break;
        }
        if (idx >= 0) e->pass_census[idx]++;
    }

    // upload the program once; earlier launches may still read the old one
    const size_t need = ops.size() * sizeof(PassOp<T>) + 256;
    if (dst) {   // caller-owned buffer: never grown
        if (ops.size() * sizeof(PassOp<T>) > dst_cap)
            return e->fail(SQ_ERR_ARG, "the register-pass program (" + std::to_string(ops.size() * sizeof(PassOp<T>)) +
                                           " bytes) exceeds the " + std::to_string(dst_cap) + "-byte scratch");
        prog.d_ops = dst;
        if (!ops.empty()) {
            if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
            hipError_t he = hipMemcpy(dst, ops.data(), ops.size() * sizeof(PassOp<T>), hipMemcpyHostToDevice);
            if (he != hipSuccess) return e->hip_fail(he, "hipMemcpy(pass program to scratch)");
        }
        return SQ_OK;
    }
    if (!ops.empty()) {
        if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
        if (need > e->passprog_cap) {
            if (e->d_passprog) (void)hipFree(e->d_passprog);
            e->d_passprog = nullptr;
            e->passprog_cap = 0;
            hipError_t he = hipMalloc(&e->d_passprog, need * 2);
            if (he != hipSuccess) return e->hip_fail(he, "hipMalloc(pass program)");
            e->passprog_cap = need * 2;
        }
        char *d = static_cast<char *>(e->d_passprog);
        hipError_t he = hipMemcpy(d, ops.data(), ops.size() * sizeof(PassOp<T>), hipMemcpyHostToDevice);
        if (he != hipSuccess) return e->hip_fail(he, "hipMemcpy(pass program)");
    }
    prog.d_ops = e->d_passprog;
    return SQ_OK;
}

// Runs a program on the engine's current state buffer (e->n must be the
// program's n), and on e->d_state2 too when that is set (one launch covers
// both); its op records are at prog.d_ops.
template <typename T>
int run_plain_t(sq_engine *e, const PlainProgram &prog, const std::vector<LoweredGate> &low) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    const int n = e->n, kh = prog.kh;
    if (n != prog.n) return e->fail(SQ_ERR_ARG, "internal: pass program planned for another size");
    // one launch per gate: on each state in turn
    auto apply_one = [&](const LoweredGate &g) {
        int rc = e->apply_lowered(g);
        if (rc || !e->d_state2) return rc;
        std::swap(e->d_state, e->d_state2);
        rc = e->apply_lowered(g);
        std::swap(e->d_state, e->d_state2);
        return rc;
    };
    if (prog.per_gate) {
        for (const auto &g : low) {
            int rc = apply_one(g);
            if (rc) return rc;
        }
        return SQ_OK;
    }
    const PassOp<T> *dops = reinterpret_cast<const PassOp<T> *>(prog.d_ops);
    char nm[48];
    std::snprintf(nm, sizeof nm, "pass%d_%s", kh, e->dtype == SQ_C128 ? "c128" : "c64");
    const double vecs = e->d_state2 ? 2.0 : 1.0;
    const double bytes = 2.0 * e->amp_bytes() * std::ldexp(1.0, n) * vecs;
    for (const auto &l : prog.launches) {
        if (!l.merged) {
            int rc = apply_one(low[l.gate]);
            if (rc) return rc;
            continue;
        }
        if (l.n_ops == 0) continue;   // the kernel reads its first op unconditionally
        PassArgs<T> a{};
        a.s2 = reinterpret_cast<cx<T> *>(e->d_state2);
        a.ops = dops + l.op_begin;
        a.n_ops = l.n_ops;
        a.n_tile_bits = n - S - kh;
        a.n_tiles = 1ULL << a.n_tile_bits;
        if ((int)l.v2l.size() != n || a.n_tile_bits > kPassMaxTileBits) return e->fail(SQ_ERR_ARG, "internal: pass layout");
        // the qubit a virtual position stands for lies at its own index bit
        for (int j = 0; j < kh; j++) {
            a.hpos[j] = l.hpos[j];
            a.reg_pos[j] = l.v2l[l.hpos[j]];
        }
        for (int k = 0; k < 6; k++) {
            // the kernel's lane offset holds slot bits up to kPassMaxSlotBit only
            a.lane_pos[k] = l.v2l[VB + k] - VB;
            if (a.lane_pos[k] < 0 || a.lane_pos[k] > kPassMaxSlotBit)
                return e->fail(SQ_ERR_ARG, "internal: lane position out of range");
        }
        for (int v = S, k = 0; v < n; v++) {
            if (std::find(l.hpos, l.hpos + kh, v) != l.hpos + kh) continue;
            a.tile_pos[k++] = l.v2l[v];
        }
        int slot = e->stat_begin(nm, bytes, l.flops_per_amp * std::ldexp(1.0, n) * vecs);
        int rc;
        switch (kh) {
        case 1: rc = launch_pass<T, 1>(e, a); break;
        case 2: rc = launch_pass<T, 2>(e, a); break;
        case 3: rc = launch_pass<T, 3>(e, a); break;
        case 4: rc = launch_pass<T, 4>(e, a); break;
        case 5: rc = launch_pass<T, 5>(e, a); break;
        default: rc = e->fail(SQ_ERR_ARG, "internal: bad pass width");
        }
        e->stat_end(slot);
        if (rc) return rc;
    }
    return SQ_OK;
}

}  // namespace

int sq_engine::build_plain(const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low, PlainProgram &prog,
                           void *dst, size_t dst_cap) {
    return dtype == SQ_C128 ? build_plain_t<double>(this, gates, low, prog, dst, dst_cap)
                            : build_plain_t<float>(this, gates, low, prog, dst, dst_cap);
}

// The second state buffer of the out-of-place permutation pass, allocated
// once when a second copy of the state fits in HBM with room to spare.
// SQ_ERR_UNSUPPORTED (no error recorded) when it does not.
int sq_engine::ensure_alt() {
    if (d_alt) return SQ_OK;
    if (out_of_core() || !bitperm_ok) return SQ_ERR_UNSUPPORTED;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void)hipGetLastError(); return SQ_ERR_UNSUPPORTED; }
    if (free_b < state_bytes() + ((size_t)4 << 30)) return SQ_ERR_UNSUPPORTED;
    if (hipMalloc(&d_alt, state_bytes()) != hipSuccess) {
        (void)hipGetLastError();
        d_alt = nullptr;
        return SQ_ERR_UNSUPPORTED;
    }
    return SQ_OK;
}

int sq_engine::run_plain_program(const PlainProgram &prog, const std::vector<LoweredGate> &low) {
    return dtype == SQ_C128 ? run_plain_t<double>(this, prog, low) : run_plain_t<float>(this, prog, low);
}

namespace {

// The target of a U2 whose matrix is exactly a CNOT (control either qubit),
// or -1.  Local index bit 0 is t[0], bit 1 is t[1].
int cnot_target(const LoweredGate &g) {
    if (g.cls != OpClass::Dense || g.k != 2) return -1;
    const std::complex<double> one(1.0, 0.0), zero(0.0, 0.0);
    auto is = [&](const int *perm) {   // row r has its one at column perm[r]
        for (int r = 0; r < 4; r++)
            for (int c = 0; c < 4; c++)
                if (g.m[r * 4 + c] != (c == perm[r] ? one : zero)) return false;
        return true;
    };
    static const int flip1[4] = {0, 3, 2, 1};   // control t[0]: bit 1 flips where bit 0 is set
    static const int flip0[4] = {0, 1, 3, 2};   // control t[1]: bit 0 flips where bit 1 is set
    if (is(flip1)) return g.t[1];
    if (is(flip0)) return g.t[0];
    return -1;
}

}  // namespace

// Register passes carry one-qubit dense and diagonal gates; a CNOT (written
// as U2, the reference's way to give one) is H_t CZ H_t with exact 1/sqrt(2)
// coefficients, so it joins a pass (on complex128 the CZ folds into the
// second H) instead of taking a pass of its own.  The products differ from
// the dense 4x4 application by rounding only (|H H - I| = 2^-52).
bool sq_engine::expand_cnots(const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low,
                             std::vector<sq_gate> &g2, std::vector<LoweredGate> &l2) const {
    const int S = (dtype == SQ_C64 ? 1 : 0) + 6;
    if (sweep_mode != 0 || n < S + pass_qbits) return false;   // no register passes
    bool any = false;
    for (const auto &g : low) any = any || cnot_target(g) >= 0;
    if (!any) return false;
    const double s = 0.70710678118654752440;
    g2.clear();
    l2.clear();
    for (size_t i = 0; i < low.size(); i++) {
        const int t = cnot_target(low[i]);
        if (t < 0) {
            g2.push_back(gates[i]);
            l2.push_back(low[i]);
            continue;
        }
        LoweredGate h{};
        h.cls = OpClass::Dense;
        h.k = 1;
        h.t[0] = t;
        h.m[0] = h.m[1] = h.m[2] = s;
        h.m[3] = -s;
        h.pass_fraction = 1.0;
        h.name = "H (CNOT)";
        LoweredGate cz{};
        cz.cls = OpClass::Diag;
        cz.k = 2;
        cz.t[0] = low[i].t[0];
        cz.t[1] = low[i].t[1];
        cz.m[0] = cz.m[1] = cz.m[2] = 1.0;
        cz.m[3] = -1.0;
        cz.diag_active = 8u;
        cz.pass_fraction = 0.25;
        cz.name = "CZ (CNOT)";
        sq_gate gh = gates[i], gz = gates[i];
        gh.kind = SQ_GATE_U1;
        gh.nq = 1;
        gh.q[0] = t;
        gz.kind = SQ_GATE_CP;
        gz.q[0] = cz.t[0];
        gz.q[1] = cz.t[1];
        g2.insert(g2.end(), {gh, gz, gh});
        l2.insert(l2.end(), {h, cz, h});
    }
    return true;
}

int sq_engine::apply_plain(const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low) {
    std::vector<sq_gate> g2;
    std::vector<LoweredGate> l2;
    const bool x = expand_cnots(gates, low, g2, l2);
    PlainProgram prog;
    int rc = build_plain(x ? g2 : gates, x ? l2 : low, prog);
    return rc ? rc : run_plain_program(prog, x ? l2 : low);
}

// ---------------------------------------------------------------------------
// sq_apply_gates_noalloc: the plain sweep with its register-pass program in a
// fixed device array instead of the engine's growable buffer, so the call
// allocates nothing.  Every engine on a device shares that device's copy of
// the array; it keeps the last call's program.  Calls that overlap on one
// device (two host threads, or two engines on different streams) race on it.
// ---------------------------------------------------------------------------

namespace {

constexpr size_t kScratchBytes = (size_t)4 << 20;   // 32768 complex128 / 65536 complex64 op records
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex scratch_mu;
void *scratch_addr[kScratchDevices] = {};   // per device: the array's device address

}  // namespace

// Called by sq_create with the engine's device current.  The array's name is
// a host-side stand-in; its device address (one per device) comes from
// hipGetSymbolAddress, once per device.
int sq_engine::setup_scratch() {
    if (device < 0 || device >= kScratchDevices) return fail(SQ_ERR_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t he = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (he != hipSuccess || !p) return hip_fail(he != hipSuccess ? he : hipErrorInvalidValue, "hipGetSymbolAddress(scratch)");
        scratch_addr[device] = p;
    }
    d_scratch = scratch_addr[device];
    return SQ_OK;
}

extern "C" int sq_apply_gates_noalloc(sq_engine *e, const sq_gate *gates, uint64_t n) {
    if (!e || (!gates && n)) return SQ_ERR_ARG;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "sq_apply_gates_noalloc needs an HBM-resident state");
    if (!e->d_scratch) return e->fail(SQ_ERR_HIP, "scratch array not set up for this device");
    if (hipSetDevice(e->device) != hipSuccess) return e->fail(SQ_ERR_HIP, "hipSetDevice");
    // validate everything first so a bad line (or a program too large for the
    // scratch) leaves the state untouched
    std::vector<LoweredGate> low(n);
    for (uint64_t i = 0; i < n; i++) {
        int rc = e->lower(gates[i], low[i]);
        if (rc) { e->err += " (gate " + std::to_string(i) + ")"; return rc; }
    }
    const std::vector<sq_gate> gl(gates, gates + n);
    std::vector<sq_gate> g2;
    std::vector<LoweredGate> l2;
    const bool x = e->expand_cnots(gl, low, g2, l2);
    PlainProgram prog;
    int rc = e->build_plain(x ? g2 : gl, x ? l2 : low, prog, e->d_scratch, kScratchBytes);
    return rc ? rc : e->run_plain_program(prog, x ? l2 : low);
}

extern "C" int sq_pass_census(sq_engine *e, int64_t *counts, int32_t cap, int32_t *n_out) {
    if (!e || (cap > 0 && !counts)) return SQ_ERR_ARG;
    const int32_t n = (int32_t)e->pass_census.size();
    if (n_out) *n_out = n;
    for (int32_t i = 0; i < std::min(cap, n); i++) counts[i] = e->pass_census[i];
    return cap < n ? SQ_ERR_ARG : SQ_OK;
}
