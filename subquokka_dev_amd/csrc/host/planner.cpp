// planner.cpp -- see planner.hpp.
#include "planner.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "gate_lowering.hpp"

namespace sq {

namespace {
const char *kind_name(int k) {
    static const char *names[] = {"?", "H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U1", "U2",
                                  "SWAP", "CP", "RZZ", "VSWAP_1_1", "VSWAP_2_2"};
    return (k >= 1 && k <= 15) ? names[k] : "?";
}

void appendf(std::string &s, const char *fmt, double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, fmt, v);
    s += buf;
}
}  // namespace

std::vector<int> gate_qubits(const sq_gate &g) {
    return std::vector<int>(g.q, g.q + std::max(0, std::min(4, g.nq)));
}

std::string gate_to_line(const sq_gate &g) {
    std::string s = kind_name(g.kind);
    for (int i = 0; i < g.nq; i++) s += " " + std::to_string(g.q[i]);
    switch (g.kind) {
    case SQ_GATE_P: case SQ_GATE_RX: case SQ_GATE_RY: case SQ_GATE_RZ: case SQ_GATE_CP: case SQ_GATE_RZZ:
        appendf(s, " %.17g", g.param);
        break;
    case SQ_GATE_U1:
        for (int i = 0; i < 8; i++) appendf(s, " %.17g", g.mat[i]);
        break;
    case SQ_GATE_U2:
        for (int i = 0; i < 32; i++) appendf(s, " %.17g", g.mat[i]);
        break;
    default:
        break;
    }
    return s;
}

namespace {

// Gates runnable with qubit set `in` from the pending list, in order.
int count_runnable(const std::vector<sq_gate> &gates, const std::vector<int> &pending,
                   const std::vector<char> &in, std::vector<char> &blocked, std::vector<int> *out) {
    std::fill(blocked.begin(), blocked.end(), 0);
    int cnt = 0;
    for (int gi : pending) {
        const sq_gate &g = gates[gi];
        bool ok = true;
        for (int i = 0; i < g.nq; i++) if (!in[g.q[i]] || blocked[g.q[i]]) { ok = false; break; }
        if (ok) {
            cnt++;
            if (out) out->push_back(gi);
        } else {
            for (int i = 0; i < g.nq; i++) blocked[g.q[i]] = 1;
        }
    }
    return cnt;
}

}  // namespace

std::vector<PlanPass> plan_passes(const std::vector<sq_gate> &gates, int n, int budget,
                                  const std::vector<int> &fixed) {
    std::vector<PlanPass> passes;
    std::vector<int> pending(gates.size());
    for (size_t i = 0; i < gates.size(); i++) pending[i] = (int)i;
    std::vector<char> in(n), blocked(n);
    while (!pending.empty()) {
        std::fill(in.begin(), in.end(), 0);
        int size = 0;
        for (int q : fixed) if (!in[q]) { in[q] = 1; size++; }
        int best = count_runnable(gates, pending, in, blocked, nullptr);
        while (size < budget) {
            int best_q = -1, best_c = best;
            for (int q = 0; q < n; q++) {
                if (in[q]) continue;
                in[q] = 1;
                int c = count_runnable(gates, pending, in, blocked, nullptr);
                in[q] = 0;
                if (c > best_c) { best_c = c; best_q = q; }
            }
            if (best_q >= 0) {
                in[best_q] = 1; size++; best = best_c;
                continue;
            }
            // No single qubit helps: add every missing qubit of the first
            // blocked gate if they fit (a two-qubit gate needs both at once).
            bool added = false;
            for (int gi : pending) {
                const sq_gate &g = gates[gi];
                int missing = 0;
                for (int i = 0; i < g.nq; i++) missing += !in[g.q[i]];
                if (missing == 0) continue;
                if (size + missing > budget) break;
                for (int i = 0; i < g.nq; i++) if (!in[g.q[i]]) { in[g.q[i]] = 1; size++; }
                best = count_runnable(gates, pending, in, blocked, nullptr);
                added = true;
                break;
            }
            if (!added) break;
        }
        PlanPass p;
        count_runnable(gates, pending, in, blocked, &p.gates);
        if (p.gates.empty()) {
            // The first pending gate alone exceeds the budget with the fixed set:
            // schedule it by itself (callers size budget >= |fixed| + 4).
            p.gates.push_back(pending.front());
            for (int q : gate_qubits(gates[pending.front()])) in[q] = 1;
        }
        for (int q = 0; q < n; q++) if (in[q]) p.qubits.push_back(q);
        std::vector<int> rest;
        size_t k = 0;
        for (int gi : pending) {
            if (k < p.gates.size() && p.gates[k] == gi) { k++; continue; }
            rest.push_back(gi);
        }
        pending.swap(rest);
        passes.push_back(std::move(p));
    }
    return passes;
}

namespace {

// Gates of `pending` that a register pass with resident qubit set `in` can
// run, in order.  Stops once every qubit is blocked.
int count_pass(const std::vector<sq_gate> &gates, const std::vector<int> &pending, const std::vector<char> &in,
               const std::vector<char> &mergeable, const std::vector<char> &anywhere, std::vector<char> &blocked,
               std::vector<int> *out, int *dense = nullptr) {
    std::fill(blocked.begin(), blocked.end(), 0);
    const int n = (int)blocked.size();
    int cnt = 0, n_blocked = 0, n_dense = 0;
    for (int gi : pending) {
        const sq_gate &g = gates[gi];
        bool ok = mergeable[gi] != 0;
        for (int i = 0; ok && i < g.nq; i++)
            if (blocked[g.q[i]] || (!anywhere[gi] && !in[g.q[i]])) ok = false;
        if (ok) {
            cnt++;
            n_dense += !anywhere[gi];
            if (out) out->push_back(gi);
            continue;
        }
        for (int i = 0; i < g.nq; i++)
            if (!blocked[g.q[i]]) { blocked[g.q[i]] = 1; n_blocked++; }
        if (n_blocked == n) break;
    }
    if (dense) *dense = n_dense;
    return cnt;
}

}  // namespace

namespace {

// Splits a chosen qubit set between the register slots (at most kh) and the
// movable lane positions [lo, lo + nlh): qubits already in a lane position
// stay there; when more than kh others remain, those with the fewest dense
// gates (lane gates cost more) and at most lh_max take the free lane
// positions.  Returns false when the set does not fit.
bool assign_lanes(const std::vector<int> &chosen, const std::vector<int> &dense_cnt, int lo, int nlh, int kh,
                  int lh_max, std::vector<int> *regs, std::vector<int> *lanes) {
    std::vector<int> lane(nlh, -1), rest;
    for (int q : chosen) {
        if (q >= lo && q < lo + nlh) lane[q - lo] = q;
        else rest.push_back(q);
    }
    int need = (int)rest.size() - kh;
    if (need > 0) {
        std::vector<int> cand;
        for (int q : rest) if (q <= lh_max) cand.push_back(q);
        std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) { return dense_cnt[x] < dense_cnt[y]; });
        int free_slots = 0;
        for (int k = 0; k < nlh; k++) free_slots += lane[k] < 0;
        if ((int)cand.size() < need || free_slots < need) return false;
        std::vector<char> moved(dense_cnt.size(), 0);
        for (int i = 0, k = 0; i < need; i++) {
            while (lane[k] >= 0) k++;
            lane[k] = cand[i];
            moved[cand[i]] = 1;
        }
        std::vector<int> keep;
        for (int q : rest) if (!moved[q]) keep.push_back(q);
        rest.swap(keep);
    }
    if (regs) *regs = rest;
    if (lanes) {
        // a free position keeps its own qubit unless that qubit is a register
        std::vector<char> used(dense_cnt.size(), 0);
        for (int q : rest) used[q] = 1;
        for (int k = 0; k < nlh; k++) if (lane[k] >= 0) used[lane[k]] = 1;
        for (int k = 0; k < nlh; k++) {
            if (lane[k] >= 0) continue;
            if (!used[lo + k]) { lane[k] = lo + k; used[lo + k] = 1; }
        }
        // positions whose own qubit became a register take an untouched qubit
        int pad = lo + nlh;
        for (int k = 0; k < nlh; k++) {
            if (lane[k] >= 0) continue;
            while (pad < (int)used.size() && (used[pad] || pad > lh_max)) pad++;
            if (pad >= (int)used.size()) return false;
            lane[k] = pad;
            used[pad] = 1;
        }
        *lanes = lane;
    }
    return true;
}

// The other split: the |chosen| - kh lane-capable qubits with the fewest
// dense gates take lane positions, even when that sends a qubit already in a
// lane position to a register.  Returns false when it does not apply.
bool assign_lanes_min_dense(const std::vector<int> &chosen, const std::vector<int> &dense_cnt, int lo, int nlh,
                            int kh, int lh_max, std::vector<int> *regs, std::vector<int> *lanes) {
    const int m = (int)chosen.size() - kh;
    if (m <= 0 || m > nlh) return false;
    auto in_lane_pos = [&](int q) { return q >= lo && q < lo + nlh; };
    std::vector<int> cand;
    for (int q : chosen) if (q <= lh_max || in_lane_pos(q)) cand.push_back(q);
    if ((int)cand.size() < m) return false;
    std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) {
        if (dense_cnt[x] != dense_cnt[y]) return dense_cnt[x] < dense_cnt[y];
        return in_lane_pos(x) && !in_lane_pos(y);
    });
    const int nq = (int)dense_cnt.size();
    std::vector<char> is_lane(nq, 0), is_chosen(nq, 0), used(nq, 0);
    for (int i = 0; i < m; i++) is_lane[cand[i]] = 1;
    for (int q : chosen) is_chosen[q] = used[q] = 1;
    std::vector<int> lane(nlh, -1), rest;
    for (int q : chosen) {
        if (!is_lane[q]) rest.push_back(q);
        else if (in_lane_pos(q)) lane[q - lo] = q;
    }
    for (int q : chosen) {
        if (!is_lane[q] || in_lane_pos(q)) continue;
        int k = 0;
        while (k < nlh && (lane[k] >= 0 || is_lane[lo + k])) k++;
        if (k == nlh) return false;
        lane[k] = q;
    }
    int pad = lo + nlh;
    for (int k = 0; k < nlh; k++) {
        if (lane[k] >= 0) continue;
        if (!is_chosen[lo + k]) { lane[k] = lo + k; continue; }
        while (pad < nq && (used[pad] || pad > lh_max)) pad++;
        if (pad >= nq) return false;
        lane[k] = pad;
        used[pad] = 1;
    }
    *regs = rest;
    *lanes = lane;
    return true;
}

// Dense gates left in lane positions plus positions moved (each move splits
// the rows into shorter pieces): both cost about the same per unit.
int lane_cost(const std::vector<int> &lanes, const std::vector<int> &dense_cnt, int lo) {
    int c = 0;
    for (size_t k = 0; k < lanes.size(); k++) c += dense_cnt[lanes[k]] + (lanes[k] != lo + (int)k);
    return c;
}

}  // namespace

namespace {

// count_pass on qubit bit masks (n <= 64), for the planner's inner loop: the
// gates of `pending` a pass with resident set `in` can run, stopping once
// every qubit is blocked.
struct MaskCounter {
    std::vector<uint64_t> qmask;   // per gate: its qubits
    const std::vector<char> &mergeable, &anywhere;
    uint64_t all;
    MaskCounter(const std::vector<sq_gate> &gates, const std::vector<char> &m, const std::vector<char> &a, int n)
        : qmask(gates.size()), mergeable(m), anywhere(a), all(n >= 64 ? ~0ULL : (1ULL << n) - 1) {
        for (size_t i = 0; i < gates.size(); i++)
            for (int j = 0; j < gates[i].nq; j++) qmask[i] |= 1ULL << gates[i].q[j];
    }
    int count(const std::vector<int> &pending, uint64_t in, int *dense) const {
        uint64_t blocked = 0;
        int cnt = 0, n_dense = 0;
        for (int gi : pending) {
            const uint64_t m = qmask[gi];
            if (mergeable[gi] && !(m & blocked) && (anywhere[gi] || !(m & ~in))) {
                cnt++;
                n_dense += !anywhere[gi];
                continue;
            }
            blocked |= m;
            if (blocked == all) break;
        }
        *dense = n_dense;
        return cnt;
    }
};

}  // namespace

std::vector<SweepStep> plan_register_passes(const std::vector<sq_gate> &gates, const std::vector<char> &mergeable,
                                            const std::vector<char> &anywhere, int n, int low, int kh, int nlh,
                                            int lh_max, int beam) {
    std::vector<SweepStep> steps;
    std::vector<int> pending(gates.size());
    for (size_t i = 0; i < gates.size(); i++) pending[i] = (int)i;
    std::vector<char> in(n), blocked(n);
    if (nlh < 0 || nlh > low) nlh = 0;
    const int fixed = std::min(low - nlh, n);
    const int lo = low - nlh;
    const MaskCounter counter(gates, mergeable, anywhere, n);
    const uint64_t lane_pos = nlh ? (((1ULL << nlh) - 1) << lo) : 0;
    const uint64_t lane_ok = lh_max >= 63 ? ~0ULL : (lh_max < 0 ? 0 : (2ULL << lh_max) - 1);
    // Whether a chosen set fits kh register slots plus the free lane positions
    // (the test assign_lanes makes, on masks): qubits already in a lane position
    // stay there, and the excess over kh needs free positions and qubits no
    // higher than lh_max.
    auto fits = [&](uint64_t chosen) {
        const int in_lane = __builtin_popcountll(chosen & lane_pos);
        const uint64_t rest = chosen & ~lane_pos;
        const int need = __builtin_popcountll(rest) - kh;
        if (need <= 0) return true;
        return nlh - in_lane >= need && __builtin_popcountll(rest & lane_ok) >= need;
    };
    while (!pending.empty()) {
        const int g0 = pending.front();
        if (!mergeable[g0]) {
            SweepStep st;
            st.gates.push_back(g0);
            steps.push_back(std::move(st));
            pending.erase(pending.begin());
            continue;
        }
        std::fill(in.begin(), in.end(), 0);
        for (int q = 0; q < fixed; q++) in[q] = 1;
        if (nlh == 0) {
            for (int q = fixed; q < std::min(low, n); q++) in[q] = 1;
        }
        uint64_t in_mask = 0;
        for (int q = 0; q < n; q++) if (in[q]) in_mask |= 1ULL << q;
        // dense gates per qubit among the pending ones, for the lane choice
        std::vector<int> dense_cnt(n, 0);
        // Beam search over the chosen qubits: keep the `beam` best sets at each
        // size (most dense gates runnable, then most gates), grow each by every
        // qubit that still fits the register and lane slots.  beam = 1 is the
        // finder's greedy rule.
        struct State {
            int dense = 0, total = 0;
            uint64_t mask = 0;
            std::vector<int> chosen;
        };
        auto better = [](const State &x, const State &y) {
            return x.dense != y.dense ? x.dense > y.dense : x.total > y.total;
        };
        std::vector<State> states(1);
        states[0].total = counter.count(pending, in_mask, &states[0].dense);
        const int picks = kh + nlh;
        std::unordered_set<uint64_t> seen;
        for (int step = 0; step < picks; step++) {
            std::vector<State> next;
            seen.clear();
            for (const State &cur : states) {
                for (int q = (nlh ? lo : low); q < n; q++) {
                    const uint64_t bit = 1ULL << q;
                    if ((in_mask | cur.mask) & bit) continue;
                    const uint64_t m = cur.mask | bit;
                    if (nlh && !fits(m)) continue;
                    if (beam > 1 && !seen.insert(m).second) continue;
                    State ns;
                    ns.mask = m;
                    ns.total = counter.count(pending, in_mask | m, &ns.dense);
                    ns.chosen = cur.chosen;
                    ns.chosen.push_back(q);
                    next.push_back(std::move(ns));
                }
            }
            if (next.empty()) break;
            std::stable_sort(next.begin(), next.end(), better);
            if ((int)next.size() > std::max(1, beam)) next.resize(std::max(1, beam));
            if (!better(next[0], states[0]) && beam <= 1) break;   // greedy: stop when no qubit helps
            states.swap(next);
        }
        // qubits that carry no dense gate of the pass take no slot (diagonal
        // gates run on any qubit)
        std::vector<int> chosen;
        {
            for (int x : states[0].chosen) in[x] = 1;
            std::vector<int> sel;
            count_pass(gates, pending, in, mergeable, anywhere, blocked, &sel);
            std::vector<char> used(n, 0);
            for (int gi : sel)
                if (!anywhere[gi])
                    for (int i = 0; i < gates[gi].nq; i++) used[gates[gi].q[i]] = 1;
            for (int x : states[0].chosen) {
                if (used[x]) chosen.push_back(x);
                else in[x] = 0;
            }
        }
        SweepStep st;
        st.merged = true;
        count_pass(gates, pending, in, mergeable, anywhere, blocked, &st.gates);
        if (st.gates.empty()) {   // cannot happen for a mergeable first gate; stay safe
            st.merged = false;
            st.gates.push_back(g0);
        }
        if (nlh) {
            for (int gi : st.gates)
                if (!anywhere[gi])
                    for (int i = 0; i < gates[gi].nq; i++) dense_cnt[gates[gi].q[i]]++;
            if (!assign_lanes(chosen, dense_cnt, lo, nlh, kh, lh_max, &st.regs, &st.lanes)) {
                // unreachable: every chosen set was checked; fall back to one gate
                st = SweepStep{};
                st.gates.push_back(g0);
            } else {
                std::vector<int> r2, l2;
                if (assign_lanes_min_dense(chosen, dense_cnt, lo, nlh, kh, lh_max, &r2, &l2) &&
                    lane_cost(l2, dense_cnt, lo) < lane_cost(st.lanes, dense_cnt, lo)) {
                    st.regs = r2;
                    st.lanes = l2;
                }
            }
        } else {
            st.regs = chosen;
        }
        std::sort(st.regs.begin(), st.regs.end());
        std::vector<int> rest;
        size_t k = 0;
        for (int gi : pending) {
            if (k < st.gates.size() && st.gates[k] == gi) { k++; continue; }
            rest.push_back(gi);
        }
        pending.swap(rest);
        steps.push_back(std::move(st));
    }
    return steps;
}

std::string plan_subcircuit_text(const std::vector<sq_gate> &gates, int n, int chunk, bool restore) {
    std::vector<int> pos(n), at(n);
    for (int q = 0; q < n; q++) pos[q] = at[q] = q;
    std::string out;
    auto emit_swaps = [&](std::vector<std::pair<int, int>> sw) {
        // sw: (chunk slot, high slot) physical pairs; pair them into VSWAP_2_2
        // where possible (a < b < chunk <= c < d), else VSWAP_1_1.
        std::vector<int> lo, hi;
        for (auto &p : sw) { lo.push_back(p.first); hi.push_back(p.second); }
        std::sort(lo.begin(), lo.end());
        std::sort(hi.begin(), hi.end());
        size_t i = 0;
        for (; i + 1 < lo.size(); i += 2) {
            out += "1\nVSWAP_2_2 " + std::to_string(lo[i]) + " " + std::to_string(lo[i + 1]) + " " +
                   std::to_string(hi[i]) + " " + std::to_string(hi[i + 1]) + "\n";
            for (int j = 0; j < 2; j++) {
                int a = lo[i + j], b = hi[i + j];
                std::swap(at[a], at[b]);
                pos[at[a]] = a; pos[at[b]] = b;
            }
        }
        for (; i < lo.size(); i++) {
            out += "1\nVSWAP_1_1 " + std::to_string(lo[i]) + " " + std::to_string(hi[i]) + "\n";
            int a = lo[i], b = hi[i];
            std::swap(at[a], at[b]);
            pos[at[a]] = a; pos[at[b]] = b;
        }
    };
    auto passes = plan_passes(gates, n, chunk, {});
    for (auto &p : passes) {
        std::vector<char> need(n, 0);
        for (int q : p.qubits) need[q] = 1;
        std::vector<int> free_slots;
        for (int s = 0; s < chunk; s++) if (!need[at[s]]) free_slots.push_back(s);
        std::vector<std::pair<int, int>> sw;
        size_t f = 0;
        for (int q : p.qubits)
            if (pos[q] >= chunk) sw.push_back({free_slots[f++], pos[q]});
        if (!sw.empty()) emit_swaps(sw);
        out += std::to_string(p.gates.size()) + "\n";
        for (int gi : p.gates) {
            sq_gate g = gates[gi];
            for (int i = 0; i < g.nq; i++) g.q[i] = pos[g.q[i]];
            if (g.nq == 2 && g.q[0] > g.q[1]) {
                std::swap(g.q[0], g.q[1]);
                if (g.kind == SQ_GATE_U2) {
                    // reorder the 4x4 basis |first second> -> |second first>
                    static const int perm[4] = {0, 2, 1, 3};
                    double m[32];
                    for (int r = 0; r < 4; r++)
                        for (int c = 0; c < 4; c++)
                            for (int z = 0; z < 2; z++) m[2 * (4 * r + c) + z] = g.mat[2 * (4 * perm[r] + perm[c]) + z];
                    std::memcpy(g.mat, m, sizeof m);
                }
            }
            out += gate_to_line(g) + "\n";
        }
    }
    if (restore) {
        // Return every logical qubit to its own index with pairwise swaps.
        for (int p = 0; p < n; p++) {
            while (at[p] != p) {
                int q = at[p];          // logical qubit sitting at p belongs at q
                int a = std::min(p, q), b = std::max(p, q);
                out += "1\nVSWAP_1_1 " + std::to_string(a) + " " + std::to_string(b) + "\n";
                std::swap(at[a], at[b]);
                pos[at[a]] = a; pos[at[b]] = b;
            }
        }
    }
    return out;
}

}  // namespace sq

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" int sq_plan(const char *circuit_text, int32_t n, int32_t mode, int32_t param, char *buf,
                       uint64_t cap, uint64_t *len) {
    if (!circuit_text || n < 1 || n > 64) return SQ_ERR_ARG;
    uint64_t ng = 0, nb = 0;
    int32_t line = 0;
    int rc = sq_parse_circuit(circuit_text, n, 0, nullptr, 0, &ng, nullptr, 0, &nb, &line);
    if (rc != SQ_OK && rc != SQ_ERR_ARG) return rc;
    std::vector<sq_gate> gates(std::max<uint64_t>(ng, 1));
    rc = sq_parse_circuit(circuit_text, n, 0, gates.data(), ng, &ng, nullptr, 0, &nb, &line);
    if (rc != SQ_OK) return rc;
    gates.resize(ng);
    std::string out;
    if (mode == 0 || mode == 3) {
        // finder role: reference subcircuit text for chunk_qbit = param
        // (mode 3 also restores the qubit order at the end)
        if (param < 4 || param > n) return SQ_ERR_ARG;
        out = sq::plan_subcircuit_text(gates, n, param, mode == 3);
    } else if (mode == 1) {
        // fused blocks: every block touches at most param qubits (no relabelling)
        if (param < 1 || param > 6) return SQ_ERR_ARG;
        for (auto &p : sq::plan_passes(gates, n, param, {})) {
            out += std::to_string(p.gates.size()) + "\n";
            for (int gi : p.gates) out += sq::gate_to_line(gates[gi]) + "\n";
        }
    } else if (mode == 2) {
        // register passes of the complex128 plain sweep: qubits 0-2 plus at most
        // param + 3 chosen qubits (param in registers, up to three in the top
        // lane positions); other gates stand alone
        if (param < 1 || param > 5) return SQ_ERR_ARG;
        std::vector<char> mergeable(gates.size(), 0), anywhere(gates.size(), 0);
        for (size_t i = 0; i < gates.size(); i++) {
            sq::LoweredGate lg;
            if (sq::lower_gate(gates[i], n, lg, nullptr) != SQ_OK) return SQ_ERR_ARG;
            if (lg.cls == sq::OpClass::Diag && lg.k <= 2) mergeable[i] = anywhere[i] = 1;
            else if (lg.cls == sq::OpClass::Dense && lg.k == 1) mergeable[i] = 1;
        }
        for (auto &st : sq::plan_register_passes(gates, mergeable, anywhere, n, 6, param, 3, std::min(n - 1, 27),
                                                 sq::kDefaultPlanBeam)) {
            out += std::to_string(st.gates.size()) + "\n";
            for (int gi : st.gates) out += sq::gate_to_line(gates[gi]) + "\n";
        }
    } else {
        return SQ_ERR_ARG;
    }
    if (len) *len = out.size() + 1;
    if (!buf || cap < out.size() + 1) return SQ_ERR_ARG;
    std::memcpy(buf, out.c_str(), out.size() + 1);
    return SQ_OK;
}
