// planner.hpp -- groups a gate sequence into passes whose gates all act inside
// a small qubit set.  This is the role of the reference's subcircuit finder
// (subcircuitFinder/finder.cpp:629-877): it greedily picks the chunk set that
// lets the most pending gates run (findChunkbyMaxNodePrinciple, 651-698) and
// pops them (popNodes, 629-649).  Here the same greedy rule is parameterised:
//   * budget: qubits per pass (chunk_qbit, LDS tile width, or fused k),
//   * fixed:  qubits every pass must contain (the low index bits, so a pass's
//             tile is read from HBM in whole contiguous runs).
// Dependencies are per qubit: a gate may join a pass only if every earlier
// gate sharing a qubit with it has already been scheduled.
#pragma once
#include <string>
#include <vector>

#include "subquokka_hip.h"

namespace sq {

struct PlanPass {
    std::vector<int> qubits;     // sorted; |qubits| <= budget
    std::vector<int> gates;      // indices into the input, in execution order
};

std::vector<int> gate_qubits(const sq_gate &g);

// Greedy pass partition.  fixed must be a subset of [0, n) of size <= budget.
std::vector<PlanPass> plan_passes(const std::vector<sq_gate> &gates, int n, int budget,
                                  const std::vector<int> &fixed);

// Plain-sweep schedule of register passes.  A merged step applies its gates
// in one read and one write of the state with the qubits [0, low) plus at
// most kh register qubits resident; any other step is one gate on its own.
// With nlh > 0 the top nlh of those low positions, [low - nlh, low), may hold
// other qubits instead (at most lh_max): the step then has [0, low - nlh)
// plus up to kh + nlh chosen qubits resident.
//   mergeable[i]: gate i may run inside a register pass;
//   anywhere[i]:  gate i may run whatever the pass's qubit set (diagonal
//                 gates: the phase needs only the amplitude's index).
// Per-qubit order is kept: a gate joins a step only after every earlier
// gate sharing a qubit with it.
struct SweepStep {
    bool merged = false;
    std::vector<int> gates;      // indices into the input, in execution order
    std::vector<int> regs;       // register qubits of a merged step, ascending
    std::vector<int> lanes;      // nlh > 0: qubit in position low - nlh + k (itself: unchanged)
};
// beam: sets kept per size in the search for a step's qubits (1: greedy).
// The engine's default: on the 30-qubit random-layer circuit 16 cuts the
// passes from 25 (greedy) to 19.
constexpr int kDefaultPlanBeam = 16;
std::vector<SweepStep> plan_register_passes(const std::vector<sq_gate> &gates, const std::vector<char> &mergeable,
                                            const std::vector<char> &anywhere, int n, int low, int kh,
                                            int nlh = 0, int lh_max = -1, int beam = 1);

// Reference subcircuit text (finder.cpp:69-333 format) for a chunk of
// chunk_qbit low qubits: gate blocks with relabelled targets separated by
// VSWAP_1_1 / VSWAP_2_2 passes.  restore=true appends swaps that return every
// logical qubit to its own position, which the reference finder leaves out
// (finder.cpp:304-330 is commented out).
std::string plan_subcircuit_text(const std::vector<sq_gate> &gates, int n, int chunk_qbit, bool restore);

// Gate line in the MEM grammar with %.17g reals (parses back to the same doubles).
std::string gate_to_line(const sq_gate &g);

}  // namespace sq
