// sq_blocks.hip -- subcircuit (blocked) sweep: MEM_Runner::run(subcircuits)
// (MEM_Runner.cpp:136-185).
//
// A subcircuit file alternates VSWAP blocks (qubit-swap passes) and gate
// blocks.  The reference applies a gate block tile by tile
// (MEM_Runner.cpp:173-182) and physically moves the data at every VSWAP so
// that the next block's targets fall in its cache-sized chunk.
//
// With the whole state in HBM there is no chunk to fill.  The default sweep
// (sweep_mode 0) therefore leaves the data where it is: it records the
// permutation the VSWAPs make (sigma[p] = the index bit that holds the
// reference's bit p), applies every gate on its mapped targets as the plain
// sweep does (register passes), and moves the data once at the end, with one
// bit-permutation pass (kernels/bitperm.hpp; qubit swaps when no second copy
// of the state fits in HBM), so the final layout is exactly the reference's.
// A plan whose swaps restore the qubit order moves no data at all.
//
// sweep_mode 1 keeps the reference's shape: VSWAP blocks are permutation
// sweeps and each gate block is one LDS tile pass (kernels/tile.hpp) over a
// tile qubit set Q that contains the block's targets and the lowest index
// bits.  A block too wide for one tile is split by the planner into several
// tile passes; the result is the same because the gates are applied in an
// order that respects every per-qubit dependency.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "host/planner.hpp"
#include "kernels/bitperm.hpp"
#include "kernels/tile.hpp"

using sq::LoweredGate;
using sq::OpClass;
using namespace sqk;

namespace {

constexpr int kTileMinK = 10;   // 4 amplitudes per thread with 256 threads

// SQ_TILE_PASSES=0 builds the blocked sweep without the LDS tile pass (gate
// blocks then run as per-gate sweeps, the same map since every gate of a
// block is tile-local).
#ifndef SQ_TILE_PASSES
#define SQ_TILE_PASSES 1
#endif
constexpr bool kTilePassesEnabled = SQ_TILE_PASSES != 0;
constexpr int kTilePrefK = 10;  // 16 KiB (complex128) per workgroup: high occupancy

// Widest tile whose 2^K amplitudes fit the device's per-workgroup LDS limit.
int tile_max_k(const sq_engine *e) {
    int k = e->dtype == SQ_C128 ? 12 : 13;
    while (k > kTileMinK && ((size_t)1 << k) * e->amp_bytes() > e->lds_per_block) k--;
    return k;
}
int low_fixed(int dtype) { return dtype == SQ_C128 ? 6 : 7; }      // one wave's 1 KiB run

struct TilePass {
    bool sweep = false;                 // true: apply gates one by one (VSWAP blocks, tiny states)
    bool plain = false;                 // true: the block as register passes (the plain sweep's form)
    std::vector<int> gates;             // indices into the lowered list
    std::vector<int> q;                 // tile qubits ascending (tile passes)
    int op_begin = 0, n_ops = 0;        // slice of the op array
};

template <typename T>
int launch_tile(sq_engine *e, int K, const TileArgs<T> &a, unsigned grid) {
    cx<T> *s = reinterpret_cast<cx<T> *>(e->d_state);
    const size_t lds = ((size_t)1 << K) * sizeof(cx<T>);
    if (lds > e->lds_per_block) return e->fail(SQ_ERR_ARG, "internal: tile exceeds the LDS limit");
    switch (K) {
    case 10: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile<T, 10, 256>), grid, 256, lds, e->stream, s, a); break;
    case 11: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile<T, 11, 256>), grid, 256, lds, e->stream, s, a); break;
    case 12: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile<T, 12, 256>), grid, 256, lds, e->stream, s, a); break;
    case 13:
        if constexpr (sizeof(T) == 4) { hipLaunchKernelGGL(HIP_KERNEL_NAME(k_tile<T, 13, 256>), grid, 256, lds, e->stream, s, a); break; }
        return e->fail(SQ_ERR_ARG, "internal: tile too wide for complex128");
    default:
        return e->fail(SQ_ERR_ARG, "internal: bad tile width");
    }
    return e->check_launch("k_tile");
}

// Appends one gate as a tile op (local bit positions) and its coefficients.
template <typename T>
void encode_op(const LoweredGate &g, const std::vector<int> &q, std::vector<TileOp> &ops, std::vector<cx<T>> &coef) {
    auto local = [&](int t) { return (int32_t)(std::lower_bound(q.begin(), q.end(), t) - q.begin()); };
    auto push_coef = [&](int cnt) {
        for (int i = 0; i < cnt; i++) coef.push_back({(T)g.m[i].real(), (T)g.m[i].imag()});
    };
    TileOp o{};
    o.coef = (uint32_t)coef.size();
    if (g.cls == OpClass::Perm) {
        for (int p = 0; p < g.n_pairs; p++) {
            TileOp w{};
            int a = local(g.t[2 * p]), b = local(g.t[2 * p + 1]);
            w.kind = TK_SWAP; w.t0 = std::min(a, b); w.t1 = std::max(a, b);
            w.coef = (uint32_t)coef.size();
            ops.push_back(w);
        }
        return;
    }
    o.t0 = local(g.t[0]);
    o.t1 = g.k > 1 ? local(g.t[1]) : 0;
    if (g.cls == OpClass::Dense) {
        o.kind = g.k == 1 ? TK_DENSE1 : TK_DENSE2;
        push_coef(g.k == 1 ? 4 : 16);
    } else {
        o.kind = g.k == 1 ? TK_DIAG1 : TK_DIAG2;
        push_coef(g.k == 1 ? 2 : 4);
    }
    ops.push_back(o);
}

int run_blocks_mapped(sq_engine *e, const sq_gate *gates, const std::vector<LoweredGate> &low,
                      const uint32_t *block_sizes, uint64_t n_blocks) {
    const int n = e->n;
    std::vector<int> sigma(n);
    for (int p = 0; p < n; p++) sigma[p] = p;
    std::vector<sq_gate> flat;
    std::vector<LoweredGate> flat_low;
    size_t first = 0;
    for (uint64_t b = 0; b < n_blocks; b++) {
        for (size_t i = first; i < first + block_sizes[b]; i++) {
            const sq_gate &g = gates[i];
            if (g.kind == SQ_GATE_VSWAP_1_1 || g.kind == SQ_GATE_VSWAP_2_2) {
                // the reference's bit a now holds what its bit b held, and b what a held
                for (int j = 0; j < low[i].n_pairs; j++) std::swap(sigma[low[i].t[2 * j]], sigma[low[i].t[2 * j + 1]]);
                continue;
            }
            LoweredGate lg = low[i];
            sq::relabel(lg, sigma);
            sq_gate mg = g;
            for (int j = 0; j < g.nq; j++) mg.q[j] = sigma[g.q[j]];
            flat.push_back(mg);
            flat_low.push_back(lg);
        }
        first += block_sizes[b];
    }
    int rc = e->apply_plain(flat, flat_low);
    if (rc) return rc;
    // The reference's bit p is at bit sigma[p]: one permutation pass moves
    // every bit home (out of place, when a second copy of the state fits)
    bool moved = false;
    for (int p = 0; p < n; p++) moved = moved || sigma[p] != p;
    if (!moved) return SQ_OK;
    if (e->ensure_alt() == SQ_OK) {
        std::vector<int> perm(n);
        for (int p = 0; p < n; p++) perm[sigma[p]] = p;
        return e->bit_permute(perm);
    }
    // otherwise swap index bits until sigma is the identity (at most n - 1 swaps)
    for (int p = 0; p < n; p++) {
        while (sigma[p] != p) {
            const int x = p, y = sigma[p];
            sq_gate sw{};
            sw.kind = SQ_GATE_SWAP;
            sw.nq = 2;
            sw.q[0] = std::min(x, y);
            sw.q[1] = std::max(x, y);
            LoweredGate lg;
            rc = e->lower(sw, lg);
            if (!rc) rc = e->apply_lowered(lg);
            if (rc) return rc;
            for (int r = 0; r < n; r++) {
                if (sigma[r] == x) sigma[r] = y;
                else if (sigma[r] == y) sigma[r] = x;
            }
        }
    }
    return SQ_OK;
}

template <typename T>
int run_blocks(sq_engine *e, const sq_gate *gates, const std::vector<LoweredGate> &low,
               const uint32_t *block_sizes, uint64_t n_blocks) {
    const int n = e->n;
    const int kmax = std::min(tile_max_k(e), n);
    std::vector<int> fixed;
    for (int i = 0; i < std::min(low_fixed(e->dtype), n); i++) fixed.push_back(i);

    // 1. passes
    std::vector<TilePass> passes;
    size_t first = 0;
    for (uint64_t b = 0; b < n_blocks; b++) {
        const size_t cnt = block_sizes[b];
        if (cnt == 0) continue;
        const int kind0 = gates[first].kind;
        const bool vswap = kind0 == SQ_GATE_VSWAP_1_1 || kind0 == SQ_GATE_VSWAP_2_2;
        // A block of one-qubit dense and diagonal gates runs as register
        // passes: its gates stay in order, each pass reads and writes HBM once,
        // and the gates are applied in registers rather than in an LDS tile
        // (whose per-gate LDS traffic bounds the tile pass).
        bool mergeable = !vswap && e->sweep_mode == 0;
        for (size_t i = 0; mergeable && i < cnt; i++) {
            const LoweredGate &g = low[first + i];
            mergeable = (g.cls == OpClass::Diag && g.k <= 2) || (g.cls == OpClass::Dense && g.k == 1);
        }
        if (mergeable) {
            TilePass p;
            p.plain = true;
            for (size_t i = 0; i < cnt; i++) p.gates.push_back((int)(first + i));
            passes.push_back(std::move(p));
        } else if (vswap || n < kTileMinK || !kTilePassesEnabled) {
            TilePass p;
            p.sweep = true;
            for (size_t i = 0; i < cnt; i++) p.gates.push_back((int)(first + i));
            passes.push_back(std::move(p));
        } else {
            std::vector<char> need(n, 0);
            for (int f : fixed) need[f] = 1;
            for (size_t i = 0; i < cnt; i++)
                for (int j = 0; j < gates[first + i].nq; j++) need[gates[first + i].q[j]] = 1;
            const int req = (int)std::count(need.begin(), need.end(), 1);
            std::vector<std::vector<int>> groups;   // (gate indices, required qubits)
            std::vector<std::vector<int>> qsets;
            if (req <= kmax) {
                std::vector<int> gi, qs;
                for (size_t i = 0; i < cnt; i++) gi.push_back((int)(first + i));
                for (int q = 0; q < n; q++) if (need[q]) qs.push_back(q);
                groups.push_back(gi);
                qsets.push_back(qs);
            } else {
                std::vector<sq_gate> blk(gates + first, gates + first + cnt);
                for (auto &pp : sq::plan_passes(blk, n, kmax, fixed)) {
                    std::vector<int> gi;
                    for (int x : pp.gates) gi.push_back((int)first + x);
                    groups.push_back(gi);
                    qsets.push_back(pp.qubits);
                }
            }
            for (size_t gidx = 0; gidx < groups.size(); gidx++) {
                TilePass p;
                p.gates = groups[gidx];
                std::vector<char> in(n, 0);
                for (int q : qsets[gidx]) in[q] = 1;
                for (int f : fixed) in[f] = 1;
                int size = (int)std::count(in.begin(), in.end(), 1);
                const int want = std::min(std::max(size, kTilePrefK), kmax);
                for (int q = 0; q < n && size < want; q++) if (!in[q]) { in[q] = 1; size++; }
                for (int q = 0; q < n; q++) if (in[q]) p.q.push_back(q);
                passes.push_back(std::move(p));
            }
        }
        first += cnt;
    }

    // 2. tile programs, uploaded once
    std::vector<TileOp> ops;
    std::vector<cx<T>> coef;
    for (auto &p : passes) {
        if (p.sweep || p.plain) continue;
        p.op_begin = (int)ops.size();
        for (int gi : p.gates) encode_op<T>(low[gi], p.q, ops, coef);
        p.n_ops = (int)ops.size() - p.op_begin;
    }
    const size_t ops_bytes = ((ops.size() * sizeof(TileOp) + 255) / 256) * 256;
    const size_t need_bytes = ops_bytes + coef.size() * sizeof(cx<T>) + 256;
    if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
    if (need_bytes > e->blockprog_cap) {
        if (e->d_blockprog) hipFree(e->d_blockprog);
        e->d_blockprog = nullptr;
        e->blockprog_cap = 0;
        hipError_t he = hipMalloc(&e->d_blockprog, need_bytes * 2);
        if (he != hipSuccess) return e->hip_fail(he, "hipMalloc(tile program)");
        e->blockprog_cap = need_bytes * 2;
    }
    char *dprog = static_cast<char *>(e->d_blockprog);
    if (!ops.empty()) {
        hipError_t he = hipMemcpy(dprog, ops.data(), ops.size() * sizeof(TileOp), hipMemcpyHostToDevice);
        if (he == hipSuccess && !coef.empty())
            he = hipMemcpy(dprog + ops_bytes, coef.data(), coef.size() * sizeof(cx<T>), hipMemcpyHostToDevice);
        if (he != hipSuccess) return e->hip_fail(he, "hipMemcpy(tile program)");
    }
    const TileOp *dops = reinterpret_cast<const TileOp *>(dprog);
    const cx<T> *dcoef = reinterpret_cast<const cx<T> *>(dprog + ops_bytes);

    // 3. launches in order
    for (auto &p : passes) {
        if (p.sweep) {
            for (int gi : p.gates) {
                int rc = e->apply_lowered(low[gi]);
                if (rc) return rc;
            }
            continue;
        }
        if (p.plain) {
            std::vector<sq_gate> bg;
            std::vector<LoweredGate> bl;
            for (int gi : p.gates) {
                bg.push_back(gates[gi]);
                bl.push_back(low[gi]);
            }
            int rc = e->apply_plain(bg, bl);
            if (rc) return rc;
            continue;
        }
        TileArgs<T> a{};
        a.ops = dops + p.op_begin;
        a.coef = dcoef;
        a.n_ops = p.n_ops;
        const int K = (int)p.q.size();
        for (int i = 0; i < K; i++) a.q[i] = p.q[i];
        a.n_tiles = 1ULL << (n - K);
        char nm[48];
        std::snprintf(nm, sizeof nm, "tile%d_%s", K, e->dtype == SQ_C128 ? "c128" : "c64");
        int slot = e->stat_begin(nm, 2.0 * e->amp_bytes() * std::ldexp(1.0, n));
        int rc = launch_tile<T>(e, K, a, (unsigned)a.n_tiles);
        e->stat_end(slot);
        if (rc) return rc;
        e->n_tile_passes++;
    }
    return SQ_OK;
}

}  // namespace

extern "C" int sq_apply_blocks(sq_engine *e, const sq_gate *gates, const uint32_t *block_sizes,
                               uint64_t n_blocks) {
    if (!e || (!gates && n_blocks) || (!block_sizes && n_blocks)) return SQ_ERR_ARG;
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    uint64_t total = 0;
    for (uint64_t b = 0; b < n_blocks; b++) total += block_sizes[b];
    std::vector<LoweredGate> low(total);
    for (uint64_t i = 0; i < total; i++) {
        int rc = e->lower(gates[i], low[i]);
        if (rc) { e->err += " (gate " + std::to_string(i) + ")"; return rc; }
    }
    if (e->out_of_core())   // same map as the gates in order
        return e->stream_apply(std::vector<sq_gate>(gates, gates + total), low);
    if (e->sweep_mode == 0) return run_blocks_mapped(e, gates, low, block_sizes, n_blocks);
    return e->dtype == SQ_C128 ? run_blocks<double>(e, gates, low, block_sizes, n_blocks)
                               : run_blocks<float>(e, gates, low, block_sizes, n_blocks);
}

int sq_engine::bit_permute(const std::vector<int> &perm) {
    if ((int)perm.size() != n) return fail(SQ_ERR_ARG, "internal: permutation of the wrong size");
    std::vector<char> seen(n, 0);
    for (int q = 0; q < n; q++) {
        if (perm[q] < 0 || perm[q] >= n || seen[perm[q]]) return fail(SQ_ERR_ARG, "internal: not a permutation");
        seen[perm[q]] = 1;
    }
    int rc = ensure_alt();
    if (rc) return fail(SQ_ERR_OOM, "no room for the permutation pass's second state buffer");
    // source-contiguous low bits Ls = [0, c) and the bits that land on the
    // destination's [0, c); the tile is their union
    int c = std::min(6, n);
    std::vector<char> in_t(n, 0);
    auto tile_bits = [&]() {
        std::fill(in_t.begin(), in_t.end(), 0);
        int tb = 0;
        for (int q = 0; q < n; q++)
            if (q < c || perm[q] < c) { in_t[q] = 1; tb++; }
        return tb;
    };
    int tb = tile_bits();
    while (c > 1 && ((size_t)1 << tb) * amp_bytes() > lds_per_block) { c--; tb = tile_bits(); }
    if (tb > kPermMaxTile || n - tb > kPermMaxBits || ((size_t)1 << tb) * amp_bytes() > lds_per_block)
        return fail(SQ_ERR_ARG, "internal: permutation tile too wide");
    BitPermArgs a{};
    a.tb = tb;
    a.n_other = n - tb;
    std::vector<int> load_pos(n, -1);
    int k = 0;
    for (int q = 0; q < c; q++) { a.ld_src[k] = q; load_pos[q] = k++; }
    for (int q = c; q < n; q++) if (in_t[q]) { a.ld_src[k] = q; load_pos[q] = k++; }
    std::vector<int> inv(n);
    for (int q = 0; q < n; q++) inv[perm[q]] = q;
    k = 0;
    for (int d = 0; d < c; d++) { a.st_loc[k] = load_pos[inv[d]]; a.st_dst[k] = d; k++; }
    for (int q = 0; q < n; q++)
        if (in_t[q] && perm[q] >= c) { a.st_loc[k] = load_pos[q]; a.st_dst[k] = perm[q]; k++; }
    if (k != tb) return fail(SQ_ERR_ARG, "internal: permutation tile bits");
    k = 0;
    for (int q = 0; q < n; q++)
        if (!in_t[q]) { a.other_src[k] = q; a.other_dst[k] = perm[q]; k++; }
    const uint64_t groups = 1ULL << (n - tb);
    if (groups > 0x7fffffffULL) return fail(SQ_ERR_ARG, "internal: permutation grid");
    const size_t lds = ((size_t)1 << tb) * amp_bytes();
    const int slot = stat_begin(dtype == SQ_C128 ? "bitperm_c128" : "bitperm_c64", 2.0 * (double)state_bytes());
    if (dtype == SQ_C128)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bitperm<double>), (unsigned)groups, kBlock, lds, stream,
                           static_cast<const cx<double> *>(d_state), static_cast<cx<double> *>(d_alt), a);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_bitperm<float>), (unsigned)groups, kBlock, lds, stream,
                           static_cast<const cx<float> *>(d_state), static_cast<cx<float> *>(d_alt), a);
    stat_end(slot);
    rc = check_launch("k_bitperm");
    if (rc) return rc;
    std::swap(d_state, d_alt);   // the state now lives in the other buffer
    return SQ_OK;
}

void sq_engine::free_extras() {
    for (void **p : {&d_blockprog, &d_fused, &d_aux, &d_ckpt, &d_gradpart, &d_passprog, &d_sample}) {
        if (*p) hipFree(*p);
        *p = nullptr;
    }
    blockprog_cap = fused_cap = aux_cap = ckpt_cap = gradpart_cap = passprog_cap = sample_cap = 0;
}

