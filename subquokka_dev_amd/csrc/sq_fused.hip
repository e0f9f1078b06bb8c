// sq_fused.hip -- dense five-qubit blocks: FP64 composition on the GPU, then
// application on MFMA (kernels/fused.hpp).  Entry points:
//   sq_run_fused      gates grouped in blocks -> compose (FP64) -> apply
//   sq_apply_fused    caller-supplied FP64 matrices -> apply
//   sq_compose_block  one block's FP64 matrix back to the host
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "host/planner.hpp"
#include "kernels/fpass.hpp"
#include "kernels/fused.hpp"

using sq::LoweredGate;
using sq::OpClass;
using namespace sqk;

// 16-column batches per wave and loop iteration (loads in flight), per
// element type: complex<float> tiles are half the registers of complex<double>.
#ifndef SQ_FUSED_U_F32
#define SQ_FUSED_U_F32 2
#endif
#ifndef SQ_FUSED_U_F64
#define SQ_FUSED_U_F64 2
#endif

namespace {

struct FusedProgram {
    std::vector<ComposeOp> ops;
    std::vector<cx<double>> coef;
    std::vector<uint32_t> op_begin;            // n_blocks + 1
    std::vector<std::vector<int>> qsets;       // per block, kFusedK qubits ascending
};

// Appends the gate as compose ops on the local bits of q (ascending).
void encode_compose(const LoweredGate &g, const std::vector<int> &q, FusedProgram &p) {
    auto local = [&](int t) { return (int32_t)(std::lower_bound(q.begin(), q.end(), t) - q.begin()); };
    auto push = [&](const std::complex<double> *m, int cnt) {
        for (int i = 0; i < cnt; i++) p.coef.push_back({m[i].real(), m[i].imag()});
    };
    if (g.cls == OpClass::Perm) {
        for (int pr = 0; pr < g.n_pairs; pr++) {
            ComposeOp o{};
            int a = local(g.t[2 * pr]), b = local(g.t[2 * pr + 1]);
            o.kind = 2; o.t0 = std::min(a, b); o.t1 = std::max(a, b);
            o.coef = (uint32_t)p.coef.size();
            p.ops.push_back(o);
        }
        return;
    }
    ComposeOp o{};
    o.coef = (uint32_t)p.coef.size();
    o.t0 = local(g.t[0]);
    o.t1 = g.k > 1 ? local(g.t[1]) : 0;
    o.kind = g.k == 1 ? 0 : 1;
    if (g.cls == OpClass::Dense) {
        push(g.m, g.k == 1 ? 4 : 16);
    } else {   // diagonal -> dense matrix
        const int d = 1 << g.k;
        std::complex<double> m[16] = {};
        for (int i = 0; i < d; i++) m[i * d + i] = g.m[i];
        push(m, d * d);
    }
    p.ops.push_back(o);
}

// Qubit set of a block padded to kFusedK qubits (lowest unused qubits).
std::vector<int> padded_set(const std::vector<int> &used, int n) {
    std::vector<char> in(n, 0);
    for (int q : used) in[q] = 1;
    int size = (int)used.size();
    for (int q = 0; q < n && size < kFusedK; q++) if (!in[q]) { in[q] = 1; size++; }
    std::vector<int> out;
    for (int q = 0; q < n; q++) if (in[q]) out.push_back(q);
    return out;
}

int ensure_fused_buffer(sq_engine *e, size_t bytes) {
    if (bytes <= e->fused_cap) return SQ_OK;
    if (e->d_fused) hipFree(e->d_fused);
    e->d_fused = nullptr;
    e->fused_cap = 0;
    hipError_t he = hipMalloc(&e->d_fused, bytes);
    if (he != hipSuccess) return e->hip_fail(he, "hipMalloc(fused program)");
    e->fused_cap = bytes;
    return SQ_OK;
}

// Applies block b's 32x32 matrix (device, FP64) to the state.
template <typename T>
int launch_apply(sq_engine *e, const cx<double> *u_dev, const std::vector<int> &q) {
    cx<T> *s = reinterpret_cast<cx<T> *>(e->d_state);
    const int n = e->n;
    const double bytes = 2.0 * e->amp_bytes() * std::ldexp(1.0, n);
    char nm[48];
    std::snprintf(nm, sizeof nm, "fused5_%s", e->dtype == SQ_C128 ? "c128" : "c64");
    if (n < 9) {
        FusedSmallArgs a{};
        a.u = u_dev;
        a.n_cols = 1ULL << (n - kFusedK);
        for (int i = 0; i < kFusedK; i++) a.q[i] = q[i];
        int slot = e->stat_begin("fused5_small", bytes);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fused5_small<T>), e->grid_flat(a.n_cols, 256), 256, 0, e->stream, s, a);
        e->stat_end(slot);
        return e->check_launch("k_fused5_small");
    }
    FusedArgs a{};
    a.u = u_dev;
    a.n_batches = 1ULL << (n - kFusedK - 4);
    for (int i = 0; i < kFusedK; i++) a.q[i] = q[i];
    constexpr int U = sizeof(T) == 4 ? SQ_FUSED_U_F32 : SQ_FUSED_U_F64;
    const unsigned grid = e->grid_flat(a.n_batches, 4 * U);
    int slot = e->stat_begin(nm, bytes);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fused5<T, U>), grid, 256, 0, e->stream, s, a);
    e->stat_end(slot);
    return e->check_launch("k_fused5");
}

// Uploads the program, composes every block in FP64 (one workgroup each) into
// the matrix area of the fused buffer.  Returns the device matrix array, and
// in *extra_dev a 256-byte aligned area of extra bytes after it.
int compose_on_device(sq_engine *e, const FusedProgram &p, int k, cx<double> **u_dev, size_t extra = 0,
                      char **extra_dev = nullptr) {
    const size_t nb = p.qsets.size();
    const size_t ops_b = ((p.ops.size() * sizeof(ComposeOp) + 255) / 256) * 256;
    const size_t coef_b = ((p.coef.size() * sizeof(cx<double>) + 255) / 256) * 256;
    const size_t beg_b = ((p.op_begin.size() * sizeof(uint32_t) + 255) / 256) * 256;
    const size_t mat_b = nb * 1024 * sizeof(cx<double>);
    if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
    int rc = ensure_fused_buffer(e, ops_b + coef_b + beg_b + mat_b + extra + 256);
    if (rc) return rc;
    char *base = static_cast<char *>(e->d_fused);
    hipError_t he = hipSuccess;
    if (!p.ops.empty()) he = hipMemcpy(base, p.ops.data(), p.ops.size() * sizeof(ComposeOp), hipMemcpyHostToDevice);
    if (he == hipSuccess && !p.coef.empty())
        he = hipMemcpy(base + ops_b, p.coef.data(), p.coef.size() * sizeof(cx<double>), hipMemcpyHostToDevice);
    if (he == hipSuccess)
        he = hipMemcpy(base + ops_b + coef_b, p.op_begin.data(), p.op_begin.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
    if (he != hipSuccess) return e->hip_fail(he, "hipMemcpy(fused program)");
    ComposeArgs a{};
    a.ops = reinterpret_cast<const ComposeOp *>(base);
    a.coef = reinterpret_cast<const cx<double> *>(base + ops_b);
    a.op_begin = reinterpret_cast<const uint32_t *>(base + ops_b + coef_b);
    a.u_out = reinterpret_cast<cx<double> *>(base + ops_b + coef_b + beg_b);
    a.k = k;
    int slot = e->stat_begin("compose_f64", 0.0);
    hipLaunchKernelGGL(k_compose, (unsigned)nb, 64, 0, e->stream, a);
    e->stat_end(slot);
    rc = e->check_launch("k_compose");
    if (rc) return rc;
    *u_dev = a.u_out;
    if (extra_dev) *extra_dev = base + ops_b + coef_b + beg_b + mat_b;
    return SQ_OK;
}

int build_program(sq_engine *e, const sq_gate *gates, const uint32_t *sizes, uint64_t n_blocks, int k,
                  const int32_t *fixed_q, FusedProgram &p) {
    size_t first = 0;
    p.op_begin.push_back(0);
    for (uint64_t b = 0; b < n_blocks; b++) {
        std::vector<LoweredGate> low(sizes[b]);
        std::vector<int> used;
        for (uint32_t i = 0; i < sizes[b]; i++) {
            int rc = e->lower(gates[first + i], low[i]);
            if (rc) { e->err += " (gate " + std::to_string(first + i) + ")"; return rc; }
            for (int j = 0; j < gates[first + i].nq; j++) used.push_back(gates[first + i].q[j]);
        }
        std::sort(used.begin(), used.end());
        used.erase(std::unique(used.begin(), used.end()), used.end());
        std::vector<int> q;
        if (fixed_q) {
            q.assign(fixed_q, fixed_q + k);
            std::vector<int> qs = q;
            std::sort(qs.begin(), qs.end());
            for (int u : used)
                if (!std::binary_search(qs.begin(), qs.end(), u))
                    return e->fail(SQ_ERR_ARG, "gate target outside the block's qubits");
            q = qs;
        } else {
            if ((int)used.size() > k) return e->fail(SQ_ERR_ARG, "block " + std::to_string(b) + " acts on more than 5 qubits");
            q = padded_set(used, e->n);
        }
        for (auto &lg : low) encode_compose(lg, q, p);
        p.op_begin.push_back((uint32_t)p.ops.size());
        p.qsets.push_back(q);
        first += sizes[b];
    }
    return SQ_OK;
}

// ---------------------------------------------------------------------------
// Fused passes (complex64): the plain gate list cut into passes over tiles of
// r qubits, each pass's gates grouped into four-qubit blocks (kernels/fpass.hpp).
// ---------------------------------------------------------------------------

// Value of the tile-fixed qubit set `ext` (bit i = ext[i]) in variant v.
inline int ext_bit(const std::vector<int> &ext, int q, uint32_t v) {
    for (size_t i = 0; i < ext.size(); i++) if (ext[i] == q) return (int)((v >> i) & 1u);
    return -1;
}

}  // namespace

// One pass of the plan: its tile qubits and its blocks.  A block acts on four
// tile qubits q (ascending); its diagonal gates may also read the tile-fixed
// qubits ext, one composed variant per value of them.
struct FusedPassBlock {
    std::vector<int> q, ext, gates;
};
struct FusedPassStep {
    std::vector<int> tile;      // r qubits ascending: tile bit i <-> tile[i]
    std::vector<FusedPassBlock> blocks;
};
struct FusedPassPlan {
    int r = 0;
    std::vector<FusedPassStep> passes;
    size_t n_var = 0;           // composed variants over all blocks
    size_t n_blocks = 0;
};

namespace {

// Splits the gate list into passes (the register-pass planner with diagonal
// gates allowed anywhere, r - 4 qubits chosen beside qubits 0-3) and each
// pass's gates into blocks of at most four tile qubits (the finder's greedy
// rule, plan_passes with budget 4).  A diagonal gate with qubits outside the
// tile counts as a gate on its qubits inside (none: a scalar of the tile,
// attached to the pass's first block).
int plan_fused_passes(sq_engine *e, const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low, int r,
                      FusedPassPlan &plan) {
    const int n = e->n;
    const size_t ng = gates.size();
    std::vector<char> mergeable(ng, 1), anywhere(ng, 0);
    for (size_t i = 0; i < ng; i++) anywhere[i] = low[i].cls == OpClass::Diag;
    const auto steps = sq::plan_register_passes(gates, mergeable, anywhere, n, 4, r - 4, 0, -1, e->plan_beam);
    plan = FusedPassPlan{};
    plan.r = r;
    for (const auto &st : steps) {
        FusedPassStep ps;
        std::vector<char> in(n, 0);
        for (int q = 0; q < 4; q++) in[q] = 1;
        for (int q : st.regs) in[q] = 1;
        int size = 0;
        for (int q = 0; q < n; q++) size += in[q];
        for (int q = 0; q < n && size < r; q++) if (!in[q]) { in[q] = 1; size++; }
        for (int q = 0; q < n; q++) if (in[q]) ps.tile.push_back(q);
        // proxies: the gates on their tile qubits only
        std::vector<sq_gate> proxy;
        std::vector<int> proxy_gate, scalars;
        for (int gi : st.gates) {
            const sq_gate &g = gates[gi];
            std::vector<int> inside;
            for (int j = 0; j < g.nq; j++) if (in[g.q[j]]) inside.push_back(g.q[j]);
            if ((int)inside.size() == g.nq) {
                proxy.push_back(g);
            } else if (low[gi].cls != OpClass::Diag) {
                return e->fail(SQ_ERR_ARG, "fused pass: a non-diagonal gate outside the tile");
            } else if (inside.empty()) {
                scalars.push_back(gi);
                continue;
            } else {
                sq_gate pg = g;
                pg.kind = SQ_GATE_P;
                pg.nq = 1;
                pg.q[0] = inside[0];
                proxy.push_back(pg);
            }
            proxy_gate.push_back(gi);
        }
        auto ext_of = [&](int gi, std::vector<int> &ext) {
            for (int j = 0; j < gates[gi].nq; j++) {
                const int q = gates[gi].q[j];
                if (!in[q] && std::find(ext.begin(), ext.end(), q) == ext.end()) ext.push_back(q);
            }
        };
        auto add_block = [&](const std::vector<int> &gl) {
            // consecutive pieces of at most kFpMaxExt tile-fixed qubits each
            size_t i = 0;
            while (i < gl.size()) {
                FusedPassBlock b;
                std::vector<int> used;
                for (; i < gl.size(); i++) {
                    std::vector<int> ext = b.ext;
                    ext_of(gl[i], ext);
                    if ((int)ext.size() > kFpMaxExt && !b.gates.empty()) break;
                    b.ext = ext;
                    b.gates.push_back(gl[i]);
                    for (int j = 0; j < gates[gl[i]].nq; j++) if (in[gates[gl[i]].q[j]]) used.push_back(gates[gl[i]].q[j]);
                }
                std::sort(used.begin(), used.end());
                used.erase(std::unique(used.begin(), used.end()), used.end());
                if (used.size() > 4) return false;
                // padded with the lowest other qubits: among 0-3, always tile qubits
                b.q = used;
                for (int q = 0; b.q.size() < 4; q++)
                    if (std::find(used.begin(), used.end(), q) == used.end()) b.q.push_back(q);
                std::sort(b.q.begin(), b.q.end());
                ps.blocks.push_back(std::move(b));
            }
            return true;
        };
        if (!proxy.empty()) {
            for (const auto &blk : sq::plan_passes(proxy, n, 4, {})) {
                std::vector<int> gl;
                for (int pi : blk.gates) gl.push_back(proxy_gate[pi]);
                if (!add_block(gl)) return e->fail(SQ_ERR_ARG, "fused pass: block over four qubits");
            }
        }
        if (!scalars.empty()) {
            // tile scalars commute with every gate of the pass: first block
            FusedPassBlock &b0 = ps.blocks.empty() ? (ps.blocks.emplace_back(), ps.blocks.back()) : ps.blocks.front();
            if (b0.q.empty()) b0.q = {0, 1, 2, 3};
            std::vector<int> rest;
            for (int gi : scalars) {
                std::vector<int> ext = b0.ext;
                ext_of(gi, ext);
                if ((int)ext.size() > kFpMaxExt) { rest.push_back(gi); continue; }
                b0.ext = ext;
                b0.gates.insert(b0.gates.begin(), gi);
            }
            if (!rest.empty() && !add_block(rest)) return e->fail(SQ_ERR_ARG, "fused pass: scalar block");
        }
        for (const auto &b : ps.blocks) plan.n_var += (size_t)1 << b.ext.size();
        plan.n_blocks += ps.blocks.size();
        plan.passes.push_back(std::move(ps));
    }
    return SQ_OK;
}

// LDS offsets of a block (kernels/fpass.hpp): the block's tile bits p, four
// column bits and the remaining batch bits.  Column bits are taken lowest
// first among the non-block tile bits whose bank-pair vectors (s(1 << q) & 31)
// keep the column span independent and free of the vectors of block bits 1
// and 2; failing that, independent only; failing that, any.
void fill_offsets(const std::vector<int> &tile, const std::vector<int> &q, int r, FpBlock &rec) {
    int p[4];
    for (int j = 0; j < 4; j++) p[j] = (int)(std::lower_bound(tile.begin(), tile.end(), q[j]) - tile.begin());
    std::vector<int> nb;
    for (int i = 0; i < r; i++) if (i != p[0] && i != p[1] && i != p[2] && i != p[3]) nb.push_back(i);
    auto vec = [](int i) { return fp_slot(1u << i) & 31u; };
    // GF(2) span of a vector set kept as a basis with distinct leading bits
    auto reduce = [](uint32_t v, const std::vector<uint32_t> &basis) {
        for (uint32_t b : basis) if ((v ^ b) < v) v ^= b;
        return v;
    };
    auto add = [&](std::vector<uint32_t> &basis, uint32_t v) {
        basis.push_back(reduce(v, basis));
        std::sort(basis.begin(), basis.end(), [](uint32_t x, uint32_t y) { return x > y; });
    };
    std::vector<int> col;
    for (int strict = 2; strict >= 0 && col.size() < 4; strict--) {
        col.clear();
        std::vector<uint32_t> basis;
        for (int i : nb) {
            if (col.size() == 4) break;
            const uint32_t v = reduce(vec(i), basis);
            if (strict >= 1 && !v) continue;
            if (strict == 2) {
                std::vector<uint32_t> b2 = basis;
                add(b2, vec(i));
                if (!reduce(vec(p[1]), b2) || !reduce(vec(p[2]), b2)) continue;
            }
            col.push_back(i);
            if (v) add(basis, vec(i));
        }
    }
    std::vector<int> rest;
    for (int i : nb) if (std::find(col.begin(), col.end(), i) == col.end()) rest.push_back(i);
    auto deposit = [](uint32_t x, const std::vector<int> &pos) {
        uint32_t o = 0;
        for (size_t k = 0; k < pos.size(); k++) o |= ((x >> k) & 1u) << pos[k];
        return o;
    };
    const std::vector<int> pv(p, p + 4);
    for (uint32_t i = 0; i < 16; i++) {
        rec.colw[i] = fp_slot(deposit(i, col)) * 8u;
        rec.blkw[i] = fp_slot(deposit(i, pv)) * 8u;
    }
    const uint32_t nw = (uint32_t)fp_waves(r);
    const uint32_t batches = 1u << (r - 8), per_wave = (batches + nw - 1) / nw;
    for (uint32_t w = 0; w < nw; w++)
        for (uint32_t i = 0; i < per_wave; i++)
            if (w + nw * i < batches) rec.ubw[w * per_wave + i] = fp_slot(deposit(w + nw * i, rest)) * 8u;
}

// Compose ops of one variant of a block (ext bits = v).
void encode_variant(const FusedPassBlock &b, const std::vector<LoweredGate> &low, uint32_t v, FusedProgram &p) {
    for (int gi : b.gates) {
        const LoweredGate &lg = low[gi];
        int n_ext = 0;
        for (int j = 0; j < lg.k; j++) n_ext += ext_bit(b.ext, lg.t[j], v) >= 0;
        if (lg.cls != OpClass::Diag || n_ext == 0) {
            encode_compose(lg, b.q, p);
            continue;
        }
        // the diagonal restricted to the tile: entries with the fixed bits set
        LoweredGate d{};
        d.cls = OpClass::Diag;
        d.k = 1;
        if (lg.k == 1 || n_ext == lg.k) {   // scalar of the tile
            int idx = 0;
            for (int j = 0; j < lg.k; j++) idx |= ext_bit(b.ext, lg.t[j], v) << j;
            d.t[0] = b.q[0];
            d.m[0] = d.m[1] = lg.m[idx];
        } else {                              // two targets, one inside
            const int x0 = ext_bit(b.ext, lg.t[0], v), x1 = ext_bit(b.ext, lg.t[1], v);
            if (x0 >= 0) {      // t[0] fixed, t[1] inside
                d.t[0] = lg.t[1];
                d.m[0] = lg.m[x0];
                d.m[1] = lg.m[x0 + 2];
            } else {            // t[1] fixed, t[0] inside
                d.t[0] = lg.t[0];
                d.m[0] = lg.m[2 * x1];
                d.m[1] = lg.m[2 * x1 + 1];
            }
        }
        encode_compose(d, b.q, p);
    }
}

}  // namespace

extern "C" int sq_run_fused(sq_engine *e, const sq_gate *gates, const uint32_t *block_sizes, uint64_t n_blocks) {
    if (!e || (n_blocks && (!gates || !block_sizes))) return SQ_ERR_ARG;
    if (n_blocks == 0) return SQ_OK;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "fused blocks need an HBM-resident state");
    if (e->n < kFusedK) return e->fail(SQ_ERR_ARG, "fused blocks need at least 5 qubits");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    FusedProgram p;
    int rc = build_program(e, gates, block_sizes, n_blocks, kFusedK, nullptr, p);
    if (rc) return rc;
    cx<double> *u = nullptr;
    rc = compose_on_device(e, p, kFusedK, &u);
    if (rc) return rc;
    for (size_t b = 0; b < p.qsets.size(); b++) {
        rc = e->dtype == SQ_C128 ? launch_apply<double>(e, u + b * 1024, p.qsets[b])
                                 : launch_apply<float>(e, u + b * 1024, p.qsets[b]);
        if (rc) return rc;
    }
    return SQ_OK;
}

extern "C" int sq_compose_block(sq_engine *e, const sq_gate *gates, uint64_t n, int32_t k, const int32_t *q,
                                double *u_out) {
    if (!e || !q || !u_out || k < 1 || k > kFusedK || (n && !gates)) return SQ_ERR_ARG;
    for (int i = 0; i < k; i++)
        if (q[i] < 0 || q[i] >= e->n) return e->fail(SQ_ERR_QUBIT_INDEX, "block qubit out of range");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    FusedProgram p;
    uint32_t sz = (uint32_t)n;
    int rc = build_program(e, gates, &sz, 1, k, q, p);
    if (rc) return rc;
    cx<double> *u = nullptr;
    rc = compose_on_device(e, p, k, &u);
    if (rc) return rc;
    // compose stores row-major in the block's sorted-qubit basis; report it in
    // the caller's q order (local bit j <-> q[j])
    const int d = 1 << k;
    std::vector<cx<double>> h(d * d);
    hipError_t he = hipMemcpyAsync(h.data(), u, sizeof(cx<double>) * d * d, hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "compose readback");
    std::vector<int> sorted(q, q + k);
    std::sort(sorted.begin(), sorted.end());
    std::vector<int> pos(k);   // caller bit j -> sorted bit
    for (int j = 0; j < k; j++) pos[j] = (int)(std::lower_bound(sorted.begin(), sorted.end(), q[j]) - sorted.begin());
    auto to_sorted = [&](int idx) {
        int o = 0;
        for (int j = 0; j < k; j++) if ((idx >> j) & 1) o |= 1 << pos[j];
        return o;
    };
    for (int r = 0; r < d; r++)
        for (int c = 0; c < d; c++) {
            const cx<double> z = h[to_sorted(r) * d + to_sorted(c)];
            u_out[2 * (r * d + c)] = z.re;
            u_out[2 * (r * d + c) + 1] = z.im;
        }
    return SQ_OK;
}

extern "C" int sq_apply_fused(sq_engine *e, const sq_fused *blocks, uint64_t n) {
    if (!e || (n && !blocks)) return SQ_ERR_ARG;
    if (n == 0) return SQ_OK;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "fused blocks need an HBM-resident state");
    if (e->n < kFusedK) return e->fail(SQ_ERR_ARG, "fused blocks need at least 5 qubits");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    // Embed each k-qubit matrix into the 32x32 space of its padded, sorted set.
    std::vector<cx<double>> mats(n * 1024);
    std::vector<std::vector<int>> qsets;
    for (uint64_t b = 0; b < n; b++) {
        const sq_fused &f = blocks[b];
        if (f.k < 1 || f.k > kFusedK || !f.u) return e->fail(SQ_ERR_ARG, "bad fused block");
        std::vector<int> used(f.q, f.q + f.k);
        for (int x : used) if (x < 0 || x >= e->n) return e->fail(SQ_ERR_QUBIT_INDEX, "fused qubit out of range");
        std::vector<int> su = used;
        std::sort(su.begin(), su.end());
        if (std::unique(su.begin(), su.end()) != su.end()) return e->fail(SQ_ERR_CONTRACT, "repeated fused qubit");
        std::vector<int> q = padded_set(used, e->n);
        std::vector<int> pos(f.k);   // block bit j -> bit in q
        for (int j = 0; j < f.k; j++) pos[j] = (int)(std::lower_bound(q.begin(), q.end(), f.q[j]) - q.begin());
        int other_mask = 31;
        for (int j = 0; j < f.k; j++) other_mask &= ~(1 << pos[j]);
        const int d = 1 << f.k;
        auto place = [&](int idx) { int o = 0; for (int j = 0; j < f.k; j++) if ((idx >> j) & 1) o |= 1 << pos[j]; return o; };
        cx<double> *m = mats.data() + b * 1024;
        for (int r = 0; r < d; r++)
            for (int c = 0; c < d; c++) {
                const cx<double> z = {f.u[2 * (r * d + c)], f.u[2 * (r * d + c) + 1]};
                for (int rest = 0; rest < 32; rest++) {
                    if (rest & ~other_mask) continue;
                    m[(place(r) | rest) * 32 + (place(c) | rest)] = z;
                }
            }
        qsets.push_back(q);
    }
    if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
    int rc = ensure_fused_buffer(e, mats.size() * sizeof(cx<double>) + 256);
    if (rc) return rc;
    hipError_t he = hipMemcpy(e->d_fused, mats.data(), mats.size() * sizeof(cx<double>), hipMemcpyHostToDevice);
    if (he != hipSuccess) return e->hip_fail(he, "hipMemcpy(fused matrices)");
    const cx<double> *u = static_cast<const cx<double> *>(e->d_fused);
    for (uint64_t b = 0; b < n; b++) {
        rc = e->dtype == SQ_C128 ? launch_apply<double>(e, u + b * 1024, qsets[b])
                                 : launch_apply<float>(e, u + b * 1024, qsets[b]);
        if (rc) return rc;
    }
    return SQ_OK;
}

// ---------------------------------------------------------------------------
// sq_apply_gates_fused
// ---------------------------------------------------------------------------
namespace {

template <int R>
void launch_fpass(sq_engine *e, const FpArgs &a, unsigned grid) {
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fpass<R>), grid, 64 * fp_waves(R), 0, e->stream,
                       reinterpret_cast<cx<float> *>(e->d_state), a);
}

}  // namespace

extern "C" int sq_apply_gates_fused(sq_engine *e, const sq_gate *gates, uint64_t n_gates, int32_t tile_qbit) {
    if (!e || (n_gates && !gates)) return SQ_ERR_ARG;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "fused passes need an HBM-resident state");
    if (e->dtype != SQ_C64) return e->fail(SQ_ERR_UNSUPPORTED, "fused passes run on complex64 states");
    if (e->n < kFpMinR) return e->fail(SQ_ERR_ARG, "fused passes need at least 9 qubits");
    if (tile_qbit != 0 && (tile_qbit < kFpMinR || tile_qbit > kFpMaxR))
        return e->fail(SQ_ERR_ARG, "tile_qbit must be 0 or 9..13");
    if (e->n - std::min(e->n, kFpMaxR) > kFpMaxOther) return e->fail(SQ_ERR_ARG, "too many qubits");
    if (n_gates == 0) return SQ_OK;
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    const int r = std::min(e->n, tile_qbit ? tile_qbit : 13);
    std::vector<sq_gate> gv(gates, gates + n_gates);
    std::vector<LoweredGate> low(n_gates);
    for (uint64_t i = 0; i < n_gates; i++) {
        int rc = e->lower(gv[i], low[i]);
        if (rc) { e->err += " (gate " + std::to_string(i) + ")"; return rc; }
    }
    // plan: cached by structure (qubits and gate classes), as the plain sweep's
    std::vector<int32_t> key = {e->n, r, e->plan_beam, (int32_t)n_gates};
    for (uint64_t i = 0; i < n_gates; i++) {
        key.push_back((int32_t)low[i].cls | (gv[i].nq << 2));
        for (int j = 0; j < gv[i].nq; j++) key.push_back(gv[i].q[j]);
    }
    std::shared_ptr<FusedPassPlan> plan;
    for (size_t i = 0; i < e->fpass_cache.size(); i++) {
        if (e->fpass_cache[i].first == key) {
            plan = e->fpass_cache[i].second;
            auto ent = std::move(e->fpass_cache[i]);
            e->fpass_cache.erase(e->fpass_cache.begin() + (long)i);
            e->fpass_cache.insert(e->fpass_cache.begin(), std::move(ent));
            break;
        }
    }
    if (!plan) {
        plan = std::make_shared<FusedPassPlan>();
        int rc = plan_fused_passes(e, gv, low, r, *plan);
        if (rc) return rc;
        e->fpass_cache.insert(e->fpass_cache.begin(), {key, plan});
        if (e->fpass_cache.size() > 4) e->fpass_cache.pop_back();
    }
    // compose programs of every variant, the block records
    FusedProgram p;
    p.op_begin.push_back(0);
    std::vector<FpBlock> recs;
    uint32_t a_off = 0;
    for (const auto &ps : plan->passes) {
        for (const auto &b : ps.blocks) {
            FpBlock rec{};
            rec.a_off = a_off;
            rec.n_ext = (int32_t)b.ext.size();
            for (size_t i = 0; i < b.ext.size(); i++) rec.ext_q[i] = b.ext[i];
            fill_offsets(ps.tile, b.q, r, rec);
            recs.push_back(rec);
            const uint32_t nv = 1u << b.ext.size();
            for (uint32_t v = 0; v < nv; v++) {
                encode_variant(b, low, v, p);
                p.op_begin.push_back((uint32_t)p.ops.size());
                p.qsets.push_back(b.q);
            }
            a_off += nv * kFpAmatU4;
        }
    }
    const size_t n_var = p.qsets.size();
    const size_t amat_b = ((n_var * kFpAmatU4 * sizeof(u32x4) + 255) / 256) * 256;
    const size_t rec_b = recs.size() * sizeof(FpBlock);
    cx<double> *u = nullptr;
    char *extra = nullptr;
    int rc = compose_on_device(e, p, 4, &u, amat_b + rec_b, &extra);
    if (rc) return rc;
    u32x4 *amat = reinterpret_cast<u32x4 *>(extra);
    FpBlock *d_recs = reinterpret_cast<FpBlock *>(extra + amat_b);
    hipError_t he = hipMemcpy(d_recs, recs.data(), rec_b, hipMemcpyHostToDevice);
    if (he != hipSuccess) return e->hip_fail(he, "hipMemcpy(fused pass blocks)");
    hipLaunchKernelGGL(k_fpass_amat, (unsigned)((n_var * kFpAmatU4 + 255) / 256), 256, 0, e->stream, u, amat,
                       (uint32_t)n_var);
    rc = e->check_launch("k_fpass_amat");
    if (rc) return rc;
    const int n = e->n;
    const unsigned grid = 1u << (n - r);
    const double bytes = 2.0 * 8.0 * std::ldexp(1.0, n);
    int b_begin = 0;
    for (const auto &ps : plan->passes) {
        FpArgs a{};
        a.blocks = d_recs;
        a.amat = amat;
        a.b_begin = b_begin;
        a.b_end = b_begin + (int)ps.blocks.size();
        b_begin = a.b_end;
        a.n_other = n - r;
        for (int i = 0; i < r; i++) a.tile_q[i] = ps.tile[i];
        for (int q = 0, k = 0; q < n; q++)
            if (!std::binary_search(ps.tile.begin(), ps.tile.end(), q)) a.other_q[k++] = q;
        // algorithmic flops: the complex 16x16 product, 128 real flops per
        // amplitude and block (the split products execute three times that)
        int slot = e->stat_begin("fpass_c64", bytes, 128.0 * (double)ps.blocks.size() * std::ldexp(1.0, n));
        switch (r) {
        case 9: launch_fpass<9>(e, a, grid); break;
        case 10: launch_fpass<10>(e, a, grid); break;
        case 11: launch_fpass<11>(e, a, grid); break;
        case 12: launch_fpass<12>(e, a, grid); break;
        default: launch_fpass<13>(e, a, grid); break;
        }
        e->stat_end(slot);
        rc = e->check_launch("k_fpass");
        if (rc) return rc;
    }
    return SQ_OK;
}
