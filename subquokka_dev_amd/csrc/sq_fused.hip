// sq_fused.hip -- dense five-qubit blocks: FP64 composition on the GPU, then
// application on MFMA (kernels/fused.hpp).  Entry points:
//   sq_run_fused      gates grouped in blocks -> compose (FP64) -> apply
//   sq_apply_fused    caller-supplied FP64 matrices -> apply
//   sq_compose_block  one block's FP64 matrix back to the host
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstring>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
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
// the matrix area of the fused buffer.  Returns the device matrix array.
int compose_on_device(sq_engine *e, const FusedProgram &p, int k, cx<double> **u_dev) {
    const size_t nb = p.qsets.size();
    const size_t ops_b = ((p.ops.size() * sizeof(ComposeOp) + 255) / 256) * 256;
    const size_t coef_b = ((p.coef.size() * sizeof(cx<double>) + 255) / 256) * 256;
    const size_t beg_b = ((p.op_begin.size() * sizeof(uint32_t) + 255) / 256) * 256;
    const size_t mat_b = nb * 1024 * sizeof(cx<double>);
    if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
    int rc = ensure_fused_buffer(e, ops_b + coef_b + beg_b + mat_b + 256);
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
