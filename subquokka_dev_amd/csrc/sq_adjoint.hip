// sq_adjoint.hip -- energy and adjoint gradient of the hardware-efficient
// ansatz (layers x [RY on every qubit, RZ on every qubit, CNOT ring
// i -> (i+1) mod n]) for H = sum_i Z_i Z_{i+1 mod n}.  No reference
// counterpart: it reuses the engine's rotation matrices (RY, RZ as in
// circuit.cpp:2734-2795) and its state in HBM.
//
// Forward: the gates in order on the engine's state, saving an FP64 copy of
// the state after every `checkpoint_every` layers.  Backward (reverse sweep):
// lambda = H psi; for each gate from last to first, one pass over psi and
// lambda computes the parameter's derivative Im<lambda|P|psi> (FP64 block
// partials) and un-applies the gate to both.  At a checkpoint boundary psi is
// reset to the saved FP64 state, so rounding drift of the un-applied state
// (complex64) is bounded by checkpoint_every layers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "kernels/adjoint.hpp"
#include "kernels/reduce.hpp"
#include "host/gate_lowering.hpp"

using sq::LoweredGate;
using namespace sqk;

namespace {

int ensure(sq_engine *e, void **buf, size_t *cap, size_t bytes, const char *what) {
    if (bytes <= *cap) return SQ_OK;
    if (*buf) hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    hipError_t he = hipMalloc(buf, bytes);
    if (he != hipSuccess) return e->hip_fail(he, what);
    *cap = bytes;
    return SQ_OK;
}

template <typename T>
int cnot(sq_engine *e, cx<T> *x, cx<T> *y, int c, int t) {
    AdjCnotArgs a{};
    a.n_quads = 1ULL << (e->n - 2);
    a.ctrl = c; a.targ = t;
    a.lo = std::min(c, t); a.hi = std::max(c, t);
    const double bytes = (y ? 2.0 : 1.0) * e->amp_bytes() * std::ldexp(1.0, e->n);   // half of each state
    int slot = e->stat_begin(y ? "adj_cnot2" : "cnot", bytes);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cnot2<T>), e->grid_flat(a.n_quads, kBlock), kBlock, 0, e->stream, x, y, a);
    e->stat_end(slot);
    return e->check_launch("k_cnot2");
}

sq_gate rotation(int kind, int q, double theta) {
    sq_gate g{};
    g.kind = kind;
    g.nq = 1;
    g.q[0] = q;
    g.param = theta;
    return g;
}

template <typename T>
int forward(sq_engine *e, int layers, const double *theta, int ckpt_every, cx<double> *ckpt) {
    const int n = e->n;
    int rc = sq_init_zero(e);
    if (rc) return rc;
    cx<T> *psi = reinterpret_cast<cx<T> *>(e->d_state);
    for (int l = 0; l < layers; l++) {
        for (int half = 0; half < 2; half++)
            for (int q = 0; q < n; q++) {
                LoweredGate lg;
                rc = e->lower(rotation(half == 0 ? SQ_GATE_RY : SQ_GATE_RZ, q, theta[(size_t)l * 2 * n + half * n + q]), lg);
                if (!rc) rc = e->apply_lowered(lg);
                if (rc) return rc;
            }
        if (n > 1)
            for (int q = 0; q < n; q++) {
                rc = cnot<T>(e, psi, nullptr, q, (q + 1) % n);
                if (rc) return rc;
            }
        if (ckpt && ckpt_every > 0 && (l + 1) % ckpt_every == 0 && l + 1 < layers) {
            const int slot = (l + 1) / ckpt_every - 1;
            const uint64_t na = e->n_amps();
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<T, double>), e->grid_for(na, kBlock), kBlock, 0, e->stream,
                               psi, ckpt + (size_t)slot * na, na);
            rc = e->check_launch("checkpoint");
            if (rc) return rc;
        }
    }
    return SQ_OK;
}

template <typename T>
int gradient(sq_engine *e, int layers, const double *theta, int ckpt_every, double *energy, double *grad) {
    const int n = e->n;
    const uint64_t na = e->n_amps();
    const int n_ckpt = (grad && ckpt_every > 0) ? std::max(0, (layers - 1) / ckpt_every) : 0;
    const unsigned grid = (unsigned)std::min<uint64_t>(kRedBlocks, std::max<uint64_t>(1, (na / 2 + kBlock - 1) / kBlock));
    const size_t n_params = (size_t)2 * n * layers;
    int rc = SQ_OK;
    if (grad) {
        rc = ensure(e, &e->d_aux, &e->aux_cap, e->state_bytes(), "hipMalloc(lambda)");
        if (!rc && n_ckpt)
            rc = ensure(e, &e->d_ckpt, &e->ckpt_cap, (size_t)n_ckpt * na * sizeof(cx<double>), "hipMalloc(checkpoints)");
        if (!rc) rc = ensure(e, &e->d_gradpart, &e->gradpart_cap, n_params * grid * sizeof(double), "hipMalloc(gradient partials)");
        if (rc) return rc;
    }
    cx<double> *ckpt = n_ckpt ? static_cast<cx<double> *>(e->d_ckpt) : nullptr;
    rc = forward<T>(e, layers, theta, ckpt_every, ckpt);
    if (rc) return rc;
    double v[2];
    rc = e->norm_zz(v);
    if (rc) return rc;
    if (energy) *energy = v[1];
    if (!grad) return SQ_OK;

    cx<T> *psi = reinterpret_cast<cx<T> *>(e->d_state);
    cx<T> *lam = reinterpret_cast<cx<T> *>(e->d_aux);
    double *part = static_cast<double *>(e->d_gradpart);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_apply_zz_ring<T>), e->grid_for(na, kBlock), kBlock, 0, e->stream, psi, lam, na, n);
    rc = e->check_launch("k_apply_zz_ring");
    if (rc) return rc;
    for (int l = layers - 1; l >= 0; l--) {
        if (ckpt && (l + 1) % ckpt_every == 0 && l + 1 < layers) {
            const int slot = (l + 1) / ckpt_every - 1;   // state after layer l
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<double, T>), e->grid_for(na, kBlock), kBlock, 0, e->stream,
                               ckpt + (size_t)slot * na, psi, na);
            rc = e->check_launch("re-anchor");
            if (rc) return rc;
        }
        if (n > 1)
            for (int q = n - 1; q >= 0; q--) {
                rc = cnot<T>(e, psi, lam, q, (q + 1) % n);
                if (rc) return rc;
            }
        // each qubit's RZ then RY, qubits in reverse order (rotations on
        // different qubits commute)
        for (int q = n - 1; q >= 0; q--) {
            const size_t py = (size_t)l * 2 * n + q, pz = py + n;
            AdjRot2Args a{};
            a.n_pairs = na / 2;
            a.q = q;
            a.cz = std::cos(theta[pz] / 2);
            a.sz = std::sin(theta[pz] / 2);
            a.cy = std::cos(theta[py] / 2);
            a.sy = std::sin(theta[py] / 2);
            int slot = e->stat_begin("adj_rot2", 4.0 * e->amp_bytes() * (double)na);
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_rot2<T>), grid, kBlock, 0, e->stream, psi, lam, a, part + py * grid,
                               (uint64_t)n * grid);
            e->stat_end(slot);
            rc = e->check_launch("k_adj_rot2");
            if (rc) return rc;
        }
    }
    std::vector<double> h(n_params * grid);
    hipError_t he = hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "gradient partials");
    for (size_t p = 0; p < n_params; p++) {
        double s = 0;
        for (unsigned b = 0; b < grid; b++) s += h[p * grid + b];
        grad[p] = s;
    }
    return SQ_OK;
}

}  // namespace

extern "C" int sq_hea_energy(sq_engine *e, int32_t layers, const double *theta, uint64_t n_theta, double *energy) {
    if (!e || layers < 0 || (layers && !theta)) return SQ_ERR_ARG;
    if (n_theta != (uint64_t)2 * e->n * layers)
        return e->fail(SQ_ERR_ARG, "theta must hold 2 * total_qbit * layers angles");
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "the ansatz needs an HBM-resident state");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    return e->dtype == SQ_C128 ? gradient<double>(e, layers, theta, 0, energy, nullptr)
                               : gradient<float>(e, layers, theta, 0, energy, nullptr);
}

extern "C" int sq_adjoint_grad_hea(sq_engine *e, int32_t layers, const double *theta, uint64_t n_theta,
                                   int32_t checkpoint_every, double *energy, double *grad) {
    if (!e || layers < 0 || (layers && (!theta || !grad)) || checkpoint_every < 0) return SQ_ERR_ARG;
    if (n_theta != (uint64_t)2 * e->n * layers)
        return e->fail(SQ_ERR_ARG, "theta must hold 2 * total_qbit * layers angles");
    if (e->n < 2) return e->fail(SQ_ERR_ARG, "the ansatz needs at least two qubits");
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "the ansatz needs an HBM-resident state");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    return e->dtype == SQ_C128 ? gradient<double>(e, layers, theta, checkpoint_every, energy, grad)
                               : gradient<float>(e, layers, theta, checkpoint_every, energy, grad);
}

// ---------------------------------------------------------------------------
// General parameterised circuits (sq_adjoint_grad).
// ---------------------------------------------------------------------------

namespace {

int adj_kind(int gate_kind) {
    switch (gate_kind) {
    case SQ_GATE_RX: return ADJ_RX;
    case SQ_GATE_RY: return ADJ_RY;
    case SQ_GATE_RZ: return ADJ_RZ;
    case SQ_GATE_P: return ADJ_P;
    case SQ_GATE_CP: return ADJ_CP;
    case SQ_GATE_RZZ: return ADJ_RZZ;
    default: return -1;
    }
}

// G^dag of a fixed gate as a gate line (rotations and phases with the angle
// negated, dense matrices conjugate-transposed, X Y Z SWAP themselves).
int adjoint_gate(const sq_gate &g, sq_gate &out) {
    out = g;
    switch (g.kind) {
    case SQ_GATE_X: case SQ_GATE_Y: case SQ_GATE_Z: case SQ_GATE_SWAP:
        return SQ_OK;
    case SQ_GATE_P: case SQ_GATE_RX: case SQ_GATE_RY: case SQ_GATE_RZ: case SQ_GATE_CP: case SQ_GATE_RZZ:
        out.param = -g.param;
        return SQ_OK;
    case SQ_GATE_H: {
        // the reference's H (circuit.cpp:247) as a dense matrix; it is real
        // and symmetric, so it is its own conjugate transpose
        const double c = 0.70710678118;
        out.kind = SQ_GATE_U1;
        const double m[8] = {c, 0, c, 0, c, 0, -c, 0};
        std::copy(m, m + 8, out.mat);
        return SQ_OK;
    }
    case SQ_GATE_U1:
    case SQ_GATE_U2: {
        const int d = g.kind == SQ_GATE_U1 ? 2 : 4;
        for (int r = 0; r < d; r++)
            for (int c = 0; c < d; c++) {
                out.mat[2 * (r * d + c)] = g.mat[2 * (c * d + r)];
                out.mat[2 * (r * d + c) + 1] = -g.mat[2 * (c * d + r) + 1];
            }
        return SQ_OK;
    }
    default:
        return SQ_ERR_UNSUPPORTED;
    }
}

template <typename T>
int zterms_apply(sq_engine *e, const ZTerm *d_terms, int n_terms, bool to_lambda, double *energy) {
    const uint64_t na = e->n_amps();
    const unsigned grid = (unsigned)std::min<uint64_t>(kRedBlocks, std::max<uint64_t>(1, (na + kBlock - 1) / kBlock));
    int rc = e->ensure_partials();
    if (rc) return rc;
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_apply_zterms<T>), grid, kBlock, 0, e->stream,
                       reinterpret_cast<const cx<T> *>(e->d_state), to_lambda ? reinterpret_cast<cx<T> *>(e->d_aux) : nullptr,
                       d_terms, n_terms, na, e->d_part);
    rc = e->check_launch("k_apply_zterms");
    if (rc) return rc;
    std::vector<double> h(grid);
    hipError_t he = hipMemcpyAsync(h.data(), e->d_part, grid * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "observable partials");
    double sum = 0;
    for (double x : h) sum += x;
    *energy = sum;
    return SQ_OK;
}

// Applies a list of gates on the engine's state as register passes (the
// plain sweep), on psi (lambda_buffer = false) or on lambda.
int sweep_on(sq_engine *e, const std::vector<sq_gate> &gl, bool on_lambda) {
    if (gl.empty()) return SQ_OK;
    std::vector<LoweredGate> low(gl.size());
    for (size_t i = 0; i < gl.size(); i++) {
        int rc = e->lower(gl[i], low[i]);
        if (rc) return rc;
    }
    if (on_lambda) std::swap(e->d_state, e->d_aux);
    const int rc = e->apply_plain(gl, low);
    if (on_lambda) std::swap(e->d_state, e->d_aux);
    return rc;
}

// The same list on psi and lambda at once: each register pass is one launch
// over both vectors (twice the waves, so small states fill the GPU better).
int sweep_both(sq_engine *e, const std::vector<sq_gate> &gl) {
    if (gl.empty()) return SQ_OK;
    std::vector<LoweredGate> low(gl.size());
    for (size_t i = 0; i < gl.size(); i++) {
        int rc = e->lower(gl[i], low[i]);
        if (rc) return rc;
    }
    e->d_state2 = e->d_aux;
    const int rc = e->apply_plain(gl, low);
    e->d_state2 = nullptr;
    return rc;
}

template <typename T>
int adjoint_general(sq_engine *e, const sq_gate *gates, const int32_t *param, uint64_t ng, const double *theta,
                    uint64_t n_theta, const sq_zterm *terms, uint64_t n_terms, int ckpt_every, double *energy,
                    double *grad) {
    const uint64_t na = e->n_amps();
    // the circuit with its angles, and the parameterised gates
    std::vector<sq_gate> circ(gates, gates + ng);
    std::vector<uint64_t> pgates;   // indices of parameterised gates, in order
    for (uint64_t i = 0; i < ng; i++) {
        if (param[i] < 0) continue;
        if ((uint64_t)param[i] >= n_theta) return e->fail(SQ_ERR_ARG, "parameter index beyond theta (gate " + std::to_string(i) + ")");
        if (adj_kind(circ[i].kind) < 0)
            return e->fail(SQ_ERR_ARG, "only RX RY RZ P CP RZZ take a parameter (gate " + std::to_string(i) + ")");
        circ[i].param = theta[param[i]];
        pgates.push_back(i);
    }
    for (uint64_t i = 0; i < ng; i++) {
        LoweredGate lg;
        int rc = e->lower(circ[i], lg);
        if (rc) { e->err += " (gate " + std::to_string(i) + ")"; return rc; }
        sq_gate inv;
        if (param[i] < 0 && adjoint_gate(circ[i], inv) != SQ_OK)
            return e->fail(SQ_ERR_UNSUPPORTED, "no adjoint for gate " + std::to_string(i));
    }
    // segments between FP64 checkpoints: every ckpt_every parameterised gates
    std::vector<uint64_t> seg_end;   // gate index where each segment ends (exclusive)
    if (grad && ckpt_every > 0)
        for (size_t k = ckpt_every; k < pgates.size(); k += ckpt_every) seg_end.push_back(pgates[k]);
    seg_end.push_back(ng);
    const size_t n_ckpt = grad && ckpt_every > 0 ? seg_end.size() : 0;   // states at the segment starts
    const unsigned grid = (unsigned)std::min<uint64_t>(kRedBlocks, std::max<uint64_t>(1, (na / 2 + kBlock - 1) / kBlock));
    const size_t term_bytes = n_terms * sizeof(ZTerm);
    int rc = grad ? ensure(e, &e->d_aux, &e->aux_cap, e->state_bytes(), "hipMalloc(lambda)") : SQ_OK;
    if (!rc && n_ckpt)
        rc = ensure(e, &e->d_ckpt, &e->ckpt_cap, n_ckpt * na * sizeof(cx<double>), "hipMalloc(checkpoints)");
    if (!rc) rc = ensure(e, &e->d_gradpart, &e->gradpart_cap,
                         std::max<size_t>(1, pgates.size()) * grid * sizeof(double) + term_bytes + 256,
                         "hipMalloc(gradient partials)");
    if (rc) return rc;
    double *part = static_cast<double *>(e->d_gradpart);
    ZTerm *d_terms = reinterpret_cast<ZTerm *>(static_cast<char *>(e->d_gradpart) +
                                               ((std::max<size_t>(1, pgates.size()) * grid * sizeof(double) + 255) / 256) * 256);
    {
        std::vector<ZTerm> ht(n_terms);
        for (uint64_t t = 0; t < n_terms; t++) ht[t] = {terms[t].coef, terms[t].mask};
        if (n_terms) {
            hipError_t he = hipMemcpyAsync(d_terms, ht.data(), term_bytes, hipMemcpyHostToDevice, e->stream);
            if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
            if (he != hipSuccess) return e->hip_fail(he, "observable upload");
        }
    }
    cx<double> *ckpt = n_ckpt ? static_cast<cx<double> *>(e->d_ckpt) : nullptr;
    auto save = [&](size_t slot) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<T, double>), e->grid_for(na, kBlock), kBlock, 0, e->stream,
                           reinterpret_cast<const cx<T> *>(e->d_state), ckpt + slot * na, na);
        return e->check_launch("checkpoint");
    };
    auto restore = [&](size_t slot) {
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<double, T>), e->grid_for(na, kBlock), kBlock, 0, e->stream,
                           ckpt + slot * na, reinterpret_cast<cx<T> *>(e->d_state), na);
        return e->check_launch("re-anchor");
    };
    // forward: segment by segment as register passes, an FP64 copy of the
    // state at each segment's start
    uint64_t begin = 0;
    for (size_t sgi = 0; sgi < seg_end.size(); sgi++) {
        if (ckpt && (rc = save(sgi))) return rc;
        if ((rc = sweep_on(e, std::vector<sq_gate>(circ.begin() + begin, circ.begin() + seg_end[sgi]), false))) return rc;
        begin = seg_end[sgi];
    }
    double en = 0;
    if ((rc = zterms_apply<T>(e, d_terms, (int)n_terms, grad != nullptr, &en))) return rc;
    if (energy) *energy = en;
    if (!grad) return SQ_OK;
    // reverse sweep
    cx<T> *psi = reinterpret_cast<cx<T> *>(e->d_state);
    cx<T> *lam = reinterpret_cast<cx<T> *>(e->d_aux);
    size_t seg = seg_end.size();   // segments left to sweep back
    std::vector<sq_gate> inv_run;
    auto flush = [&]() {
        int r = SQ_OK;
        if (!inv_run.empty()) {
            r = sweep_both(e, inv_run);
            inv_run.clear();
            psi = reinterpret_cast<cx<T> *>(e->d_state);
            lam = reinterpret_cast<cx<T> *>(e->d_aux);
        }
        return r;
    };
    // Consecutive one-qubit parameterised gates on distinct qubits go to one
    // k_adj_multi launch (up to four in complex64, three in complex128);
    // two-qubit ones (CP, RZZ) to k_adj_gate.
    struct Pending {
        uint64_t gi;
        size_t pk;   // slot of its gradient partials
    };
    std::vector<Pending> group;   // descending pk (the sweep's order)
    auto launch_group = [&]() -> int {
        if (group.empty()) return SQ_OK;
        const int M = (int)group.size();
        AdjMultiArgs a{};
        a.n_items = na >> M;
        for (int b = 0; b < M; b++) {   // local bit b: the b-th smallest partial slot
            const sq_gate &g = circ[group[M - 1 - b].gi];
            a.kind[b] = adj_kind(g.kind);
            a.q[b] = g.q[0];
            const double ang = a.kind[b] == ADJ_P ? g.param : g.param / 2;
            a.c[b] = std::cos(ang);
            a.s[b] = std::sin(ang);
        }
        std::copy(a.q, a.q + M, a.qs);
        std::sort(a.qs, a.qs + M);
        double *dst = part + group.back().pk * grid;   // slots pk_min .. pk_min + M - 1
        const int slot = e->stat_begin("adj_multi", 4.0 * e->amp_bytes() * (double)na);
        switch (M) {
        case 1: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_multi<T, 1>), grid, kBlock, 0, e->stream, psi, lam, a, dst, (uint64_t)grid); break;
        case 2: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_multi<T, 2>), grid, kBlock, 0, e->stream, psi, lam, a, dst, (uint64_t)grid); break;
        case 3: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_multi<T, 3>), grid, kBlock, 0, e->stream, psi, lam, a, dst, (uint64_t)grid); break;
        default: hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_multi<T, 4>), grid, kBlock, 0, e->stream, psi, lam, a, dst, (uint64_t)grid); break;
        }
        e->stat_end(slot);
        group.clear();
        return e->check_launch("k_adj_multi");
    };
    size_t pk = pgates.size();
    for (uint64_t gi = ng; gi-- > 0;) {
        // entering the previous segment: psi is re-anchored to the FP64 copy
        // taken at this segment's end (the next one's start)
        while (seg > 1 && gi < seg_end[seg - 2]) {
            if ((rc = launch_group())) return rc;
            if ((rc = flush())) return rc;
            if ((rc = restore(seg - 1))) return rc;
            seg--;
        }
        if (param[gi] < 0) {
            if ((rc = launch_group())) return rc;
            sq_gate inv;
            adjoint_gate(circ[gi], inv);
            inv_run.push_back(inv);
            continue;
        }
        if ((rc = flush())) return rc;
        pk--;
        const sq_gate &g = circ[gi];
        const int kind = adj_kind(g.kind);
        if (kind != ADJ_CP && kind != ADJ_RZZ) {
            // four gates per launch for complex64 (171 VGPRs), three for complex128
            bool clash = group.size() == (sizeof(T) == 4 ? 4u : 3u);
            for (const auto &m : group) clash = clash || circ[m.gi].q[0] == g.q[0];
            if (clash && (rc = launch_group())) return rc;
            group.push_back({gi, pk});
            continue;
        }
        if ((rc = launch_group())) return rc;
        AdjGateArgs a{};
        a.kind = kind;
        a.n_items = na;
        a.q0 = g.q[0];
        a.q1 = g.q[1];
        const double ang = a.kind == ADJ_CP ? g.param : g.param / 2;
        a.c = std::cos(ang);
        a.s = std::sin(ang);
        // every block writes its partial (zero when it has no items)
        const int slot = e->stat_begin("adj_gate", 4.0 * e->amp_bytes() * (double)na);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_adj_gate<T>), grid, kBlock, 0, e->stream, psi, lam, a, part + pk * grid);
        e->stat_end(slot);
        if ((rc = e->check_launch("k_adj_gate"))) return rc;
    }
    if ((rc = launch_group())) return rc;
    if ((rc = flush())) return rc;
    if (ckpt && (rc = restore(0))) return rc;   // the state the call started from, exactly
    std::vector<double> h(pgates.size() * grid);
    if (!h.empty()) {
        hipError_t he = hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
        if (he != hipSuccess) return e->hip_fail(he, "gradient partials");
    }
    std::fill(grad, grad + n_theta, 0.0);
    for (size_t k = 0; k < pgates.size(); k++) {
        double sum = 0;
        for (unsigned b = 0; b < grid; b++) sum += h[k * grid + b];
        grad[param[pgates[k]]] += sum;
    }
    return SQ_OK;
}

}  // namespace

extern "C" int sq_expect_z(sq_engine *e, const sq_zterm *terms, uint64_t n_terms, double *out) {
    if (!e || (!terms && n_terms) || !out) return SQ_ERR_ARG;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "the observable needs an HBM-resident state");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    return e->dtype == SQ_C128 ? adjoint_general<double>(e, nullptr, nullptr, 0, nullptr, 0, terms, n_terms, 0, out, nullptr)
                               : adjoint_general<float>(e, nullptr, nullptr, 0, nullptr, 0, terms, n_terms, 0, out, nullptr);
}

extern "C" int sq_adjoint_grad(sq_engine *e, const sq_gate *gates, const int32_t *param, uint64_t n_gates,
                               const double *theta, uint64_t n_theta, const sq_zterm *terms, uint64_t n_terms,
                               int32_t checkpoint_every, double *energy, double *grad) {
    if (!e || (!gates && n_gates) || (!param && n_gates) || (!theta && n_theta) || (!terms && n_terms) ||
        checkpoint_every < 0 || (!grad && n_theta))
        return SQ_ERR_ARG;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "the adjoint gradient needs an HBM-resident state");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    double dummy = 0;
    double *g = grad ? grad : &dummy;
    return e->dtype == SQ_C128
               ? adjoint_general<double>(e, gates, param, n_gates, theta, n_theta, terms, n_terms, checkpoint_every, energy, g)
               : adjoint_general<float>(e, gates, param, n_gates, theta, n_theta, terms, n_terms, checkpoint_every, energy, g);
}
