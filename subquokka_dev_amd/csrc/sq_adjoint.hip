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
