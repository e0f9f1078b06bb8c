// adjoint.hpp -- kernels of the adjoint (reverse-sweep) gradient.
//
// For a rotation G(theta) = exp(-i theta P / 2), P in {Y, Z} (RY, RZ with the
// reference's matrices, circuit.cpp:2736-2737, 2764-2765), and energy
// E = <psi|H|psi>, the reverse sweep holds psi_k (state after gate k) and
// lambda_k = G_{k+1}^dag ... G_L^dag H psi_L and computes
//     dE/dtheta_k = 2 Re <lambda_k| dG_k/dtheta |psi_{k-1}> = Im <lambda_k| P |psi_k>,
// then un-applies G_k to both vectors.  One kernel does this for both rotations
// of a qubit in a single pass over psi and lambda; the inner products are
// accumulated in FP64.
#pragma once
#include "common.hpp"
#include "reduce.hpp"

namespace sqk {

struct AdjRot2Args {
    uint64_t n_pairs;
    int q;
    double cz, sz;     // cos, sin of theta/2 for RZ(q)
    double cy, sy;     // cos, sin of theta/2 for RY(q)
};

// The two rotations of one qubit in one pass: the reverse sweep meets RZ(q)
// and then RY(q) of a layer.  Rotations on other qubits commute with both
// (and with their generators), so the sweep may take each qubit's pair
// together.  Per amplitude pair: the RZ gradient term, un-apply RZ, the RY
// gradient term from the updated values, un-apply RY -- the same arithmetic,
// in the same order and precision, as one pass per rotation.  acc[0] is the RY
// parameter's partial sum and acc[1] the RZ parameter's, written `stride`
// apart.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_adj_rot2(cx<T> *__restrict__ psi, cx<T> *__restrict__ lam,
                                                     const AdjRot2Args a, double *__restrict__ part, uint64_t stride) {
    double acc[2] = {0.0, 0.0};
    const uint64_t gstride = (uint64_t)gridDim.x * kBlock;
    const T cz = (T)a.cz, sz = (T)a.sz, cy = (T)a.cy, sy = (T)a.sy;
    const cx<T> ep = {cz, sz}, en = {cz, -sz};
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < a.n_pairs; p += gstride) {
        const uint64_t i0 = insert_zero(p, a.q), i1 = i0 | (1ULL << a.q);
        cx<T> x0 = psi[i0], x1 = psi[i1], l0 = lam[i0], l1 = lam[i1];
        {
            // <l|Z x> = conj(l0) x0 - conj(l1) x1; Im(conj(l) x) = lr*xi - li*xr
            const double x0r = x0.re, x0i = x0.im, x1r = x1.re, x1i = x1.im;
            const double l0r = l0.re, l0i = l0.im, l1r = l1.re, l1i = l1.im;
            acc[1] += (l0r * x0i - l0i * x0r) - (l1r * x1i - l1i * x1r);
            // RZ(theta)^dag = diag(e^{+i theta/2}, e^{-i theta/2})
            x0 = cmul(x0, ep);
            x1 = cmul(x1, en);
            l0 = cmul(l0, ep);
            l1 = cmul(l1, en);
        }
        {
            // <l|Y x> = conj(l0)(-i x1) + conj(l1)(i x0); Im of it
            const double x0r = x0.re, x0i = x0.im, x1r = x1.re, x1i = x1.im;
            const double l0r = l0.re, l0i = l0.im, l1r = l1.re, l1i = l1.im;
            acc[0] += (l1r * x0r + l1i * x0i) - (l0r * x1r + l0i * x1i);
            // RY(theta)^dag = [[c, s], [-s, c]]
            psi[i0] = {cy * x0.re + sy * x1.re, cy * x0.im + sy * x1.im};
            psi[i1] = {cy * x1.re - sy * x0.re, cy * x1.im - sy * x0.im};
            lam[i0] = {cy * l0.re + sy * l1.re, cy * l0.im + sy * l1.im};
            lam[i1] = {cy * l1.re - sy * l0.re, cy * l1.im - sy * l0.im};
        }
    }
    block_sum<2>(acc, part + blockIdx.x, stride);
}

// CNOT (self-inverse) on two vectors: swap target 0/1 where the control is 1.
struct AdjCnotArgs {
    uint64_t n_quads;
    int ctrl, targ;
    int lo, hi;        // min/max of (ctrl, targ)
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_cnot2(cx<T> *__restrict__ x, cx<T> *__restrict__ y, const AdjCnotArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < a.n_quads; p += stride) {
        const uint64_t i = insert_zero(insert_zero(p, a.lo), a.hi) | (1ULL << a.ctrl);
        const uint64_t j = i | (1ULL << a.targ);
        const cx<T> u = x[i];
        x[i] = x[j];
        x[j] = u;
        if (y) {
            const cx<T> v = y[i];
            y[i] = y[j];
            y[j] = v;
        }
    }
}

// lambda = H psi with H = sum_i Z_i Z_{i+1 mod n} (diagonal).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_apply_zz_ring(const cx<T> *__restrict__ psi, cx<T> *__restrict__ lam,
                                                          uint64_t n_amp, int n) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t mask = (n >= 64) ? ~0ULL : ((1ULL << n) - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_amp; i += stride) {
        const uint64_t rot = ((i >> 1) | (i << (n - 1))) & mask;
        const int differ = (int)__popcll((i ^ rot) & mask);   // __popcll is unsigned: keep the sum signed
        const T zz = (T)(n - 2 * differ);
        lam[i] = {zz * psi[i].re, zz * psi[i].im};
    }
}

}  // namespace sqk
