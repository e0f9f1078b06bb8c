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


// ---------------------------------------------------------------------------
// General parameterised circuits: one kernel per parameterised gate takes its
// gradient term and un-applies it on psi and lambda in one pass.
//   RX, RY, RZ, RZZ:  G = exp(-i t P / 2)  ->  dE/dt = Im <lambda|P|psi>
//   P, CP:            G = exp(i t Pi)      ->  dE/dt = -2 Im <lambda|Pi|psi>
// (Pi the projector on |1> / |11>), with psi the state after G and lambda the
// observable applied to the final state and un-applied down to G.
// ---------------------------------------------------------------------------

enum AdjKind : int32_t { ADJ_RX = 0, ADJ_RY = 1, ADJ_RZ = 2, ADJ_P = 3, ADJ_CP = 4, ADJ_RZZ = 5 };

struct AdjGateArgs {
    uint64_t n_items;   // pairs (one-qubit kinds) or amplitudes (CP, RZZ)
    int kind;
    int q0, q1;         // qubits (q1: second qubit of CP / RZZ)
    double c, s;        // cos, sin of the half angle (RX RY RZ RZZ) or of the angle (P CP)
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_adj_gate(cx<T> *__restrict__ psi, cx<T> *__restrict__ lam,
                                                     const AdjGateArgs a, double *__restrict__ part) {
    double acc[1] = {0.0};
    const uint64_t gstride = (uint64_t)gridDim.x * kBlock;
    const T c = (T)a.c, s = (T)a.s;
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < a.n_items; p += gstride) {
        if (a.kind == ADJ_CP || a.kind == ADJ_RZZ) {
            const uint64_t i = p;
            const int b0 = (int)((i >> a.q0) & 1), b1 = (int)((i >> a.q1) & 1);
            const cx<T> x = psi[i], l = lam[i];
            // Im(conj(l) x)
            const double im = (double)l.re * (double)x.im - (double)l.im * (double)x.re;
            if (a.kind == ADJ_CP) {
                if (b0 & b1) {
                    acc[0] += -2.0 * im;
                    const cx<T> e = {c, -s};   // e^{-i t}
                    psi[i] = cmul(x, e);
                    lam[i] = cmul(l, e);
                }
            } else {
                const int z = (b0 == b1) ? 1 : -1;
                acc[0] += z * im;
                const cx<T> e = {c, z > 0 ? s : -s};   // RZZ^dag = e^{i t z / 2}
                psi[i] = cmul(x, e);
                lam[i] = cmul(l, e);
            }
            continue;
        }
        const uint64_t i0 = insert_zero(p, a.q0), i1 = i0 | (1ULL << a.q0);
        const cx<T> x0 = psi[i0], x1 = psi[i1], l0 = lam[i0], l1 = lam[i1];
        const double x0r = x0.re, x0i = x0.im, x1r = x1.re, x1i = x1.im;
        const double l0r = l0.re, l0i = l0.im, l1r = l1.re, l1i = l1.im;
        switch (a.kind) {
        case ADJ_RX: {
            // <l|X x> = conj(l0) x1 + conj(l1) x0
            acc[0] += (l0r * x1i - l0i * x1r) + (l1r * x0i - l1i * x0r);
            // RX(t)^dag = [[c, i s], [i s, c]]
            psi[i0] = {c * x0.re - s * x1.im, c * x0.im + s * x1.re};
            psi[i1] = {c * x1.re - s * x0.im, c * x1.im + s * x0.re};
            lam[i0] = {c * l0.re - s * l1.im, c * l0.im + s * l1.re};
            lam[i1] = {c * l1.re - s * l0.im, c * l1.im + s * l0.re};
            break; }
        case ADJ_RY: {
            // <l|Y x> = conj(l0)(-i x1) + conj(l1)(i x0)
            acc[0] += (l1r * x0r + l1i * x0i) - (l0r * x1r + l0i * x1i);
            // RY(t)^dag = [[c, s], [-s, c]]
            psi[i0] = {c * x0.re + s * x1.re, c * x0.im + s * x1.im};
            psi[i1] = {c * x1.re - s * x0.re, c * x1.im - s * x0.im};
            lam[i0] = {c * l0.re + s * l1.re, c * l0.im + s * l1.im};
            lam[i1] = {c * l1.re - s * l0.re, c * l1.im - s * l0.im};
            break; }
        case ADJ_RZ: {
            // <l|Z x> = conj(l0) x0 - conj(l1) x1
            acc[0] += (l0r * x0i - l0i * x0r) - (l1r * x1i - l1i * x1r);
            // RZ(t)^dag = diag(e^{i t/2}, e^{-i t/2})
            const cx<T> ep = {c, s}, en = {c, -s};
            psi[i0] = cmul(x0, ep);
            psi[i1] = cmul(x1, en);
            lam[i0] = cmul(l0, ep);
            lam[i1] = cmul(l1, en);
            break; }
        default: {   // ADJ_P: dE/dt = -2 Im(conj(l1) x1); P(t)^dag = diag(1, e^{-i t})
            acc[0] += -2.0 * (l1r * x1i - l1i * x1r);
            const cx<T> e = {c, -s};
            psi[i1] = cmul(x1, e);
            lam[i1] = cmul(l1, e);
            break; }
        }
    }
    block_sum<1>(acc, part + blockIdx.x);
}

// Up to four consecutive parameterised one-qubit gates (RX RY RZ P) on
// distinct qubits, un-applied in one pass over psi and lambda.  Gate j's
// generator acts on its own qubit and commutes with the other gates of the
// group, so its term <lambda_j|P_j|psi_j> can be taken with the other gates
// un-applied or not, in any order.  Each thread holds the 2^M amplitudes that
// differ only in the group's qubits (local bit b <-> qubit q[b]).
struct AdjMultiArgs {
    uint64_t n_items;   // 2^n / 2^M
    int kind[4];        // ADJ_RX, ADJ_RY, ADJ_RZ or ADJ_P
    int q[4];           // the gates' qubits
    int qs[4];          // the same, ascending (where zeros are inserted)
    double c[4], s[4];  // cos, sin of the half angle (RX RY RZ) or of the angle (P)
};

template <typename T, int M>
__global__ void __launch_bounds__(kBlock) k_adj_multi(cx<T> *__restrict__ psi, cx<T> *__restrict__ lam,
                                                      const AdjMultiArgs a, double *__restrict__ part,
                                                      uint64_t part_stride) {
    constexpr int D = 1 << M;
    double acc[M];
#pragma unroll
    for (int b = 0; b < M; b++) acc[b] = 0.0;
    const uint64_t gstride = (uint64_t)gridDim.x * kBlock;
    uint64_t off[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        uint64_t o = 0;
#pragma unroll
        for (int b = 0; b < M; b++) o |= (uint64_t)((j >> b) & 1) << a.q[b];
        off[j] = o;
    }
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < a.n_items; p += gstride) {
        uint64_t base = p;
#pragma unroll
        for (int b = 0; b < M; b++) base = insert_zero(base, a.qs[b]);
        cx<T> x[D], l[D];
#pragma unroll
        for (int j = 0; j < D; j++) {
            x[j] = psi[base + off[j]];
            l[j] = lam[base + off[j]];
        }
        // Gate b's term, then gate b un-applied: the later terms are unchanged,
        // since gate b and the other gates' generators act on other qubits.
#pragma unroll
        for (int b = 0; b < M; b++) {
            const T c = (T)a.c[b], s = (T)a.s[b];
            switch (a.kind[b]) {   // uniform; one loop per kind keeps one path's temporaries live
            case ADJ_RX:
#pragma unroll
                for (int j = 0; j < D; j++) {
                    if ((j >> b) & 1) continue;
                    const int j1 = j | (1 << b);
                    const cx<T> x0 = x[j], x1 = x[j1], l0 = l[j], l1 = l[j1];
                    acc[b] += ((double)l0.re * x1.im - (double)l0.im * x1.re) + ((double)l1.re * x0.im - (double)l1.im * x0.re);
                    // RX(t)^dag = [[c, i s], [i s, c]]
                    x[j] = {c * x0.re - s * x1.im, c * x0.im + s * x1.re};
                    x[j1] = {c * x1.re - s * x0.im, c * x1.im + s * x0.re};
                    l[j] = {c * l0.re - s * l1.im, c * l0.im + s * l1.re};
                    l[j1] = {c * l1.re - s * l0.im, c * l1.im + s * l0.re};
                }
                break;
            case ADJ_RY:
#pragma unroll
                for (int j = 0; j < D; j++) {
                    if ((j >> b) & 1) continue;
                    const int j1 = j | (1 << b);
                    const cx<T> x0 = x[j], x1 = x[j1], l0 = l[j], l1 = l[j1];
                    acc[b] += ((double)l1.re * x0.re + (double)l1.im * x0.im) - ((double)l0.re * x1.re + (double)l0.im * x1.im);
                    // RY(t)^dag = [[c, s], [-s, c]]
                    x[j] = {c * x0.re + s * x1.re, c * x0.im + s * x1.im};
                    x[j1] = {c * x1.re - s * x0.re, c * x1.im - s * x0.im};
                    l[j] = {c * l0.re + s * l1.re, c * l0.im + s * l1.im};
                    l[j1] = {c * l1.re - s * l0.re, c * l1.im - s * l0.im};
                }
                break;
            case ADJ_RZ:
#pragma unroll
                for (int j = 0; j < D; j++) {
                    if ((j >> b) & 1) continue;
                    const int j1 = j | (1 << b);
                    const cx<T> x0 = x[j], x1 = x[j1], l0 = l[j], l1 = l[j1];
                    acc[b] += ((double)l0.re * x0.im - (double)l0.im * x0.re) - ((double)l1.re * x1.im - (double)l1.im * x1.re);
                    // RZ(t)^dag = diag(e^{i t/2}, e^{-i t/2})
                    x[j] = cmul(x0, cx<T>{c, s});
                    x[j1] = cmul(x1, cx<T>{c, -s});
                    l[j] = cmul(l0, cx<T>{c, s});
                    l[j1] = cmul(l1, cx<T>{c, -s});
                }
                break;
            default:   // ADJ_P: dE/dt = -2 Im(conj(l1) x1); P(t)^dag = diag(1, e^{-i t})
#pragma unroll
                for (int j = 0; j < D; j++) {
                    if ((j >> b) & 1) continue;
                    const int j1 = j | (1 << b);
                    const cx<T> x1 = x[j1], l1 = l[j1];
                    acc[b] += -2.0 * ((double)l1.re * x1.im - (double)l1.im * x1.re);
                    x[j1] = cmul(x1, cx<T>{c, -s});
                    l[j1] = cmul(l1, cx<T>{c, -s});
                }
                break;
            }
        }
#pragma unroll
        for (int j = 0; j < D; j++) {
            psi[base + off[j]] = x[j];
            lam[base + off[j]] = l[j];
        }
    }
    block_sum<M>(acc, part + blockIdx.x, part_stride);
}

// A diagonal observable as Z strings: h(i) = sum_t coef_t (-1)^{popcount(i & mask_t)}.
struct ZTerm {
    double coef;
    uint64_t mask;
};

// lambda = H psi, and the block partials of <psi|H|psi> (FP64).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_apply_zterms(const cx<T> *__restrict__ psi, cx<T> *__restrict__ lam,
                                                         const ZTerm *__restrict__ terms, int n_terms, uint64_t n_amp,
                                                         double *__restrict__ part) {
    double acc[1] = {0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_amp; i += stride) {
        double h = 0.0;
        for (int t = 0; t < n_terms; t++) h += (__popcll(i & terms[t].mask) & 1) ? -terms[t].coef : terms[t].coef;
        const cx<T> x = psi[i];
        acc[0] += h * ((double)x.re * (double)x.re + (double)x.im * (double)x.im);
        if (lam) lam[i] = {(T)(h * (double)x.re), (T)(h * (double)x.im)};
    }
    block_sum<1>(acc, part + blockIdx.x);
}

}  // namespace sqk
