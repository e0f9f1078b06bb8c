// pass.hpp -- register multi-gate pass for the plain sweep.
//
// MEM_Runner::run(vector<Gate*>&) (MEM_Runner.cpp:101-134) reads and writes
// the whole state once per gate.  A register pass applies a run of gates to
// the state in one read and one write: each wave holds a "wave tile" of
// 2^KH x 64 16-byte vectors in registers -- index bits [0, VB) inside the
// vector, bits [VB, VB+6) across the lanes and KH chosen qubits hpos[] as the
// register index -- and applies every gate of the pass in order before
// storing the tile back.
//
// Gates a pass can hold:
//   * dense one-qubit gates on a vector, lane or register bit (H X Y U1 RX RY);
//   * diagonal one- and two-qubit gates on ANY qubits (Z P RZ CP RZZ): the
//     phase depends only on the amplitude's global index, which every lane
//     knows.
// Each gate uses the same operation sequence as the single-gate kernels
// (k_dense: acc = 0, then cfma over the row; k_diag: cmul(amp, d)), so a pass
// computes the same values as the gates applied one launch at a time.
#pragma once
#include "common.hpp"

namespace sqk {

enum PassOpKind : int32_t { PO_DENSE_REG = 0, PO_DENSE_LANE = 1, PO_DENSE_VEC = 2, PO_DIAG1 = 3, PO_DIAG2 = 4 };

// 32-bit fields only: the op loop reads records with scalar (SMEM) loads,
// which ignore the two low bits of their address components; byte fields
// beside the dword field let the compiler address it from a base 2 bytes off
// alignment, and the load then returns the byte fields instead.
struct PassOp {
    int32_t kind;
    int32_t a, b;      // DENSE_REG: register slot; DENSE_LANE: lane bit; DIAG: global bits (a < b)
    uint32_t coef;     // offset of the op's coefficients (4 for dense, 2 or 4 for diagonal)
};

constexpr int kPassMaxKH = 4;   // five register qubits spill on gfx950

template <typename T> struct PassArgs {
    const PassOp *ops;
    const cx<T> *coef;
    uint64_t n_tiles;        // wave tiles = 2^(n - VB - 6 - KH)
    int n_ops;
    int hpos[kPassMaxKH];    // register qubits, ascending, all >= VB + 6
};

// One dense 2x2 on the register pairs (r, r | 2^J).
template <typename T, int J, int R, int E>
__device__ inline void pass_dense_reg(Pack<T> (&v)[R], const cx<T> m0, const cx<T> m1, const cx<T> m2, const cx<T> m3) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        if ((r >> J) & 1) continue;
        const int q = r | (1 << J);
#pragma unroll
        for (int e = 0; e < E; e++) {
            const cx<T> x = v[r].e[e], y = v[q].e[e];
            cx<T> o0 = {0, 0}, o1 = {0, 0};
            o0 = cfma(o0, m0, x);
            o0 = cfma(o0, m1, y);
            o1 = cfma(o1, m2, x);
            o1 = cfma(o1, m3, y);
            v[r].e[e] = o0;
            v[q].e[e] = o1;
        }
    }
}

// Value of global index bit `pos` for register r, vector element e of a lane
// whose tile base (lane bits included) is `base`: the bit is a register slot,
// a bit inside the 16-byte vector, or a bit of the base.
template <int KH> struct BitSrc {
    int slot;     // register slot holding the bit, or -1
    int vec;      // 1: the bit is the vector element index
    bool fixed;   // the bit's value when it comes from the base
    __device__ BitSrc(int pos, const int *hpos, uint64_t base, int vb) {
        slot = -1;
#pragma unroll
        for (int j = 0; j < KH; j++) if (hpos[j] == pos) slot = j;
        vec = pos < vb;
        fixed = (base >> pos) & 1;
    }
    __device__ bool get(int r, int e) const {
        return slot >= 0 ? ((r >> slot) & 1) : (vec ? (e & 1) : fixed);
    }
};

template <typename T, int KH>
__global__ void __launch_bounds__(kBlock) k_pass(cx<T> *__restrict__ s, const PassArgs<T> a) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    constexpr int E = 1 << VB;
    constexpr int R = 1 << KH;
    const int lane = threadIdx.x & 63;
    const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint64_t n_waves = (uint64_t)gridDim.x * (kBlock / 64);

    uint64_t hb[KH];
#pragma unroll
    for (int j = 0; j < KH; j++) hb[j] = 1ULL << a.hpos[j];

    for (uint64_t t = wave; t < a.n_tiles; t += n_waves) {
        uint64_t base = t << S;
#pragma unroll
        for (int j = 0; j < KH; j++) base = insert_zero(base, a.hpos[j]);
        base += (uint64_t)lane << VB;

        Pack<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            uint64_t off = base;
#pragma unroll
            for (int j = 0; j < KH; j++) if ((r >> j) & 1) off += hb[j];
            v[r] = ld(s + off);
        }

        for (int op = 0; op < a.n_ops; op++) {
            const PassOp o = a.ops[op];
            const cx<T> *c = a.coef + o.coef;
            switch (o.kind) {
            case PO_DENSE_REG: {
                const cx<T> m0 = c[0], m1 = c[1], m2 = c[2], m3 = c[3];
                switch (o.a) {
                case 0: pass_dense_reg<T, 0, R, E>(v, m0, m1, m2, m3); break;
                case 1: if constexpr (KH > 1) pass_dense_reg<T, 1, R, E>(v, m0, m1, m2, m3); break;
                case 2: if constexpr (KH > 2) pass_dense_reg<T, 2, R, E>(v, m0, m1, m2, m3); break;
                case 3: if constexpr (KH > 3) pass_dense_reg<T, 3, R, E>(v, m0, m1, m2, m3); break;
                default: break;
                }
                break; }
            case PO_DENSE_LANE: {
                // lane bit 0: out = m0*self + m1*partner; lane bit 1: out = m2*partner + m3*self
                const bool hi = (lane >> o.a) & 1;
                const cx<T> mA = hi ? c[2] : c[0], mB = hi ? c[3] : c[1];
                const int mask = 1 << o.a;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const Pack<T> p = shfl_xor_pack(v[r], mask);
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        const cx<T> first = hi ? p.e[e] : v[r].e[e], second = hi ? v[r].e[e] : p.e[e];
                        cx<T> acc = {0, 0};
                        acc = cfma(acc, mA, first);
                        acc = cfma(acc, mB, second);
                        v[r].e[e] = acc;
                    }
                }
                break; }
            case PO_DENSE_VEC: {
                if constexpr (E == 2) {
                    const cx<T> m0 = c[0], m1 = c[1], m2 = c[2], m3 = c[3];
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        const cx<T> x = v[r].e[0], y = v[r].e[1];
                        cx<T> o0 = {0, 0}, o1 = {0, 0};
                        o0 = cfma(o0, m0, x);
                        o0 = cfma(o0, m1, y);
                        o1 = cfma(o1, m2, x);
                        o1 = cfma(o1, m3, y);
                        v[r].e[0] = o0;
                        v[r].e[1] = o1;
                    }
                }
                break; }
            case PO_DIAG1: {
                const cx<T> d0 = c[0], d1 = c[1];
                const BitSrc<KH> ba(o.a, a.hpos, base, VB);
#pragma unroll
                for (int r = 0; r < R; r++) {
#pragma unroll
                    for (int e = 0; e < E; e++) v[r].e[e] = cmul(v[r].e[e], ba.get(r, e) ? d1 : d0);
                }
                break; }
            case PO_DIAG2: {
                const cx<T> d0 = c[0], d1 = c[1], d2 = c[2], d3 = c[3];
                const BitSrc<KH> ba(o.a, a.hpos, base, VB), bb(o.b, a.hpos, base, VB);
#pragma unroll
                for (int r = 0; r < R; r++) {
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        const bool x = ba.get(r, e), y = bb.get(r, e);
                        v[r].e[e] = cmul(v[r].e[e], y ? (x ? d3 : d2) : (x ? d1 : d0));
                    }
                }
                break; }
            default:
                break;
            }
        }

#pragma unroll
        for (int r = 0; r < R; r++) {
            uint64_t off = base;
#pragma unroll
            for (int j = 0; j < KH; j++) if ((r >> j) & 1) off += hb[j];
            st(s + off, v[r]);
        }
    }
}

}  // namespace sqk
