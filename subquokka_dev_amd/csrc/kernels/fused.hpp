// fused.hpp -- dense five-qubit blocks on MFMA.
//
// The planner groups gates into blocks acting on at most five qubits Q; a
// block's gates are composed in FP64 into one 32x32 complex matrix U (the same
// map, in one pass over the state instead of one pass per gate -- the role the
// reference gives its cache-resident subcircuits, MEM_Runner.cpp:173-182).
//
// With Q fixed, the state is a 32 x 2^(n-5) matrix X (row j = local index on
// Q, column = the other n-5 index bits) and the block computes Y = U X.  In
// real form [Yr; Yi] = [[Ur, -Ui], [Ui, Ur]] [Xr; Xi]: a 64x64 real matrix M
// times 64 x 16 column batches, done with v_mfma_{f32,f64}_16x16x4 (K = 64 in
// 16 steps, four 16-row output tiles).  Operand layout of the 16x16x4 MFMA
// (lane l, g = l/16, c = l%16):  A[c][g], B[g][c]; D[4g + r][c] (f32) or
// D[g + 4r][c] (f64) for r = 0..3 (both measured on gfx950).
//   * A: lane l keeps M[16 rt + c][4 st + g] for rt < 4, st < 16 (64 values),
//     loaded once per wave and reused for every batch.
//   * B at step st is row 4 st + g of X: Re x_j (st < 8) or Im x_j (st >= 8)
//     with j = 4 (st mod 8) + g, so lane l loads amplitudes j = g + 4t, t < 8,
//     of column c.
//   * D tile rt holds row d = row(g, r) of its 16: real row 16 rt + d, so lane l
//     stores y_d = (tile 0, tile 2) and y_{16+d} = (tile 1, tile 3).
// Every column is read and written by exactly one wave, so the update is in place.
#pragma once
#include "common.hpp"

namespace sqk {

constexpr int kFusedK = 5;

template <typename T> struct MfmaTypes;
template <> struct MfmaTypes<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(float a, float b, v4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    static __host__ __device__ constexpr int row(int g, int r) { return 4 * g + r; }   // D row in register r
};
template <> struct MfmaTypes<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(double a, double b, v4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    static __host__ __device__ constexpr int row(int g, int r) { return g + 4 * r; }
};

struct FusedArgs {
    const cx<double> *u;   // 32x32 complex, row-major, FP64
    uint64_t n_batches;    // 16-column batches = 2^(n-5) / 16
    int q[kFusedK];        // block qubits ascending; local bit j <-> q[j]
};

__device__ inline uint64_t pdep5(int j, const int *q) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < kFusedK; b++) if ((j >> b) & 1) o += 1ULL << q[b];
    return o;
}

template <typename T, int U>
__global__ void __launch_bounds__(256) k_fused5(cx<T> *__restrict__ s, const FusedArgs a) {
    using V = typename MfmaTypes<T>::v4;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, c = lane & 15;
    const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const uint64_t n_waves = (uint64_t)gridDim.x * 4;

    T am[4][16];
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
#pragma unroll
        for (int st = 0; st < 16; st++) {
            const int i = 16 * rt + c, k = 4 * st + g;
            const cx<double> z = a.u[(i & 31) * 32 + (k & 31)];
            const double v = (i < 32) ? ((k < 32) ? z.re : -z.im) : ((k < 32) ? z.im : z.re);
            am[rt][st] = (T)v;
        }
    }
    uint64_t in_off[8], out_off[8];
#pragma unroll
    for (int t = 0; t < 8; t++) in_off[t] = pdep5(g + 4 * t, a.q);
#pragma unroll
    for (int r = 0; r < 4; r++) {
        out_off[r] = pdep5(MfmaTypes<T>::row(g, r), a.q);
        out_off[4 + r] = pdep5(16 + MfmaTypes<T>::row(g, r), a.q);
    }

    for (uint64_t b0 = wave * U; b0 < a.n_batches; b0 += n_waves * U) {
        uint64_t base[U];
        cx<T> x[U][8];
#pragma unroll
        for (int u = 0; u < U; u++) {
            uint64_t col = ((b0 + u) << 4) + c;
#pragma unroll
            for (int b = 0; b < kFusedK; b++) col = insert_zero(col, a.q[b]);
            base[u] = col;
            if (b0 + u >= a.n_batches) continue;   // uniform per wave
#pragma unroll
            for (int t = 0; t < 8; t++) x[u][t] = s[base[u] + in_off[t]];
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (b0 + u >= a.n_batches) continue;
            V acc[4];
#pragma unroll
            for (int rt = 0; rt < 4; rt++) acc[rt] = V{0, 0, 0, 0};
#pragma unroll
            for (int st = 0; st < 16; st++) {
                const T bv = st < 8 ? x[u][st].re : x[u][st - 8].im;
#pragma unroll
                for (int rt = 0; rt < 4; rt++) acc[rt] = MfmaTypes<T>::mma(am[rt][st], bv, acc[rt]);
            }
#pragma unroll
            for (int r = 0; r < 4; r++) {
                s[base[u] + out_off[r]] = cx<T>{acc[0][r], acc[2][r]};
                s[base[u] + out_off[4 + r]] = cx<T>{acc[1][r], acc[3][r]};
            }
        }
    }
}

// Scalar form for states below 9 qubits (fewer than one 16-column batch):
// one thread per column, FP64 arithmetic.
struct FusedSmallArgs {
    const cx<double> *u;
    uint64_t n_cols;
    int q[kFusedK];
};

template <typename T>
__global__ void __launch_bounds__(256) k_fused5_small(cx<T> *__restrict__ s, const FusedSmallArgs a) {
    const uint64_t col = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (col >= a.n_cols) return;
    uint64_t base = col;
    for (int b = 0; b < kFusedK; b++) base = insert_zero(base, a.q[b]);
    cx<double> x[32];
    for (int j = 0; j < 32; j++) {
        const cx<T> v = s[base + pdep5(j, a.q)];
        x[j] = {(double)v.re, (double)v.im};
    }
    for (int i = 0; i < 32; i++) {
        cx<double> acc = {0, 0};
        for (int j = 0; j < 32; j++) acc = cfma(acc, a.u[i * 32 + j], x[j]);
        s[base + pdep5(i, a.q)] = {(T)acc.re, (T)acc.im};
    }
}

// ---------------------------------------------------------------------------
// FP64 composition of each block's gates into its 2^k x 2^k matrix (k <= 5).
// One workgroup per block; thread c < 2^k owns column c of U, starts it as the
// basis vector e_c in the output buffer and applies the block's gates to it in
// order.  No two threads touch the same element, so no shared memory is used.
// ---------------------------------------------------------------------------
// Every field is 32 bits wide: op records are read with scalar (SMEM) loads,
// which drop the two low bits of their address components, so a dword field
// behind byte fields must not be reached through a non-dword-aligned base.
struct ComposeOp {
    int32_t kind;      // 0: dense 1q, 1: dense 2q, 2: swap
    int32_t t0, t1;    // local bits (t0 < t1)
    uint32_t coef;     // offset into the FP64 coefficient array
};

struct ComposeArgs {
    const ComposeOp *ops;
    const cx<double> *coef;
    const uint32_t *op_begin;   // per block: ops [op_begin[b], op_begin[b+1])
    cx<double> *u_out;          // per block 2^k x 2^k row-major, stride 1024
    int k;
};

__global__ void __launch_bounds__(64) k_compose(const ComposeArgs a) {
    const int c = threadIdx.x;
    const int d = 1 << a.k;
    if (c >= d) return;
    cx<double> *u = a.u_out + (size_t)blockIdx.x * 1024;
    // element (row r, column c) at u[r * d + c]
    for (int r = 0; r < d; r++) u[r * d + c] = {r == c ? 1.0 : 0.0, 0.0};
    for (uint32_t o = a.op_begin[blockIdx.x]; o < a.op_begin[blockIdx.x + 1]; o++) {
        const ComposeOp op = a.ops[o];
        const cx<double> *m = a.coef + op.coef;
        if (op.kind == 0) {
            for (int p = 0; p < d / 2; p++) {
                const int i0 = (int)insert_zero(p, op.t0), i1 = i0 | (1 << op.t0);
                const cx<double> x = u[i0 * d + c], y = u[i1 * d + c];
                u[i0 * d + c] = cfma(cmul(m[0], x), m[1], y);
                u[i1 * d + c] = cfma(cmul(m[2], x), m[3], y);
            }
        } else if (op.kind == 1) {
            for (int p = 0; p < d / 4; p++) {
                const int i = (int)insert_zero(insert_zero(p, op.t0), op.t1);
                const int i1 = i | (1 << op.t0), i2 = i | (1 << op.t1), i3 = i1 | (1 << op.t1);
                const cx<double> v0 = u[i * d + c], v1 = u[i1 * d + c], v2 = u[i2 * d + c], v3 = u[i3 * d + c];
                const cx<double> w0 = cfma(cfma(cfma(cmul(m[0], v0), m[1], v1), m[2], v2), m[3], v3);
                const cx<double> w1 = cfma(cfma(cfma(cmul(m[4], v0), m[5], v1), m[6], v2), m[7], v3);
                const cx<double> w2 = cfma(cfma(cfma(cmul(m[8], v0), m[9], v1), m[10], v2), m[11], v3);
                const cx<double> w3 = cfma(cfma(cfma(cmul(m[12], v0), m[13], v1), m[14], v2), m[15], v3);
                u[i * d + c] = w0; u[i1 * d + c] = w1; u[i2 * d + c] = w2; u[i3 * d + c] = w3;
            }
        } else {
            for (int p = 0; p < d / 4; p++) {
                const int i = (int)insert_zero(insert_zero(p, op.t0), op.t1);
                const int ia = i | (1 << op.t0), ib = i | (1 << op.t1);
                const cx<double> x = u[ia * d + c];
                u[ia * d + c] = u[ib * d + c];
                u[ib * d + c] = x;
            }
        }
    }
}

}  // namespace sqk
