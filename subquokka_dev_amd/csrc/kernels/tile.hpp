// tile.hpp -- blocked sweep: the GPU form of MEM_Runner::run(subcircuits)'s
// gate blocks (MEM_Runner.cpp:173-182), where every gate of a block runs on a
// cache-resident 2^chunk-amplitude tile before the next tile is touched.
//
// One workgroup owns one tile of 2^K amplitudes: the amplitudes whose index
// bits outside the tile qubit set Q = {q_0 < ... < q_{K-1}} equal the tile
// number.  Local index bit i <-> global bit q_i.  Q always contains the lowest
// index bits, so the tile is read and written in contiguous runs.  The tile is
// loaded into LDS, every gate of the block is applied in order (one barrier
// per gate), and the tile is written back: one HBM read and one write per
// block instead of one per gate.
#pragma once
#include "common.hpp"

namespace sqk {

enum TileKind : int32_t { TK_DENSE1 = 0, TK_DENSE2 = 1, TK_DIAG1 = 2, TK_DIAG2 = 3, TK_SWAP = 4 };

// 32-bit fields only: the op loop reads records with scalar (SMEM) loads,
// which ignore the two low bits of their address components; byte fields
// beside the dword field let the compiler address it from a base 2 bytes off
// alignment, and the load then returns the byte fields instead.
struct TileOp {
    int32_t kind;
    int32_t t0, t1;      // local target bits, t0 < t1
    uint32_t coef;       // offset of the op's coefficients in the coefficient array
};

template <typename T> struct TileArgs {
    const TileOp *ops;
    const cx<T> *coef;
    uint64_t n_tiles;
    int n_ops;
    int q[16];           // tile qubits ascending
};

template <typename T, int K, int BS>
__global__ void __launch_bounds__(BS) k_tile(cx<T> *__restrict__ s, const TileArgs<T> a) {
    constexpr int D = 1 << K;
    constexpr int M = D / BS;            // amplitudes per thread
    static_assert(M >= 4, "a tile holds at least four amplitudes per thread");
    constexpr int LB = __builtin_ctz(BS);   // local bits covered by the thread index
    // Dynamic LDS of exactly D amplitudes (launch argument); a static array in
    // this template is padded by the compiler beyond D.
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(lds_raw);
    const int tid = threadIdx.x;

    // global base of this tile: insert zeros at every tile qubit
    uint64_t base = blockIdx.x;
#pragma unroll
    for (int i = 0; i < K; i++) base = insert_zero(base, a.q[i]);
    // offset of local index j = tid + m*BS is pdep(tid) + pdep(m << LB)
    uint64_t off_lo = 0;
#pragma unroll
    for (int i = 0; i < LB; i++) if ((tid >> i) & 1) off_lo += 1ULL << a.q[i];
    uint64_t off_hi[M];
#pragma unroll
    for (int m = 0; m < M; m++) {
        uint64_t o = 0;
#pragma unroll
        for (int i = LB; i < K; i++) if (((m << LB) >> i) & 1) o += 1ULL << a.q[i];
        off_hi[m] = o;
    }
    cx<T> *g = s + base + off_lo;
#pragma unroll
    for (int m = 0; m < M; m++) lds[tid + m * BS] = g[off_hi[m]];
    __syncthreads();

    for (int op = 0; op < a.n_ops; op++) {
        const TileOp o = a.ops[op];
        const cx<T> *c = a.coef + o.coef;
        switch (o.kind) {
        case TK_DENSE1: {
            const cx<T> m0 = c[0], m1 = c[1], m2 = c[2], m3 = c[3];
            const int b = o.t0;
#pragma unroll
            for (int r = 0; r < M / 2; r++) {
                const int p = tid + r * BS;
                const int i0 = (int)insert_zero(p, b), i1 = i0 | (1 << b);
                const cx<T> x = lds[i0], y = lds[i1];
                lds[i0] = cfma(cmul(m0, x), m1, y);
                lds[i1] = cfma(cmul(m2, x), m3, y);
            }
            break; }
        case TK_DENSE2: {
            cx<T> m[16];
#pragma unroll
            for (int i = 0; i < 16; i++) m[i] = c[i];
            const int b0 = o.t0, b1 = o.t1;
#pragma unroll
            for (int r = 0; r < M / 4; r++) {
                const int p = tid + r * BS;
                const int i = (int)insert_zero(insert_zero(p, b0), b1);
                const int ix[4] = {i, i | (1 << b0), i | (1 << b1), i | (1 << b0) | (1 << b1)};
                cx<T> v[4];
#pragma unroll
                for (int k = 0; k < 4; k++) v[k] = lds[ix[k]];
#pragma unroll
                for (int row = 0; row < 4; row++) {
                    cx<T> acc = cmul(m[4 * row], v[0]);
#pragma unroll
                    for (int k = 1; k < 4; k++) acc = cfma(acc, m[4 * row + k], v[k]);
                    lds[ix[row]] = acc;
                }
            }
            break; }
        case TK_DIAG1: {
            const cx<T> d0 = c[0], d1 = c[1];
            const int b = o.t0;
#pragma unroll
            for (int r = 0; r < M; r++) {
                const int j = tid + r * BS;
                lds[j] = cmul(lds[j], ((j >> b) & 1) ? d1 : d0);
            }
            break; }
        case TK_DIAG2: {
            const cx<T> d0 = c[0], d1 = c[1], d2 = c[2], d3 = c[3];
            const int b0 = o.t0, b1 = o.t1;
#pragma unroll
            for (int r = 0; r < M; r++) {
                const int j = tid + r * BS;
                const bool x = (j >> b0) & 1, y = (j >> b1) & 1;
                const cx<T> d = y ? (x ? d3 : d2) : (x ? d1 : d0);
                lds[j] = cmul(lds[j], d);
            }
            break; }
        case TK_SWAP: {
            const int b0 = o.t0, b1 = o.t1;
#pragma unroll
            for (int r = 0; r < M / 4; r++) {
                const int p = tid + r * BS;
                const int i = (int)insert_zero(insert_zero(p, b0), b1);
                const cx<T> x = lds[i | (1 << b0)], y = lds[i | (1 << b1)];
                lds[i | (1 << b0)] = y;
                lds[i | (1 << b1)] = x;
            }
            break; }
        default:
            break;
        }
        __syncthreads();
    }
#pragma unroll
    for (int m = 0; m < M; m++) g[off_hi[m]] = lds[tid + m * BS];
}

}  // namespace sqk
