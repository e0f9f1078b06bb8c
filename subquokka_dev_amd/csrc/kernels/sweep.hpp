// sweep.hpp -- one-pass gate kernels for the plain gate-by-gate sweep
// (the GPU form of MEM_Runner::run(vector<Gate*>&), MEM_Runner.cpp:101-134,
// and of the loop shapes chunk_gate_mem ... nonchunk_nonchunk_gate_mem,
// circuit.cpp:143-218).  Each launch reads and writes the amplitudes a gate
// changes exactly once.
#pragma once
#include "common.hpp"

namespace sqk {

// ---------------------------------------------------------------------------
// Dense k-qubit gate (k = KV + KL + KH <= 2).
//   KV: targets inside the 16-byte vector (only bit 0 of a complex<float> state)
//   KL: targets inside the lane bits [VB, VB+6)        -> __shfl_xor exchange
//   KH: targets at or above VB+6                      -> register index
// Local index of an amplitude: vector bit | lane bits << KV | register << (KV+KL),
// which is ascending target order because targets are sorted.
// ---------------------------------------------------------------------------
template <typename T> struct DenseArgs {
    cx<T> m[16];        // row-major 2^k x 2^k
    uint64_t n_tiles;   // wave tiles = 2^(n - S - KH)
    int hpos[2];        // register-target bit positions, ascending
    int lmask[2];       // lane masks 1 << (t - VB) of lane targets, ascending
};

template <typename T, int KV, int KL, int KH, int U>
__global__ void __launch_bounds__(kBlock) k_dense(cx<T> *__restrict__ s, const DenseArgs<T> a) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    constexpr int K = KV + KL + KH;
    constexpr int D = 1 << K;
    constexpr int R = 1 << KH;
    constexpr int NL = 1 << KL;
    constexpr int E = 1 << VB;
    const int lane = threadIdx.x & 63;
    const uint64_t wave = (uint64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint64_t n_waves = (uint64_t)gridDim.x * (kBlock / 64);
    int c = 0;   // this lane's lane-target bits
#pragma unroll
    for (int j = 0; j < KL; j++) c |= ((lane & a.lmask[j]) ? 1 : 0) << j;

    // U consecutive wave tiles per iteration: all loads issue before any use
    for (uint64_t t0 = wave * U; t0 < a.n_tiles; t0 += n_waves * U) {
        uint64_t base[U];
        Pack<T> reg[U][R];
#pragma unroll
        for (int u = 0; u < U; u++) {
            base[u] = (t0 + u) << S;
#pragma unroll
            for (int j = 0; j < KH; j++) base[u] = insert_zero(base[u], a.hpos[j]);
            base[u] += (uint64_t)lane << VB;
            if (t0 + u >= a.n_tiles) continue;   // uniform per wave
#pragma unroll
            for (int r = 0; r < R; r++) {
                uint64_t off = base[u];
#pragma unroll
                for (int j = 0; j < KH; j++) if ((r >> j) & 1) off += 1ULL << a.hpos[j];
                reg[u][r] = ld(s + off);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (t0 + u >= a.n_tiles) continue;
            // sh[d][r]: register r of the lane that differs from this one by the
            // lane targets selected in d.
            Pack<T> sh[NL][R];
#pragma unroll
            for (int d = 0; d < NL; d++) {
                int mask = 0;
#pragma unroll
                for (int j = 0; j < KL; j++) if ((d >> j) & 1) mask |= a.lmask[j];
#pragma unroll
                for (int r = 0; r < R; r++) sh[d][r] = d == 0 ? reg[u][r] : shfl_xor_pack(reg[u][r], mask);
            }
#pragma unroll
            for (int r = 0; r < R; r++) {
                Pack<T> out;
#pragma unroll
                for (int e = 0; e < E; e++) {
                    cx<T> res = {0, 0};
#pragma unroll
                    for (int cc = 0; cc < NL; cc++) {   // candidate lane bits, selected below
                        const int row = (KV ? (e & 1) : 0) | (cc << KV) | (r << (KV + KL));
                        cx<T> acc = {0, 0};
#pragma unroll
                        for (int b = 0; b < D; b++) {
                            const int bv = KV ? (b & 1) : 0;
                            const int bc = (b >> KV) & (NL - 1);
                            const int br = b >> (KV + KL);
                            const int ee = KV ? ((e & ~1) | bv) : e;
                            acc = cfma(acc, a.m[row * D + b], sh[cc ^ bc][br].e[ee]);
                        }
                        if (cc == c) res = acc;
                    }
                    out.e[e] = res;
                }
                uint64_t off = base[u];
#pragma unroll
                for (int j = 0; j < KH; j++) if ((r >> j) & 1) off += 1ULL << a.hpos[j];
                st(s + off, out);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Scalar fallback for states smaller than one wave tile: one thread per
// k-qubit group, bit-inserted addressing (bit_string, MEM_Runner.cpp:90-99).
// ---------------------------------------------------------------------------
template <typename T> struct SmallArgs {
    cx<T> m[16];
    uint64_t n_groups;
    int t[2];
    int k;
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_dense_small(cx<T> *__restrict__ s, const SmallArgs<T> a) {
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g >= a.n_groups) return;
    uint64_t base = g;
    for (int j = 0; j < a.k; j++) base = insert_zero(base, a.t[j]);
    const int D = 1 << a.k;
    cx<T> in[4], out[4];
    for (int b = 0; b < D; b++) {
        uint64_t off = base;
        for (int j = 0; j < a.k; j++) if ((b >> j) & 1) off += 1ULL << a.t[j];
        in[b] = s[off];
    }
    for (int r = 0; r < D; r++) {
        cx<T> acc = {0, 0};
        for (int b = 0; b < D; b++) acc = cfma(acc, a.m[r * D + b], in[b]);
        out[r] = acc;
    }
    for (int b = 0; b < D; b++) {
        uint64_t off = base;
        for (int j = 0; j < a.k; j++) if ((b >> j) & 1) off += 1ULL << a.t[j];
        s[off] = out[b];
    }
}

// ---------------------------------------------------------------------------
// Diagonal gate (Z, P, RZ, CP, RZZ).  Targets whose "0" half is the identity
// (diag entries equal to 1) are fixed to 1 in the index space, so P/Z touch
// half of the state and CP a quarter, as the reference's Z/P/CP loops do
// (circuit.cpp:2640-2679, 2960-2982).  Every fixed target is >= VB.
// ---------------------------------------------------------------------------
template <typename T> struct DiagArgs {
    cx<T> d[4];          // diagonal in local-index order (bit j <-> t[j])
    uint64_t n_vec;      // 16-byte vectors to visit
    int t[2];
    int k;
    int fixed[2];        // ascending positions forced to 1
    int nfixed;
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_diag(cx<T> *__restrict__ s, const DiagArgs<T> a) {
    constexpr int VB = Traits<T>::VB;
    constexpr int E = 1 << VB;
    constexpr int U = 2;   // vectors per thread per iteration, one grid apart
    const uint64_t span = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v0 = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v0 < a.n_vec; v0 += span * U) {
        uint64_t i[U];
        Pack<T> p[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const uint64_t v = v0 + u * span;
            i[u] = v << VB;
            for (int j = 0; j < a.nfixed; j++) i[u] = insert_one(i[u], a.fixed[j]);
            if (v < a.n_vec) p[u] = ld(s + i[u]);
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (v0 + u * span >= a.n_vec) continue;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const uint64_t idx = i[u] + e;
                int l = 0;
                for (int j = 0; j < a.k; j++) l |= (int)((idx >> a.t[j]) & 1) << j;
                p[u].e[e] = cmul(p[u].e[e], a.d[l]);
            }
            st(s + i[u], p[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// Qubit-swap permutation: NP disjoint pairs (p[2j] <-> p[2j+1]), every
// position >= VB.  SWAP / VSWAP_1_1 (NP = 1, half the state moves) and
// VSWAP_2_2 (NP = 2, 12 of every 16 amplitudes move), circuit.cpp:351-355,
// 3457-3556.  Only moving amplitudes are read and written.
// ---------------------------------------------------------------------------
struct PermArgs {
    uint64_t n_vec;      // 16-byte vectors per group-slot = 2^(n - VB - 2*NP)
    int p[4];            // pair positions
    int sorted[4];       // the same positions ascending (for bit insertion)
};

template <typename T, int NP>
__global__ void __launch_bounds__(kBlock) k_perm(cx<T> *__restrict__ s, const PermArgs a) {
    constexpr int VB = Traits<T>::VB;
    constexpr int Q = 2 * NP;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t v = (uint64_t)blockIdx.x * kBlock + threadIdx.x; v < a.n_vec; v += stride) {
        uint64_t base = v << VB;
#pragma unroll
        for (int j = 0; j < Q; j++) base = insert_zero(base, a.sorted[j]);
        Pack<T> val[1 << Q];
        // local index bit 2j <-> p[2j], bit 2j+1 <-> p[2j+1]
#pragma unroll
        for (int x = 0; x < (1 << Q); x++) {
            bool moves = false;
#pragma unroll
            for (int j = 0; j < NP; j++) moves |= ((x >> (2 * j)) & 1) != ((x >> (2 * j + 1)) & 1);
            if (!moves) continue;
            uint64_t off = base;
#pragma unroll
            for (int j = 0; j < Q; j++) if ((x >> j) & 1) off += 1ULL << a.p[j];
            val[x] = ld(s + off);
        }
#pragma unroll
        for (int x = 0; x < (1 << Q); x++) {
            bool moves = false;
            int y = 0;   // source of destination x: swap the two bits of every pair
#pragma unroll
            for (int j = 0; j < NP; j++) {
                const int b0 = (x >> (2 * j)) & 1, b1 = (x >> (2 * j + 1)) & 1;
                moves |= b0 != b1;
                y |= (b1 << (2 * j)) | (b0 << (2 * j + 1));
            }
            if (!moves) continue;
            uint64_t off = base;
#pragma unroll
            for (int j = 0; j < Q; j++) if ((x >> j) & 1) off += 1ULL << a.p[j];
            st(s + off, val[y]);
        }
    }
}

}  // namespace sqk
