// fpass.hpp -- a plain gate list as fused passes over a complex64 state.
//
// The planner cuts the list into passes whose dense gates act inside a set R
// of r qubits (qubits 0-3 and r - 4 chosen ones); diagonal gates may act on
// any qubits.  A workgroup loads one tile -- the 2^r amplitudes that share
// the values of every bit outside R -- into LDS, applies the pass's
// four-qubit blocks one after another on the matrix cores and writes the tile
// back: one read and one write of the state per pass, however many blocks
// (the role the reference gives its cache-resident subcircuits,
// MEM_Runner.cpp:173-182, with the finder's blocks, finder.cpp:629-698).
//
// A block is the FP64 product of its gates (k_compose): a 16x16 complex U on
// four tile bits, applied as the real 32x32 matrix M whose rows and columns
// are ordered (amplitude m, component c) -> 2m + c:
//   M[2m+c][2m'+c'] = Re U[m][m'] if c == c';  -Im U[m][m'] if c = 0, c' = 1;
//                     Im U[m][m'] if c = 1, c' = 0.
// A diagonal gate with qubits outside R is, inside one tile, a diagonal gate
// on its qubits in R (or a scalar) whose entries depend on the tile's fixed
// bits.  The block holding it is composed once per value of those bits (a
// variant); each workgroup picks its variant from its tile's bits.
//
// Products: v_mfma_f32_16x16x32_bf16 on split operands.  x = hi + lo with
// hi = bf16(x) and lo = bf16(x - hi) (16 significant bits), and
//   M x ~ Mhi xlo + Mlo xhi + Mhi xhi,  accumulated in FP32
// (the dropped Mlo xlo term and the splits are ~2^-17 relative per block).
// The tile stays in FP32 in LDS; only the MFMA operands are split.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (lane l, g = l/16, c = l%16;
// measured on gfx950): A lane l element e = A[c][8g+e]; B lane l element
// e = B[8g+e][c]; D lane l register r = D[4g+r][c].
//   * A batch is 16 columns of the tile (a column: fixed values of the
//     non-block tile bits).  Lane l supplies K = 8g..8g+7 of column c: the
//     block amplitudes m = 4g..4g+3, (re, im) each; their split halves
//     packed in pairs are the bf16x8 operands.
//   * D of row tile t holds rows 16t + 4g + r: amplitudes 8t + 2g and
//     8t + 2g + 1, (re, im) each -- written back as they lie.
// Every column is read and written by one wave, so a block updates the tile
// in place; the workgroup synchronises between blocks.
//
// LDS slots (8 bytes, one amplitude) are XOR-swizzled:
//   s(a) = a ^ ((a >> 5) & 31) ^ ((a >> 10) & 31),
// which folds every tile bit onto the five low slot bits (the bank pair of
// 64 banks).  The map is linear under XOR, so an address is the XOR of the
// column's, the block amplitude's and the batch's swizzled parts, all
// computed on the host.  Per block, the four tile bits that form a batch's
// 16 columns are chosen so that their bank-pair bits are independent and
// stay independent with the block bit that separates lanes 0-15 from 16-31
// (block bit 2 for the reads, bit 1 for the writes): 32 lanes, 32 distinct
// bank pairs, wherever the block's own bits lie.
#pragma once
#include "common.hpp"

namespace sqk {

constexpr int kFpMinR = 9;          // 2^(r-1) amplitude pairs >= the threads of a workgroup
constexpr int kFpMaxR = 13;         // 2^13 amplitudes: 64 KiB of LDS
constexpr int kFpMaxExt = 8;        // tile-fixed qubits selecting a block's variant
constexpr int kFpMaxOther = 40;     // qubits outside the tile
constexpr int kFpAmatU4 = 256;      // 16-byte words per variant: 2 row tiles x (hi, lo) x 64 lanes
constexpr int kFpMaxBatches = 1 << (kFpMaxR - 8);
// Waves per workgroup: 8 for the 64 KiB tiles (two workgroups per CU), 4 below.
__host__ __device__ constexpr int fp_waves(int r) { return r >= 13 ? 8 : 4; }

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float fp2 __attribute__((ext_vector_type(2)));
typedef float fp4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Swizzled LDS slot of tile amplitude a (host and device).
__host__ __device__ inline uint32_t fp_slot(uint32_t a) {
    return a ^ ((a >> 5) & 31u) ^ ((a >> 10) & 31u);
}

// Every field is 32 bits wide: records are read with scalar loads (the
// uniform fields) and vector loads (colw, blkw: one entry per lane).
struct FpBlock {
    uint32_t a_off;               // 16-byte index of variant 0's A operand; variant v at a_off + 256 v
    int32_t n_ext;                // tile-fixed qubits selecting the variant
    int32_t ext_q[kFpMaxExt];     // variant bit i = bit ext_q[i] of the tile's base index
    uint32_t colw[16];            // byte offset (swizzled) of column c's tile bits
    uint32_t blkw[16];            // byte offset (swizzled) of block amplitude m's tile bits
    uint32_t ubw[kFpMaxBatches];  // byte offset (swizzled) of batch b's tile bits, stored per wave:
                                  // wave w's batches w, w + W, ... (W waves) at ubw[w * ceil(batches / W) + i]
};

struct FpArgs {
    const FpBlock *blocks;
    const u32x4 *amat;
    int32_t b_begin, b_end;       // the pass's blocks
    int32_t n_other;              // bits outside the tile: log2 of the workgroup count
    int32_t tile_q[kFpMaxR];      // qubit of tile bit i (tile_q[0..3] = 0..3)
    int32_t other_q[kFpMaxOther]; // qubit of workgroup-index bit k
};

// (re, im) -> packed bf16 hi and lo words.
__device__ __forceinline__ void fp_split(fp2 v, uint32_t &hw, uint32_t &lw) {
    hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    const fp2 h = {__builtin_bit_cast(float, hw << 16), __builtin_bit_cast(float, hw & 0xffff0000u)};
    lw = __builtin_bit_cast(uint32_t, __builtin_convertvector(v - h, bf16x2));
}

template <int R>
__global__ void __launch_bounds__(64 * fp_waves(R)) k_fpass(cx<float> *__restrict__ s, const FpArgs a) {
    static_assert(R >= kFpMinR && R <= kFpMaxR, "tile bits");
    constexpr int NW = fp_waves(R), TB = NW == 8 ? 9 : 8;   // threads = 2^TB
    // static LDS: its address is a constant the compiler folds into every access
    __shared__ __align__(16) fp2 lds[1u << R];
    unsigned char *const L = reinterpret_cast<unsigned char *>(lds);
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // Workgroups are dealt to the 8 XCDs in turn; renumbered, each XCD takes
    // one contiguous eighth of the tiles (neighbouring tiles share DRAM pages
    // and that XCD's L2).  Then the tile's fixed bits (uniform).
    uint32_t wg = blockIdx.x;
    if ((gridDim.x & 7) == 0) wg = (wg & 7) * (gridDim.x >> 3) + (wg >> 3);
    uint64_t base = 0;
    for (int k = 0; k < a.n_other; k++) base |= (uint64_t)((wg >> k) & 1u) << a.other_q[k];

    // Load: thread t takes amplitude pairs v = it * 2^TB + t (tile index 2v).
    // Tile bit i >= 1 is bit i - 1 of v: the thread index covers tile bits
    // 1..TB, the iteration the bits above.
    uint64_t toff = 0;
#pragma unroll
    for (int k = 0; k < TB; k++) toff |= (uint64_t)((tid >> k) & 1) << a.tile_q[k + 1];
    const uint32_t w0 = fp_slot(2u * (uint32_t)tid) * 8u;   // slot of amplitude 2t
    constexpr int kIters = 1 << (R - 1 - TB);
#pragma unroll
    for (int it = 0; it < kIters; it++) {
        uint64_t ioff = 0;
#pragma unroll
        for (int k = 0; k < R - 1 - TB; k++) ioff |= (uint64_t)((it >> k) & 1) << a.tile_q[k + 1 + TB];
        const fp4 v = __builtin_nontemporal_load(reinterpret_cast<const fp4 *>(s + base + ioff + toff));
        const uint32_t w = w0 ^ (fp_slot((2u << TB) * it) * 8u);
        *reinterpret_cast<fp2 *>(L + w) = fp2{v.x, v.y};
        *reinterpret_cast<fp2 *>(L + (w ^ 8u)) = fp2{v.z, v.w};   // s(2v + 1) = s(2v) ^ 1
    }
    __syncthreads();

    const uint32_t g = lane >> 4, c = lane & 15;
    constexpr uint32_t kBatches = 1u << (R - 8);          // 2^(R-4) columns, 16 per batch
    constexpr uint32_t kPerWave = (kBatches + NW - 1) / NW;   // R = 9: two batches, waves 2 and 3 idle
    // A operands and per-lane offsets of block bi (this tile's variant); the
    // next block's are loaded while the current one runs
    struct Next {
        u32x4 am[4];
        uint32_t col;
        u32x4 rd;
        u32x2 w0, w1;
    };
    auto fetch = [&](int bi, Next &n) {
        const FpBlock *B = a.blocks + bi;
        uint32_t var = 0;
        const int n_ext = B->n_ext;
        for (int k = 0; k < n_ext; k++) var |= (uint32_t)((base >> B->ext_q[k]) & 1u) << k;
        const u32x4 *am = a.amat + B->a_off + (size_t)var * kFpAmatU4;
        n.am[0] = am[lane];
        n.am[1] = am[64 + lane];
        n.am[2] = am[128 + lane];
        n.am[3] = am[192 + lane];
        n.col = B->colw[c];
        n.rd = *reinterpret_cast<const u32x4 *>(B->blkw + 4 * g);
        n.w0 = *reinterpret_cast<const u32x2 *>(B->blkw + 2 * g);
        n.w1 = *reinterpret_cast<const u32x2 *>(B->blkw + 8 + 2 * g);
    };
    Next nx;
    if (a.b_begin < a.b_end) fetch(a.b_begin, nx);
    for (int bi = a.b_begin; bi < a.b_end; bi++) {
        const FpBlock *B = a.blocks + bi;
        const bf16x8 A0h = __builtin_bit_cast(bf16x8, nx.am[0]);
        const bf16x8 A0l = __builtin_bit_cast(bf16x8, nx.am[1]);
        const bf16x8 A1h = __builtin_bit_cast(bf16x8, nx.am[2]);
        const bf16x8 A1l = __builtin_bit_cast(bf16x8, nx.am[3]);
        const uint32_t rd0 = nx.col ^ nx.rd.x, rd1 = nx.col ^ nx.rd.y, rd2 = nx.col ^ nx.rd.z, rd3 = nx.col ^ nx.rd.w;
        const uint32_t wr0 = nx.col ^ nx.w0.x, wr1 = nx.col ^ nx.w0.y, wr2 = nx.col ^ nx.w1.x, wr3 = nx.col ^ nx.w1.y;
        if (bi + 1 < a.b_end) fetch(bi + 1, nx);
        const uint32_t *ub = B->ubw + wave * kPerWave;
#pragma unroll
        for (uint32_t i = 0; i < kPerWave; i++) {
            if (kBatches < NW && (uint32_t)wave >= kBatches) break;   // uniform
            const uint32_t u = ub[i];   // uniform
            const fp2 x0 = *reinterpret_cast<const fp2 *>(L + (rd0 ^ u));
            const fp2 x1 = *reinterpret_cast<const fp2 *>(L + (rd1 ^ u));
            const fp2 x2 = *reinterpret_cast<const fp2 *>(L + (rd2 ^ u));
            const fp2 x3 = *reinterpret_cast<const fp2 *>(L + (rd3 ^ u));
            uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
            fp_split(x0, h0, l0);
            fp_split(x1, h1, l1);
            fp_split(x2, h2, l2);
            fp_split(x3, h3, l3);
            const u32x4 xh = {h0, h1, h2, h3}, xl = {l0, l1, l2, l3};
            const bf16x8 Bh = __builtin_bit_cast(bf16x8, xh), Bl = __builtin_bit_cast(bf16x8, xl);
            fp4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0h, Bl, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1h, Bl, d1, 0, 0, 0);
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0l, Bh, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1l, Bh, d1, 0, 0, 0);
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0h, Bh, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1h, Bh, d1, 0, 0, 0);
            *reinterpret_cast<fp2 *>(L + (wr0 ^ u)) = fp2{d0.x, d0.y};
            *reinterpret_cast<fp2 *>(L + (wr1 ^ u)) = fp2{d0.z, d0.w};
            *reinterpret_cast<fp2 *>(L + (wr2 ^ u)) = fp2{d1.x, d1.y};
            *reinterpret_cast<fp2 *>(L + (wr3 ^ u)) = fp2{d1.z, d1.w};
        }
        __syncthreads();
    }

    // Store the tile back.
#pragma unroll
    for (int it = 0; it < kIters; it++) {
        uint64_t ioff = 0;
#pragma unroll
        for (int k = 0; k < R - 1 - TB; k++) ioff |= (uint64_t)((it >> k) & 1) << a.tile_q[k + 1 + TB];
        const uint32_t w = w0 ^ (fp_slot((2u << TB) * it) * 8u);
        const fp2 p0 = *reinterpret_cast<const fp2 *>(L + w);
        const fp2 p1 = *reinterpret_cast<const fp2 *>(L + (w ^ 8u));
        __builtin_nontemporal_store(fp4{p0.x, p0.y, p1.x, p1.y}, reinterpret_cast<fp4 *>(s + base + ioff + toff));
    }
}

// A operands of each composed block variant: 16x16 complex U (FP64, row-major,
// stride 1024 elements per variant, as k_compose writes) -> the real M above
// in the MFMA A layout, split into bf16 hi and lo.  One thread per
// (variant, row tile t, hi/lo, lane): out[v*256 + (2t + hl)*64 + lane].
__global__ void __launch_bounds__(256) k_fpass_amat(const cx<double> *__restrict__ u, u32x4 *__restrict__ out,
                                                     uint32_t n_var) {
    const uint32_t id = blockIdx.x * 256 + threadIdx.x;
    if (id >= n_var * kFpAmatU4) return;
    const uint32_t v = id / kFpAmatU4, rem = id % kFpAmatU4;
    const uint32_t t = rem / 128, hl = (rem / 64) & 1, lane = rem % 64;
    const uint32_t g = lane / 16, c = lane % 16;
    const cx<double> *U = u + (size_t)v * 1024;
    const uint32_t row = 16 * t + c, m = row / 2, cr = row % 2;
    uint32_t w[4];
    for (int e2 = 0; e2 < 4; e2++) {
        uint32_t half[2];
        for (int h = 0; h < 2; h++) {
            const uint32_t k = 8 * g + 2 * e2 + h, mp = k / 2, ck = k % 2;
            const cx<double> z = U[m * 16 + mp];
            const double x = cr == ck ? z.re : (cr ? z.im : -z.im);
            const __bf16 hi = (__bf16)(float)x;
            const __bf16 lo = (__bf16)(float)(x - (double)(float)hi);
            half[h] = (uint32_t)__builtin_bit_cast(uint16_t, hl ? lo : hi);
        }
        w[e2] = half[0] | (half[1] << 16);
    }
    out[id] = u32x4{w[0], w[1], w[2], w[3]};
}

}  // namespace sqk
