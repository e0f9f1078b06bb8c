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
// The tile lives in LDS as two planes of packed bf16 pairs (re | im << 16),
// hi and lo; the pass's store writes hi + lo, which is exact in FP32.
//
// Operand layout of v_mfma_f32_16x16x32_bf16 (lane l, g = l/16, c = l%16;
// measured on gfx950): A lane l element e = A[c][8g+e]; B lane l element
// e = B[8g+e][c]; D lane l register r = D[4g+r][c].
//   * A batch is 16 columns of the tile (a column: fixed values of the
//     non-block tile bits).  Lane l supplies K = 8g..8g+7 of column c: the
//     block amplitudes m = 4g..4g+3, (re, im) each -- four packed words per
//     plane, which are the bf16x8 operand as they lie.
//   * D of row tile t holds rows 16t + 4g + r: amplitudes 8t + 2g and
//     8t + 2g + 1, (re, im) each.
// Every column is read and written by one wave, so a block updates the tile
// in place; the workgroup synchronises between blocks.
//
// LDS words are XOR-swizzled, w ^ ((((w >> 5) ^ (w >> 9)) & 15) << 1): block
// bits at high tile positions no longer put every lane group on the same
// banks.  The map is linear under XOR, so an address is the XOR of the
// column's, the batch's and the block offset's swizzled parts, and bit 0 is
// kept (amplitude pairs stay 8-byte aligned).
#pragma once
#include "common.hpp"

namespace sqk {

constexpr int kFpMinR = 9;          // 2^(r-1) amplitude pairs >= 256 threads
constexpr int kFpMaxR = 13;         // 2^13 amplitudes: 64 KiB of LDS (two planes)
constexpr int kFpMaxExt = 8;        // tile-fixed qubits selecting a block's variant
constexpr int kFpMaxOther = 40;     // qubits outside the tile
constexpr int kFpAmatU4 = 256;      // 16-byte words per variant: 2 row tiles x (hi, lo) x 64 lanes

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float fp2 __attribute__((ext_vector_type(2)));
typedef float fp4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Every field is 32 bits wide: records are read with scalar loads.
struct FpBlock {
    uint32_t a_off;               // 16-byte index of variant 0's A operand; variant v at a_off + 256 v
    int32_t n_ext;                // tile-fixed qubits selecting the variant
    int32_t ext_q[kFpMaxExt];     // variant bit i = bit ext_q[i] of the tile's base index
    int32_t p[4];                 // tile bits of the block's qubits, ascending
};

struct FpArgs {
    const FpBlock *blocks;
    const u32x4 *amat;
    int32_t b_begin, b_end;       // the pass's blocks
    int32_t r;                    // tile bits
    int32_t n_other;              // bits outside the tile: log2 of the workgroup count
    int32_t tile_q[kFpMaxR];      // qubit of tile bit i (tile_q[0..3] = 0..3)
    int32_t other_q[kFpMaxOther]; // qubit of workgroup-index bit k
};

__device__ __forceinline__ uint32_t fp_swz(uint32_t w) { return w ^ ((((w >> 5) ^ (w >> 9)) & 15u) << 1); }

// (re, im) -> packed bf16 hi and lo words.
__device__ __forceinline__ void fp_split(float re, float im, uint32_t &hw, uint32_t &lw) {
    const fp2 v = {re, im};
    hw = __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
    const fp2 h = {__builtin_bit_cast(float, hw << 16), __builtin_bit_cast(float, hw & 0xffff0000u)};
    lw = __builtin_bit_cast(uint32_t, __builtin_convertvector(v - h, bf16x2));
}
__device__ __forceinline__ float fp_lo_half(uint32_t w) { return __builtin_bit_cast(float, w << 16); }
__device__ __forceinline__ float fp_hi_half(uint32_t w) { return __builtin_bit_cast(float, w & 0xffff0000u); }

// Hides x's known low zero bits from the compiler, which would otherwise undo
// the byte scaling folded into the per-lane offsets and shift every LDS
// address again after its XOR.
__device__ __forceinline__ uint32_t fp_opaque(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm volatile("" : "+v"(x));
#endif
    return x;
}

// Tile bits p0 < p1 < p2 < p3 of a block: deposits x over the other tile bits.
__device__ __forceinline__ uint32_t fp_cols(uint32_t x, const int32_t *p) {
    return (uint32_t)insert_zero(insert_zero(insert_zero(insert_zero(x, p[0]), p[1]), p[2]), p[3]);
}
// ... and the block-local index m over the block's bits.
__device__ __forceinline__ uint32_t fp_blk(uint32_t m, const int32_t *p) {
    return ((m & 1u) << p[0]) | (((m >> 1) & 1u) << p[1]) | (((m >> 2) & 1u) << p[2]) | (((m >> 3) & 1u) << p[3]);
}

template <int R>
__global__ void __launch_bounds__(256) k_fpass(cx<float> *__restrict__ s, const FpArgs a) {
    static_assert(R >= kFpMinR && R <= kFpMaxR, "tile bits");
    // static LDS: its address is a constant the compiler folds into every access
    __shared__ __align__(16) uint32_t lds_w[2u << R];
    unsigned char *const lds_raw = reinterpret_cast<unsigned char *>(lds_w);
    constexpr uint32_t kPlaneB = 4u << R;                 // bytes per plane (hi at 0, lo at kPlaneB)
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // the tile's fixed bits (uniform)
    uint64_t base = 0;
    for (int k = 0; k < a.n_other; k++) base |= (uint64_t)((blockIdx.x >> k) & 1u) << a.other_q[k];

    // Load: thread t takes amplitude pairs v = it * 256 + t (tile index 2v).
    // Tile bit i >= 1 is bit i - 1 of v: the thread index covers tile bits
    // 1..8, the iteration the bits above.
    uint64_t toff = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) toff |= (uint64_t)((tid >> k) & 1) << a.tile_q[k + 1];
    constexpr int kIters = 1 << (R - 9);
#pragma unroll
    for (int it = 0; it < kIters; it++) {
        uint64_t ioff = 0;
#pragma unroll
        for (int k = 0; k < R - 9; k++) ioff |= (uint64_t)((it >> k) & 1) << a.tile_q[k + 9];
        const fp4 v = __builtin_nontemporal_load(reinterpret_cast<const fp4 *>(s + base + ioff + toff));
        uint32_t h0, l0, h1, l1;
        fp_split(v.x, v.y, h0, l0);
        fp_split(v.z, v.w, h1, l1);
        const uint32_t wb = fp_swz(2u * (uint32_t)(it * 256 + tid)) * 4u;
        *reinterpret_cast<u32x2 *>(lds_raw + wb) = u32x2{h0, h1};
        *reinterpret_cast<u32x2 *>(lds_raw + kPlaneB + wb) = u32x2{l0, l1};
    }
    __syncthreads();

    const uint32_t g = lane >> 4, c = lane & 15;
    constexpr uint32_t kBatches = 1u << (R - 8);          // 2^(R-4) columns, 16 per batch
    for (int bi = a.b_begin; bi < a.b_end; bi++) {
        const FpBlock *B = a.blocks + bi;
        int32_t p[4] = {B->p[0], B->p[1], B->p[2], B->p[3]};
        uint32_t var = 0;
        const int n_ext = B->n_ext;
        for (int k = 0; k < n_ext; k++) var |= (uint32_t)((base >> B->ext_q[k]) & 1u) << k;
        const u32x4 *am = a.amat + B->a_off + (size_t)var * kFpAmatU4;
        const bf16x8 A0h = __builtin_bit_cast(bf16x8, am[lane]);
        const bf16x8 A0l = __builtin_bit_cast(bf16x8, am[64 + lane]);
        const bf16x8 A1h = __builtin_bit_cast(bf16x8, am[128 + lane]);
        const bf16x8 A1l = __builtin_bit_cast(bf16x8, am[192 + lane]);
        // byte offsets: column c of batch 0 with the block amplitude folded in
        const uint32_t col = fp_swz(fp_cols(c, p));
        uint32_t rd[4], wr[4];
#pragma unroll
        for (uint32_t t = 0; t < 4; t++) rd[t] = fp_opaque((col ^ fp_swz(fp_blk(4u * g + t, p))) * 4u);
#pragma unroll
        for (uint32_t t = 0; t < 2; t++) {
            wr[2 * t] = fp_opaque((col ^ fp_swz(fp_blk(8u * t + 2u * g, p))) * 4u);
            wr[2 * t + 1] = fp_opaque((col ^ fp_swz(fp_blk(8u * t + 2u * g + 1u, p))) * 4u);
        }
        for (uint32_t b = wave; b < kBatches; b += 4) {
            const uint32_t ub = fp_swz(fp_cols(b << 4, p)) * 4u;   // uniform
            u32x4 xh, xl;
            xh.x = *reinterpret_cast<const uint32_t *>(lds_raw + (rd[0] ^ ub));
            xh.y = *reinterpret_cast<const uint32_t *>(lds_raw + (rd[1] ^ ub));
            xh.z = *reinterpret_cast<const uint32_t *>(lds_raw + (rd[2] ^ ub));
            xh.w = *reinterpret_cast<const uint32_t *>(lds_raw + (rd[3] ^ ub));
            xl.x = *reinterpret_cast<const uint32_t *>(lds_raw + kPlaneB + (rd[0] ^ ub));
            xl.y = *reinterpret_cast<const uint32_t *>(lds_raw + kPlaneB + (rd[1] ^ ub));
            xl.z = *reinterpret_cast<const uint32_t *>(lds_raw + kPlaneB + (rd[2] ^ ub));
            xl.w = *reinterpret_cast<const uint32_t *>(lds_raw + kPlaneB + (rd[3] ^ ub));
            const bf16x8 Bh = __builtin_bit_cast(bf16x8, xh), Bl = __builtin_bit_cast(bf16x8, xl);
            fp4 d0 = {0, 0, 0, 0}, d1 = {0, 0, 0, 0};
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0h, Bl, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1h, Bl, d1, 0, 0, 0);
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0l, Bh, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1l, Bh, d1, 0, 0, 0);
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A0h, Bh, d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A1h, Bh, d1, 0, 0, 0);
            uint32_t h, l;
            fp_split(d0.x, d0.y, h, l);
            *reinterpret_cast<uint32_t *>(lds_raw + (wr[0] ^ ub)) = h;
            *reinterpret_cast<uint32_t *>(lds_raw + kPlaneB + (wr[0] ^ ub)) = l;
            fp_split(d0.z, d0.w, h, l);
            *reinterpret_cast<uint32_t *>(lds_raw + (wr[1] ^ ub)) = h;
            *reinterpret_cast<uint32_t *>(lds_raw + kPlaneB + (wr[1] ^ ub)) = l;
            fp_split(d1.x, d1.y, h, l);
            *reinterpret_cast<uint32_t *>(lds_raw + (wr[2] ^ ub)) = h;
            *reinterpret_cast<uint32_t *>(lds_raw + kPlaneB + (wr[2] ^ ub)) = l;
            fp_split(d1.z, d1.w, h, l);
            *reinterpret_cast<uint32_t *>(lds_raw + (wr[3] ^ ub)) = h;
            *reinterpret_cast<uint32_t *>(lds_raw + kPlaneB + (wr[3] ^ ub)) = l;
        }
        __syncthreads();
    }

    // Store: hi + lo of each amplitude pair back to the state.
#pragma unroll
    for (int it = 0; it < kIters; it++) {
        uint64_t ioff = 0;
#pragma unroll
        for (int k = 0; k < R - 9; k++) ioff |= (uint64_t)((it >> k) & 1) << a.tile_q[k + 9];
        const uint32_t wb = fp_swz(2u * (uint32_t)(it * 256 + tid)) * 4u;
        const u32x2 h = *reinterpret_cast<const u32x2 *>(lds_raw + wb);
        const u32x2 l = *reinterpret_cast<const u32x2 *>(lds_raw + kPlaneB + wb);
        fp4 v;
        v.x = fp_lo_half(h.x) + fp_lo_half(l.x);
        v.y = fp_hi_half(h.x) + fp_hi_half(l.x);
        v.z = fp_lo_half(h.y) + fp_lo_half(l.y);
        v.w = fp_hi_half(h.y) + fp_hi_half(l.y);
        __builtin_nontemporal_store(v, reinterpret_cast<fp4 *>(s + base + ioff + toff));
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
