// pass.hpp -- register multi-gate pass for the plain sweep.
//
// MEM_Runner::run(vector<Gate*>&) (MEM_Runner.cpp:101-134) reads and writes
// the whole state once per gate.  A register pass applies a run of gates to
// the state in one read and one write: each wave holds a "wave tile" of
// 2^KH x 64 16-byte vectors in registers -- index bits [0, VB) inside the
// vector, bits [VB, VB+6) across the lanes (the top three of them may stand
// for other qubits; PassArgs::lh) and KH chosen qubits hpos[] as the register
// index -- and applies every gate of the pass in order before storing the
// tile back.
//
// Gates a pass can hold:
//   * dense one-qubit gates on a vector, lane or register bit (H X Y U1 RX RY);
//   * diagonal one- and two-qubit gates on ANY qubits (Z P RZ CP RZZ): the
//     phase depends only on the amplitude's global index, which every lane
//     knows.
// Each gate is its own 2x2 (or diagonal) step in FP arithmetic of the state's
// type -- nothing is composed except a controlled phase folded into the next
// dense gate on one of its qubits (its phase scales two coefficients), and a
// dense gate with complex coefficients may arrive split into a real rotation
// and one-qubit phases that the host collects into PO_PHASES ops (the product
// is the gate; only the rounding differs) -- so a pass agrees with the gates
// applied one launch at a time up to the rounding of individual products.
// Per-op work
// is shaped so that nothing is selected per amplitude: lane-gate
// coefficients are chosen once per lane, shuffle masks are constants, and a
// diagonal factor is chosen per lane, per register at compile time, or once
// per wave (identity entries then skip the op).
#pragma once
#include "common.hpp"

namespace sqk {

enum PassOpKind : int32_t {
    PO_DENSE_REG = 0,    // a = register slot of the target; c = 15 when all four coefficients
                         //   are real, 1 when m0 is, else 0 (the kernel leaves out zero terms)
    PO_DENSE_LANE = 1,   // a = lane bit of the target (0..5)
    PO_DENSE_VEC = 2,    // target inside the 16-byte vector (complex<float>, bit 0)
    PO_DIAG_FREE = 3,    // diagonal, no bit in a register slot: a, b = global bits (b < 0: one-qubit);
                         //   c = 1 when both bits are wave-uniform (tile bits), so identity
                         //   entries skip the whole op
    PO_DIAG_REG1 = 4,    // diagonal, one bit in register slot a, the other at global bit b (< 0: none);
                         //   coefficients e[(register bit << 1) | other bit]; c = 1 when the other
                         //   bit is a tile bit or absent (wave-uniform)
    PO_DIAG_REG2 = 5,    // diagonal, both bits in register slots a < b
    PO_DENSE_REG_C = 6,  // complex<double> only: a controlled phase folded into the dense
                         //   gate on register slot a that follows it: where the control
                         //   bit is 1 the gate is U diag(1, ph), i.e. m1 and m3 times ph
                         //   (coef[4]).  b = the control's position; active = 0: register
                         //   slot b, 1: lane bit b - VB, 2: tile bit b.  c as PO_DENSE_REG
                         //   (1: m0 real)
    PO_DENSE_LANE_C = 7, // complex<double> only: the same fold into a dense gate on lane bit a;
                         //   b, active and coef[4] as PO_DENSE_REG_C
    PO_PHASES = 8,       // a product of one-qubit phases: amplitude i times L(i) T(i), where
                         //   L = coef[0] * prod over virtual positions p < VB + 6 with bit p of
                         //   a set of (bit p of i ? z_p : 1), and T = t[register index] when
                         //   b = 1 (else 1).  c coefficient records (8 coefficients each, no
                         //   header) follow the op: when a != 0, one with z_p in slot p; when
                         //   b = 1, the 2^KH table t in the next ceil(2^KH / 8)
    PO_ROT_REG = 9,      // the rotation [[c, -s], [s, c]] (c, s >= 0) on register slot a (or,
                         //   with c = 1, inside the complex64 vector) as three shears:
                         //   x += k y, y += m x, x += k y, k = coef[0].re = -s / (1 + c),
                         //   m = coef[0].im = s; each an FMA into its own input's registers
    PO_ROT_LANE = 10,    // the same rotation on lane bit a.  Lane bits 0-3, scaled by the host
                         //   (the scale goes into a later PO_PHASES scalar): c = 0:
                         //   [[1, -k], [k, 1]] with k = s / c, c = 1: [[k, -1], [1, k]] with
                         //   k = c / s, k = coef[0].re.  Lane bits 4-5: shears as PO_ROT_REG
    PO_SIGN = 11,        // amplitude i negated where every bit of i named below is 1 (CP(pi),
                         //   CZ): a = mask of virtual positions < VB + 6, b = mask of register
                         //   slots, pad[0], pad[1] = tile positions (-1: none)
};
// Dense ops (PO_DENSE_REG/_C, PO_DENSE_LANE/_C, PO_DENSE_VEC) with c = 15 have
// four real coefficients (after a fold, also the folded ones): the kernel
// leaves out the zero imaginary terms.

// 32-bit fields only: the op loop reads records with scalar (SMEM) loads,
// which ignore the two low bits of their address components; byte fields
// beside a dword field let the compiler address it from a base 2 bytes off
// alignment, and the load then returns the byte fields instead.
// The four coefficients travel inside the record, so one scalar fetch brings
// the whole op (the offset-then-load of a separate array costs a second
// dependent round trip per op).
template <typename T> struct alignas(sizeof(T) == 8 ? 128 : 64) PassOp {
    int32_t kind;
    int32_t a, b, c;
    uint32_t active;     // diagonal: bit k set when coefficient k is not 1
    int32_t pad[3];
    cx<T> coef[sizeof(T) == 8 ? 5 : 4];   // complex<double>: coef[4] = PO_DENSE_REG_C's phase
};
typedef int pass_v16i __attribute__((ext_vector_type(16)));

// Address spaces of the device build: 1 is global memory, 4 constant memory.
// The host build of these kernels (the CPU emulation) has one flat space.
#if defined(__HIP_DEVICE_COMPILE__)
#define SQK_AS_GLOBAL __attribute__((address_space(1)))
#define SQK_AS_CONST __attribute__((address_space(4)))
#else
#define SQK_AS_GLOBAL
#define SQK_AS_CONST
#endif
typedef SQK_AS_GLOBAL v4f pass_gv4f;

// The whole record as dwords, read with one batch of aligned scalar loads.
template <typename T> struct PassRec {
    static constexpr int NV = sizeof(PassOp<T>) / 64;
    pass_v16i w[NV];
    __device__ explicit PassRec(const PassOp<T> *p) {
        // constant memory: the program is read-only while a pass runs, so the
        // records stay scalar loads whatever the compiler can prove about the
        // state stores (which go through addresses rebuilt from integers)
        typedef SQK_AS_CONST const pass_v16i cpv;
        cpv *q = (cpv *)p;
#pragma unroll
        for (int i = 0; i < NV; i++) w[i] = q[i];
    }
    __device__ int dw(int i) const { return w[i / 16][i % 16]; }
    // coefficient j of an op record (after its eight header dwords)
    __device__ cx<T> coef(int j) const { return at((sizeof(T) == 8 ? 2 : 4) + j); }
    // coefficient j of a coefficient record (no header), j < 8
    __device__ cx<T> xcoef(int j) const { return at(j); }
    __device__ cx<T> at(int j) const {
        if constexpr (sizeof(T) == 8) {
            const int b = 4 * j;
            const uint64_t r = (uint64_t)(uint32_t)dw(b) | ((uint64_t)(uint32_t)dw(b + 1) << 32);
            const uint64_t m = (uint64_t)(uint32_t)dw(b + 2) | ((uint64_t)(uint32_t)dw(b + 3) << 32);
            return {__builtin_bit_cast(double, r), __builtin_bit_cast(double, m)};
        } else {
            const int b = 2 * j;
            return {__builtin_bit_cast(float, dw(b)), __builtin_bit_cast(float, dw(b + 1))};
        }
    }
};

constexpr int kPassMaxKH = 5;
constexpr int kPassLaneHi = 3;   // top lane bits a pass may give to other qubits
constexpr int kPassFoldPhase = 4;   // PassOp<double>::coef slot of PO_DENSE_REG_C's phase
constexpr int kPassMaxTileBits = 32;
constexpr int kPassMaxSlotBit = 27;   // a lane's 16-byte slot offset stays below 2^28 (32-bit byte offset)

// The ops work on a virtual index: bits [0, VB) inside the 16-byte vector,
// [VB, VB+6) the lane, the register slots at hpos[] and the tile number's bits
// in the remaining positions, ascending.  Where each of those bits lies in
// the state is given per pass: the qubit a virtual position stands for (the
// top three lane positions may stand for any qubit) is that qubit's own index
// bit.  Vector bits never move.
template <typename T> struct PassArgs {
    const PassOp<T> *ops;
    uint64_t n_tiles;        // wave tiles = 2^(n - VB - 6 - KH)
    int n_ops;
    int n_tile_bits;         // log2(n_tiles)
    int hpos[kPassMaxKH];    // register slots' virtual positions, ascending, all >= VB + 6
    int lane_pos[6];                     // slot-index bit (index bit - VB) of lane bit k, <= 27
    int reg_pos[kPassMaxKH];             // index bit of register slot j
    int tile_pos[kPassMaxTileBits];      // index bit of tile-number bit k
    cx<T> *s2;               // a second state the same program runs on (tiles n_tiles.. 2 n_tiles - 1), or null
};

// In-place updates.  The op loop carries the state in fixed registers, so
// every op must leave each result in the register its input came from; a
// result computed while its input is still needed lands elsewhere and costs
// a copy per value at the end of the op.  Each output below is therefore a
// partial sum of the other terms, formed while all inputs are live, plus one
// final FMA that reads the input the output replaces, last.

template <typename T> __device__ inline T fmaT(T a, T b, T c) {
    if constexpr (sizeof(T) == 8) return __builtin_fma(a, b, c);
    else return __builtin_fmaf(a, b, c);
}

// Loads and stores field by field, so the op loop carries plain scalars (a
// 16-byte copy into the Pack makes it carry <2 x i64> vectors that every op
// unpacks and repacks).  One 16-byte slot per lane: a cx<double> or two
// cx<float>.
template <typename T> __device__ inline void pass_ld(const pass_gv4f *p, Pack<T> &r) {
    const v4f v = __builtin_nontemporal_load(p);
    if constexpr (sizeof(T) == 8) {
        typedef double v2d __attribute__((ext_vector_type(2)));
        const v2d d = __builtin_bit_cast(v2d, v);
        r.e[0].re = d.x;
        r.e[0].im = d.y;
    } else {
        r.e[0].re = v.x; r.e[0].im = v.y; r.e[1].re = v.z; r.e[1].im = v.w;
    }
}
template <typename T> __device__ inline void pass_st(pass_gv4f *p, const Pack<T> &r) {
    v4f v;
    if constexpr (sizeof(T) == 8) {
        typedef double v2d __attribute__((ext_vector_type(2)));
        v2d d;
        d.x = r.e[0].re;
        d.y = r.e[0].im;
        v = __builtin_bit_cast(v4f, d);
    } else {
        v.x = r.e[0].re; v.y = r.e[0].im; v.z = r.e[1].re; v.w = r.e[1].im;
    }
    __builtin_nontemporal_store(v, p);
}

// Start of register row r of a tile, as a wave-uniform global address.  It
// lives in scalar registers, so each of the 2^KH loads and stores takes it
// plus one shared 32-bit lane offset, instead of a 64-bit vector address per
// row kept live across the op loop (2^KH x 2 registers: at KH = 5 those
// registers are what held the kernel to two waves per SIMD; it runs three).
template <typename T, int KH>
__device__ inline pass_gv4f *row_ptr(const cx<T> *s, uint64_t tb, const uint64_t (&hb)[KH], int r) {
    uint64_t off = tb;
#pragma unroll
    for (int j = 0; j < KH; j++) if ((r >> j) & 1) off += hb[j];
    const uint64_t p = reinterpret_cast<uint64_t>(s + off);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)p);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(p >> 32));
    return (pass_gv4f *)((uint64_t)lo | ((uint64_t)hi << 32));
}

// Cross-lane exchanges without the LDS.
// Lane bits 0 and 1: one DPP quad_perm per dword.  Lane bit 2: two DPP moves
// that both read the source -- row_shl by 4 written only to the 4-lane banks
// whose lane bit is 0, then row_shr written to the others -- so no DPP reads a
// register a VALU op has just written (that would need wait states).  Lane
// bit 3: one row_ror by 8, since rotating a 16-lane row by half its length is
// lane ^ 8.  Lane bits 4 and 5: v_permlane16_swap / v_permlane32_swap, which
// exchange the odd 16-lane rows (upper 32 lanes) of one register with the
// even rows (lower 32 lanes) of another; applied to a register pair that
// differs in register bit 0, this swaps lane bit 4 (5) with register bit 0,
// and a second application swaps them back.
template <int MASK> __device__ inline int lane_xor_dword(int x) {
    static_assert(MASK == 1 || MASK == 2 || MASK == 4 || MASK == 8, "DPP exchange covers lane bits 0-3");
    if constexpr (MASK == 1) return __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
    else if constexpr (MASK == 2) return __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
    else if constexpr (MASK == 4) {
        const int lo = __builtin_amdgcn_mov_dpp(x, 0x104, 0xF, 0x5, false);      // row_shl:4 -> banks 0, 2
        return __builtin_amdgcn_update_dpp(lo, x, 0x114, 0xF, 0xA, false);       // row_shr:4 -> banks 1, 3
    } else {
        return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);              // row_ror:8
    }
}
template <int MASK, typename T> __device__ inline T lane_xor(T x) {
    if constexpr (sizeof(T) == 8) {
        typedef int v2i __attribute__((ext_vector_type(2)));
        v2i w = __builtin_bit_cast(v2i, x);
        w.x = lane_xor_dword<MASK>(w.x);
        w.y = lane_xor_dword<MASK>(w.y);
        return __builtin_bit_cast(T, w);
    } else {
        return __builtin_bit_cast(T, lane_xor_dword<MASK>(__builtin_bit_cast(int, x)));
    }
}
template <int MASK> __device__ inline void permlane_swap_dword(int &a, int &b) {
    static_assert(MASK == 16 || MASK == 32, "permlane swaps cover lane bits 4 and 5");
    if constexpr (MASK == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(a, b, false, false);
        a = r[0]; b = r[1];
    } else {
        const auto r = __builtin_amdgcn_permlane32_swap(a, b, false, false);
        a = r[0]; b = r[1];
    }
}
template <int MASK, typename T> __device__ inline void permlane_swap(T &a, T &b) {
    if constexpr (sizeof(T) == 8) {
        typedef int v2i __attribute__((ext_vector_type(2)));
        v2i x = __builtin_bit_cast(v2i, a), y = __builtin_bit_cast(v2i, b);
        int x0 = x.x, x1 = x.y, y0 = y.x, y1 = y.y;
        permlane_swap_dword<MASK>(x0, y0);
        permlane_swap_dword<MASK>(x1, y1);
        x.x = x0; x.y = x1; y.x = y0; y.y = y1;
        a = __builtin_bit_cast(T, x); b = __builtin_bit_cast(T, y);
    } else {
        int x = __builtin_bit_cast(int, a), y = __builtin_bit_cast(int, b);
        permlane_swap_dword<MASK>(x, y);
        a = __builtin_bit_cast(T, x); b = __builtin_bit_cast(T, y);
    }
}

// x = m x + p into x's own registers.  Left to itself the compiler emits the
// two-address v_fmac, whose result goes to the addend's (the partial sum's)
// registers, so every op would move the state out of its loop registers and
// copy it back at the end of the op.
template <typename T> __device__ inline void fma_into(T &x, const T m, const T p) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 8) asm("v_fma_f64 %0, %1, %0, %2" : "+v"(x) : "v"(m), "v"(p));
    else asm("v_fma_f32 %0, %1, %0, %2" : "+v"(x) : "v"(m), "v"(p));
#else
    x = fmaT(m, x, p);
#endif
}

// x += k y into x's own registers (k wave-uniform: scalar operand; fma_acc_v:
// k per lane).
template <typename T> __device__ inline void fma_acc_s(T &x, const T k, const T y) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 8) asm("v_fma_f64 %0, %1, %2, %0" : "+v"(x) : "s"(k), "v"(y));
    else asm("v_fma_f32 %0, %1, %2, %0" : "+v"(x) : "s"(k), "v"(y));
#else
    x = fmaT(k, y, x);
#endif
}
template <typename T> __device__ inline void fma_acc_v(T &x, const T k, const T y) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 8) asm("v_fma_f64 %0, %1, %2, %0" : "+v"(x) : "v"(k), "v"(y));
    else asm("v_fma_f32 %0, %1, %2, %0" : "+v"(x) : "v"(k), "v"(y));
#else
    x = fmaT(k, y, x);
#endif
}

// The same with a wave-uniform multiplier read from scalar registers: the
// FMA then reads two vector operands instead of three (and needs no copy of
// the coefficient into vector registers).  Only for values that are the same
// in every lane -- the inline asm asks for scalar registers, and a value the
// compiler holds in vector registers is read from the first lane.
template <typename T> __device__ inline void fma_into_s(T &x, const T m, const T p) {
#if defined(__HIP_DEVICE_COMPILE__)
    if constexpr (sizeof(T) == 8) asm("v_fma_f64 %0, %1, %0, %2" : "+v"(x) : "s"(m), "v"(p));
    else asm("v_fma_f32 %0, %1, %0, %2" : "+v"(x) : "s"(m), "v"(p));
#else
    x = fmaT(m, x, p);
#endif
}
// complex<double> only: measured on the 30-qubit circuit, 225 -> 222 ms for
// complex128 and 136 -> 140 ms for complex64, which keeps the vector form.
template <bool SU, typename T> __device__ inline void fma_into_u(T &x, const T m, const T p) {
    if constexpr (SU && sizeof(T) == 8) fma_into_s(x, m, p);
    else fma_into(x, m, p);
}

// The four coefficients of an op as separate values.  Selections between them
// are made per component (v_cndmask on values): a select between array
// elements would select addresses, and the array would then live in memory
// (LDS or scratch) instead of registers.
template <typename T> struct Coef4 {
    cx<T> c0, c1, c2, c3;
};
template <typename T> __device__ inline cx<T> csel(bool p, const cx<T> a, const cx<T> b) {
    return {p ? a.re : b.re, p ? a.im : b.im};
}
// entry (y << 1) | x
template <typename T> __device__ inline cx<T> csel4(bool y, bool x, const Coef4<T> &k) {
    return csel(y, csel(x, k.c3, k.c2), csel(x, k.c1, k.c0));
}

// x *= f: 2 temporaries, 4 FP64 ops.
template <typename T> __device__ inline void ip_cmul(cx<T> &x, const cx<T> f) {
    const T pr = -(x.im * f.im), pi = x.re * f.im;
    fma_into(x.re, f.re, pr);
    fma_into(x.im, f.re, pi);
}
// The same with a wave-uniform factor (scalar-register operands).
template <typename T> __device__ inline void ip_cmul_s(cx<T> &x, const cx<T> f) {
    const T pr = -(x.im * f.im), pi = x.re * f.im;
    fma_into_s(x.re, f.re, pr);
    fma_into_s(x.im, f.re, pi);
}

// (x, y) <- (m0 x + m1 y, m2 x + m3 y): 4 temporaries, 16 FP64 ops.  SU:
// the coefficients are wave-uniform (see fma_into_s).
template <bool SU = false, typename T>
__device__ inline void ip_pair(cx<T> &x, cx<T> &y, const cx<T> m0, const cx<T> m1, const cx<T> m2, const cx<T> m3) {
    const T pxr = fmaT(m1.re, y.re, fmaT(-m1.im, y.im, -(m0.im * x.im)));
    const T pxi = fmaT(m1.re, y.im, fmaT(m1.im, y.re, m0.im * x.re));
    const T pyr = fmaT(m2.re, x.re, fmaT(-m2.im, x.im, -(m3.im * y.im)));
    const T pyi = fmaT(m2.re, x.im, fmaT(m2.im, x.re, m3.im * y.re));
    fma_into_u<SU>(x.re, m0.re, pxr);
    fma_into_u<SU>(x.im, m0.re, pxi);
    fma_into_u<SU>(y.re, m3.re, pyr);
    fma_into_u<SU>(y.im, m3.re, pyi);
}

// The same update when coefficients are known to be real (imaginary part
// exactly 0): RM = 1, m0 real (U3 and the U1 form of it); RM = 15, all four
// real (H, X, RY, ...).  The terms left out are exact zeros, so the results
// equal ip_pair's (up to the sign of a zero): 14 and 8 FP64 operations per
// pair instead of 16.
template <int RM, bool SU = false, typename T>
__device__ inline void ip_pair_real(cx<T> &x, cx<T> &y, const cx<T> m0, const cx<T> m1, const cx<T> m2,
                                    const cx<T> m3) {
    if constexpr (RM == 15) {
        const T pxr = m1.re * y.re, pxi = m1.re * y.im;
        const T pyr = m2.re * x.re, pyi = m2.re * x.im;
        fma_into_u<SU>(x.re, m0.re, pxr);
        fma_into_u<SU>(x.im, m0.re, pxi);
        fma_into_u<SU>(y.re, m3.re, pyr);
        fma_into_u<SU>(y.im, m3.re, pyi);
    } else if constexpr (RM == 1) {
        const T pxr = fmaT(m1.re, y.re, -(m1.im * y.im));
        const T pxi = fmaT(m1.re, y.im, m1.im * y.re);
        const T pyr = fmaT(m2.re, x.re, fmaT(-m2.im, x.im, -(m3.im * y.im)));
        const T pyi = fmaT(m2.re, x.im, fmaT(m2.im, x.re, m3.im * y.re));
        fma_into_u<SU>(x.re, m0.re, pxr);
        fma_into_u<SU>(x.im, m0.re, pxi);
        fma_into_u<SU>(y.re, m3.re, pyr);
        fma_into_u<SU>(y.im, m3.re, pyi);
    } else {
        ip_pair<SU>(x, y, m0, m1, m2, m3);
    }
}

// x <- ms x + mp p (p: the partner's value, already in other registers).
// RM = 15: both coefficients real, 4 FP64 ops instead of 8 (the terms left
// out are exact zeros).
template <int RM = 0, typename T>
__device__ inline void ip_lane(cx<T> &x, const cx<T> p, const cx<T> ms, const cx<T> mp) {
    if constexpr (RM == 15) {
        const T pr = mp.re * p.re, pi = mp.re * p.im;
        fma_into(x.re, ms.re, pr);
        fma_into(x.im, ms.re, pi);
    } else {
        const T pr = fmaT(mp.re, p.re, fmaT(-mp.im, p.im, -(ms.im * x.im)));
        const T pi = fmaT(mp.re, p.im, fmaT(mp.im, p.re, ms.im * x.re));
        fma_into(x.re, ms.re, pr);
        fma_into(x.im, ms.re, pi);
    }
}

// A rotation [[c, -s], [s, c]] (c, s >= 0) on a pair as three shears,
// [[1, k], [0, 1]] [[1, 0], [m, 1]] [[1, k], [0, 1]] with k = -s / (1 + c) and
// m = s (their product is the rotation when c^2 + s^2 = 1; |k| <= 1): six
// FMAs per pair, each written over its own input, against eight operations
// for the rotation's 2x2 product.
template <typename T> __device__ inline void ip_shear3(cx<T> &x, cx<T> &y, const T k, const T m) {
    fma_acc_s(x.re, k, y.re);
    fma_acc_s(x.im, k, y.im);
    fma_acc_s(y.re, m, x.re);
    fma_acc_s(y.im, m, x.im);
    fma_acc_s(x.re, k, y.re);
    fma_acc_s(x.im, k, y.im);
}
template <typename T, int J, int R, int E>
__device__ inline void pass_rot_reg(Pack<T> (&v)[R], const T k, const T m) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        if ((r >> J) & 1) continue;
#pragma unroll
        for (int e = 0; e < E; e++) ip_shear3(v[r].e[e], v[r | (1 << J)].e[e], k, m);
    }
}
template <typename T, int KH, int R, int E>
__device__ inline void pass_rot_reg_at(Pack<T> (&v)[R], int j, const T k, const T m) {
    switch (j) {
    case 0: pass_rot_reg<T, 0, R, E>(v, k, m); break;
    case 1: if constexpr (KH > 1) pass_rot_reg<T, 1, R, E>(v, k, m); break;
    case 2: if constexpr (KH > 2) pass_rot_reg<T, 2, R, E>(v, k, m); break;
    case 3: if constexpr (KH > 3) pass_rot_reg<T, 3, R, E>(v, k, m); break;
    case 4: if constexpr (KH > 4) pass_rot_reg<T, 4, R, E>(v, k, m); break;
    default: break;
    }
}

// Sign flip of a value where the lane's mask word is 0x80000000.
template <typename T> __device__ inline void flip_sign(cx<T> &x, uint32_t mk) {
    if constexpr (sizeof(T) == 8) {
        const uint64_t m = (uint64_t)mk << 32;
        x.re = __builtin_bit_cast(T, __builtin_bit_cast(uint64_t, x.re) ^ m);
        x.im = __builtin_bit_cast(T, __builtin_bit_cast(uint64_t, x.im) ^ m);
    } else {
        x.re = __builtin_bit_cast(T, __builtin_bit_cast(uint32_t, x.re) ^ mk);
        x.im = __builtin_bit_cast(T, __builtin_bit_cast(uint32_t, x.im) ^ mk);
    }
}

// PO_ROT_LANE on lane bits 0-3, scaled (see PO_ROT_LANE).  form 0: the lane
// with the bit clear gets x - k p, its partner x + k p: one FMA per part.
// form 1: k x - p and k x + p: the partner's value is negated on the lanes
// with the bit clear (sign bits), then one FMA per part.
template <typename T, int MASK, int R, int E>
__device__ inline void pass_rot_lane(Pack<T> (&v)[R], bool hi, int form, const T k) {
    if (form == 0) {
        const T kl = hi ? k : -k;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                const cx<T> p = {lane_xor<MASK>(v[r].e[e].re), lane_xor<MASK>(v[r].e[e].im)};
                fma_acc_v(v[r].e[e].re, kl, p.re);
                fma_acc_v(v[r].e[e].im, kl, p.im);
            }
    } else {
        const uint32_t mk = hi ? 0u : 0x80000000u;
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int e = 0; e < E; e++) {
                cx<T> p = {lane_xor<MASK>(v[r].e[e].re), lane_xor<MASK>(v[r].e[e].im)};
                flip_sign(p, mk);
                fma_into_s(v[r].e[e].re, k, p.re);
                fma_into_s(v[r].e[e].im, k, p.im);
            }
    }
}
// PO_ROT_LANE on lane bits 4-5: swap with register bit 0, shears, swap back.
template <typename T, int MASK, int R, int E>
__device__ inline void pass_rot_lane_swap(Pack<T> (&v)[R], const T k, const T m) {
#pragma unroll
    for (int r = 0; r < R; r += 2) {
#pragma unroll
        for (int e = 0; e < E; e++) {
            cx<T> &x = v[r].e[e], &y = v[r + 1].e[e];
            permlane_swap<MASK>(x.re, y.re);
            permlane_swap<MASK>(x.im, y.im);
            ip_shear3(x, y, k, m);
            permlane_swap<MASK>(x.re, y.re);
            permlane_swap<MASK>(x.im, y.im);
        }
    }
}

// One dense 2x2 on the register pairs (r, r | 2^J).  SU: coefficients the
// same in every lane.
template <typename T, int J, int R, int E, int RM = 0, bool SU = false>
__device__ inline void pass_dense_reg(Pack<T> (&v)[R], const cx<T> m0, const cx<T> m1, const cx<T> m2, const cx<T> m3) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        if ((r >> J) & 1) continue;
        const int q = r | (1 << J);
#pragma unroll
        for (int e = 0; e < E; e++) ip_pair_real<RM, SU>(v[r].e[e], v[q].e[e], m0, m1, m2, m3);
    }
}

// PO_DENSE_REG_C with the control in register slot sj (wave-uniform): each
// pair takes the coefficients of its control value, one uniform branch each.
template <typename T, int J, int R, int E, int RM>
__device__ inline void pass_dense_reg_ctl(Pack<T> (&v)[R], int sj, const cx<T> m0, const cx<T> m1, const cx<T> m2,
                                          const cx<T> m3, const cx<T> m1p, const cx<T> m3p) {
#pragma unroll
    for (int r = 0; r < R; r++) {
        if ((r >> J) & 1) continue;
        const int q = r | (1 << J);
        const bool c = (r >> sj) & 1;   // wave-uniform: scalar selects, one copy of the code
        const cx<T> a1 = {c ? m1p.re : m1.re, c ? m1p.im : m1.im}, a3 = {c ? m3p.re : m3.re, c ? m3p.im : m3.im};
#pragma unroll
        for (int e = 0; e < E; e++) ip_pair_real<RM, true>(v[r].e[e], v[q].e[e], m0, a1, m2, a3);
    }
}

template <typename T, int KH, int R, int E, int RM>
__device__ inline void pass_dense_reg_ctl_at(Pack<T> (&v)[R], int j, int sj, const cx<T> m0, const cx<T> m1,
                                             const cx<T> m2, const cx<T> m3, const cx<T> m1p, const cx<T> m3p) {
    switch (j) {
    case 0: pass_dense_reg_ctl<T, 0, R, E, RM>(v, sj, m0, m1, m2, m3, m1p, m3p); break;
    case 1: if constexpr (KH > 1) pass_dense_reg_ctl<T, 1, R, E, RM>(v, sj, m0, m1, m2, m3, m1p, m3p); break;
    case 2: if constexpr (KH > 2) pass_dense_reg_ctl<T, 2, R, E, RM>(v, sj, m0, m1, m2, m3, m1p, m3p); break;
    case 3: if constexpr (KH > 3) pass_dense_reg_ctl<T, 3, R, E, RM>(v, sj, m0, m1, m2, m3, m1p, m3p); break;
    case 4: if constexpr (KH > 4) pass_dense_reg_ctl<T, 4, R, E, RM>(v, sj, m0, m1, m2, m3, m1p, m3p); break;
    default: break;
    }
}

template <typename T, int KH, int R, int E, int RM, bool SU = false>
__device__ inline void pass_dense_reg_at(Pack<T> (&v)[R], int j, const cx<T> m0, const cx<T> m1, const cx<T> m2,
                                         const cx<T> m3) {
    switch (j) {
    case 0: pass_dense_reg<T, 0, R, E, RM, SU>(v, m0, m1, m2, m3); break;
    case 1: if constexpr (KH > 1) pass_dense_reg<T, 1, R, E, RM, SU>(v, m0, m1, m2, m3); break;
    case 2: if constexpr (KH > 2) pass_dense_reg<T, 2, R, E, RM, SU>(v, m0, m1, m2, m3); break;
    case 3: if constexpr (KH > 3) pass_dense_reg<T, 3, R, E, RM, SU>(v, m0, m1, m2, m3); break;
    case 4: if constexpr (KH > 4) pass_dense_reg<T, 4, R, E, RM, SU>(v, m0, m1, m2, m3); break;
    default: break;
    }
}

// Dense 2x2 on lane bit MASK.
// Lane bits 0-3: a lane whose bit is 0 holds x0 and gets m0 x0 + m1 x1; its
// partner holds x1 and gets m3 x1 + m2 x0.  The per-lane coefficients are
// chosen once per op; the partner's value comes from a DPP permutation.
// Lane bits 4-5: each register pair (r, r | 1) swaps the lane bit with register
// bit 0, takes the gate as a register-pair op, and swaps back.
template <typename T, int MASK, int R, int E, bool SU = false, int RM = 0>
__device__ inline void pass_dense_lane(Pack<T> (&v)[R], bool hi, const Coef4<T> k) {
    if constexpr (MASK <= 8) {
        const cx<T> ms = csel(hi, k.c3, k.c0), mp = csel(hi, k.c2, k.c1);
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const cx<T> p = {lane_xor<MASK>(v[r].e[e].re), lane_xor<MASK>(v[r].e[e].im)};
                ip_lane<RM>(v[r].e[e], p, ms, mp);
            }
        }
    } else {
        static_assert(R >= 2, "the swap needs a register bit");
#pragma unroll
        for (int r = 0; r < R; r += 2) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                cx<T> &x = v[r].e[e], &y = v[r + 1].e[e];
                permlane_swap<MASK>(x.re, y.re);
                permlane_swap<MASK>(x.im, y.im);
                ip_pair_real<RM == 15 ? 15 : 0, SU>(x, y, k.c0, k.c1, k.c2, k.c3);
                permlane_swap<MASK>(x.re, y.re);
                permlane_swap<MASK>(x.im, y.im);
            }
        }
    }
}

// PO_DENSE_LANE_C with the control in register slot sj: k0 / k1 are the
// coefficients for control 0 / 1.  Lane bits 0-3: one uniform branch per
// register.  Lane bits 4-5: after the swap, a pair (r, r | 1) shares its
// control unless sj = 0, whose bit the swap has moved into lane bit MASK.
template <typename T, int MASK, int R, int E, int RM = 0>
__device__ inline void pass_dense_lane_ctl(Pack<T> (&v)[R], bool hi, int sj, const Coef4<T> k0, const Coef4<T> k1,
                                           bool lane_bit_set) {
    constexpr int PM = RM == 15 ? 15 : 0;
    if constexpr (MASK <= 8) {
        const cx<T> ms0 = csel(hi, k0.c3, k0.c0), mp0 = csel(hi, k0.c2, k0.c1);
        const cx<T> ms1 = csel(hi, k1.c3, k1.c0), mp1 = csel(hi, k1.c2, k1.c1);
#pragma unroll
        for (int r = 0; r < R; r++) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                const cx<T> p = {lane_xor<MASK>(v[r].e[e].re), lane_xor<MASK>(v[r].e[e].im)};
                if ((r >> sj) & 1) ip_lane<PM>(v[r].e[e], p, ms1, mp1);
                else ip_lane<PM>(v[r].e[e], p, ms0, mp0);
            }
        }
    } else {
        static_assert(R >= 2, "the swap needs a register bit");
        // sj = 0: the control is the lane's bit MASK once swapped
        const Coef4<T> kl = {k0.c0, csel(lane_bit_set, k1.c1, k0.c1), k0.c2, csel(lane_bit_set, k1.c3, k0.c3)};
#pragma unroll
        for (int r = 0; r < R; r += 2) {
#pragma unroll
            for (int e = 0; e < E; e++) {
                cx<T> &x = v[r].e[e], &y = v[r + 1].e[e];
                permlane_swap<MASK>(x.re, y.re);
                permlane_swap<MASK>(x.im, y.im);
                if (sj == 0) ip_pair_real<PM>(x, y, kl.c0, kl.c1, kl.c2, kl.c3);
                else if ((r >> sj) & 1) ip_pair_real<PM>(x, y, k1.c0, k1.c1, k1.c2, k1.c3);
                else ip_pair_real<PM>(x, y, k0.c0, k0.c1, k0.c2, k0.c3);
                permlane_swap<MASK>(x.re, y.re);
                permlane_swap<MASK>(x.im, y.im);
            }
        }
    }
}

// Diagonal with a bit in register slot J: ph0 / ph1 are the lane's factors
// for register bit 0 / 1 (per vector element).  A half whose factors are all
// 1 (act0 / act1 false, wave-uniform) is skipped.
template <typename T, int J, int R, int E>
__device__ inline void pass_diag_reg1(Pack<T> (&v)[R], const cx<T> (&ph0)[E], const cx<T> (&ph1)[E], bool act0,
                                      bool act1) {
    if (act0) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if (!((r >> J) & 1))
#pragma unroll
                for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], ph0[e]);
    }
    if (act1) {
#pragma unroll
        for (int r = 0; r < R; r++)
            if ((r >> J) & 1)
#pragma unroll
                for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], ph1[e]);
    }
}

// Diagonal on register slots J0 < J1: the factor of register r is d[l(r)],
// known per r at compile time.  One uniform branch per entry skips the
// entries equal to 1 (a controlled phase touches a quarter of the registers).
template <typename T, int J0, int J1, int R, int E>
__device__ inline void pass_diag_reg2(Pack<T> (&v)[R], const Coef4<T> d, uint32_t active) {
#pragma unroll
    for (int l = 0; l < 4; l++) {
        if (!((active >> l) & 1)) continue;
        const cx<T> f = csel(l == 0, d.c0, csel(l == 1, d.c1, csel(l == 2, d.c2, d.c3)));   // l folds
#pragma unroll
        for (int r = 0; r < R; r++)
            if ((((r >> J0) & 1) | (((r >> J1) & 1) << 1)) == l)
#pragma unroll
                for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], f);
    }
}

template <typename T, int R, int E, bool SU, int RM>
__device__ inline void pass_dense_lane_at(Pack<T> (&v)[R], int a, bool hi, const Coef4<T> k) {
    switch (a) {
    case 0: pass_dense_lane<T, 1, R, E, false, RM>(v, hi, k); break;
    case 1: pass_dense_lane<T, 2, R, E, false, RM>(v, hi, k); break;
    case 2: pass_dense_lane<T, 4, R, E, false, RM>(v, hi, k); break;
    case 3: pass_dense_lane<T, 8, R, E, false, RM>(v, hi, k); break;
    case 4: pass_dense_lane<T, 16, R, E, SU, RM>(v, hi, k); break;
    case 5: pass_dense_lane<T, 32, R, E, SU, RM>(v, hi, k); break;
    default: break;
    }
}
template <typename T, int R, int E, int RM>
__device__ inline void pass_dense_lane_ctl_at(Pack<T> (&v)[R], int a, bool hi, int sj, const Coef4<T> k0,
                                              const Coef4<T> k1, bool lb) {
    switch (a) {
    case 0: pass_dense_lane_ctl<T, 1, R, E, RM>(v, hi, sj, k0, k1, lb); break;
    case 1: pass_dense_lane_ctl<T, 2, R, E, RM>(v, hi, sj, k0, k1, lb); break;
    case 2: pass_dense_lane_ctl<T, 4, R, E, RM>(v, hi, sj, k0, k1, lb); break;
    case 3: pass_dense_lane_ctl<T, 8, R, E, RM>(v, hi, sj, k0, k1, lb); break;
    case 4: pass_dense_lane_ctl<T, 16, R, E, RM>(v, hi, sj, k0, k1, lb); break;
    case 5: pass_dense_lane_ctl<T, 32, R, E, RM>(v, hi, sj, k0, k1, lb); break;
    default: break;
    }
}

template <typename T, int KH>
__global__ void __launch_bounds__(kBlock) k_pass(cx<T> *__restrict__ s, const PassArgs<T> a) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    constexpr int E = 1 << VB;
    constexpr int R = 1 << KH;
    const int lane = threadIdx.x & 63;
    // the wave index is uniform; readfirstlane lets the compiler know, so the
    // tile loop and every branch on tile bits are scalar branches, not
    // divergent regions that merge the state registers under exec masks
    // Workgroups are dealt to the 8 XCDs in turn; renumbering them gives each
    // XCD one contiguous eighth of the tiles.  This keeps the access pattern
    // even when the register qubits are all high ones (32 rows a MiB or more
    // apart), where the dealt order ran up to 1.35x slower.
    uint64_t bid = blockIdx.x;
    if ((gridDim.x & 7) == 0) bid = (bid & 7) * (gridDim.x >> 3) + (bid >> 3);
    // one tile per wave (the grid covers every tile); with a second state,
    // the waves past the first state's tiles take the second one's
    uint64_t t = bid * (kBlock / 64) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    cx<T> *st = s;
    if (t >= a.n_tiles) {
        if (!a.s2 || t >= 2 * a.n_tiles) return;
        t -= a.n_tiles;
        st = a.s2;
    }

    // state addresses of the tile: the tile number's bits placed by tile_pos
    // (uniform), the register rows by reg_pos, the lane's 16-byte slot by
    // lane_pos (one 32-bit offset per lane)
    uint64_t pt = 0;
#pragma unroll
    for (int k = 0; k < kPassMaxTileBits; k++)
        if (k < a.n_tile_bits) pt |= ((t >> k) & 1) << a.tile_pos[k];
    pt = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)pt) |
         ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(pt >> 32)) << 32);
    uint32_t slot = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) slot |= (uint32_t)((lane >> k) & 1) << a.lane_pos[k];
    slot &= (1u << (kPassMaxSlotBit + 1)) - 1;
    uint64_t hb[KH];
#pragma unroll
    for (int j = 0; j < KH; j++) hb[j] = 1ULL << a.reg_pos[j];
    {
        uint64_t tb = t << S;   // virtual tile bits only (uniform)
#pragma unroll
        for (int j = 0; j < KH; j++) tb = insert_zero(tb, a.hpos[j]);
        const uint64_t base = tb + ((uint64_t)lane << VB);   // virtual index of register 0

        Pack<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) pass_ld<T>(row_ptr<T, KH>(st, pt, hb, r) + slot, v[r]);

        for (int op = 0; op < a.n_ops; op++) {
            // every field is read up front: one batch of scalar loads, one wait
            const PassRec<T> rec(a.ops + op);
            struct { int kind, a, b, c; uint32_t active; } o = {rec.dw(0), rec.dw(1), rec.dw(2), rec.dw(3),
                                                                (uint32_t)rec.dw(4)};
            const Coef4<T> c = {rec.coef(0), rec.coef(1), rec.coef(2), rec.coef(3)};
            switch (o.kind) {
            case PO_DENSE_REG: {
                // o.c: 15 all coefficients real, 1 m0 real, 0 general
                const cx<T> m0 = c.c0, m1 = c.c1, m2 = c.c2, m3 = c.c3;
                if (o.c == 15) pass_dense_reg_at<T, KH, R, E, 15, true>(v, o.a, m0, m1, m2, m3);
                else if (o.c == 1) pass_dense_reg_at<T, KH, R, E, 1, true>(v, o.a, m0, m1, m2, m3);
                else pass_dense_reg_at<T, KH, R, E, 0, true>(v, o.a, m0, m1, m2, m3);
                break; }
            case PO_DENSE_REG_C: {
                if constexpr (sizeof(T) == 8) {
                    const cx<T> m0 = c.c0, m1 = c.c1, m2 = c.c2, m3 = c.c3, ph = rec.coef(kPassFoldPhase);
                    const cx<T> m1p = {m1.re * ph.re - m1.im * ph.im, m1.re * ph.im + m1.im * ph.re};
                    const cx<T> m3p = {m3.re * ph.re - m3.im * ph.im, m3.re * ph.im + m3.im * ph.re};
                    if (o.active == 0) {   // control in a register slot
                        if (o.c == 15) pass_dense_reg_ctl_at<T, KH, R, E, 15>(v, o.a, o.b, m0, m1, m2, m3, m1p, m3p);
                        else if (o.c == 1) pass_dense_reg_ctl_at<T, KH, R, E, 1>(v, o.a, o.b, m0, m1, m2, m3, m1p, m3p);
                        else pass_dense_reg_ctl_at<T, KH, R, E, 0>(v, o.a, o.b, m0, m1, m2, m3, m1p, m3p);
                    } else {               // lane bit (per lane) or tile bit (per wave)
                        const bool on = o.active == 1 ? ((base >> o.b) & 1) != 0 : ((tb >> o.b) & 1) != 0;
                        const cx<T> m1s = csel(on, m1p, m1), m3s = csel(on, m3p, m3);
                        if (o.c == 15) pass_dense_reg_at<T, KH, R, E, 15>(v, o.a, m0, m1s, m2, m3s);
                        else if (o.c == 1) pass_dense_reg_at<T, KH, R, E, 1>(v, o.a, m0, m1s, m2, m3s);
                        else pass_dense_reg_at<T, KH, R, E, 0>(v, o.a, m0, m1s, m2, m3s);
                    }
                }
                break; }
            case PO_DENSE_LANE: {
                const bool hi = (lane >> o.a) & 1;
                if (o.c == 15) pass_dense_lane_at<T, R, E, true, 15>(v, o.a, hi, c);
                else pass_dense_lane_at<T, R, E, true, 0>(v, o.a, hi, c);
                break; }
            case PO_DENSE_LANE_C: {
                if constexpr (sizeof(T) == 8) {
                    const bool hi = (lane >> o.a) & 1;
                    const cx<T> ph = rec.coef(kPassFoldPhase);
                    const cx<T> m1p = {c.c1.re * ph.re - c.c1.im * ph.im, c.c1.re * ph.im + c.c1.im * ph.re};
                    const cx<T> m3p = {c.c3.re * ph.re - c.c3.im * ph.im, c.c3.re * ph.im + c.c3.im * ph.re};
                    const Coef4<T> k1 = {c.c0, m1p, c.c2, m3p};
                    if (o.active == 0) {   // control in register slot o.b
                        const bool lb = ((lane >> o.a) & 1) != 0;
                        if (o.c == 15) pass_dense_lane_ctl_at<T, R, E, 15>(v, o.a, hi, o.b, c, k1, lb);
                        else pass_dense_lane_ctl_at<T, R, E, 0>(v, o.a, hi, o.b, c, k1, lb);
                    } else {               // lane bit (per lane) or tile bit (per wave)
                        const bool on = o.active == 1 ? ((base >> o.b) & 1) != 0 : ((tb >> o.b) & 1) != 0;
                        const Coef4<T> k = {c.c0, csel(on, m1p, c.c1), c.c2, csel(on, m3p, c.c3)};
                        if (o.c == 15) pass_dense_lane_at<T, R, E, false, 15>(v, o.a, hi, k);
                        else pass_dense_lane_at<T, R, E, false, 0>(v, o.a, hi, k);
                    }
                }
                break; }
            case PO_DENSE_VEC: {
                if constexpr (E == 2) {
                    const cx<T> m0 = c.c0, m1 = c.c1, m2 = c.c2, m3 = c.c3;
                    if (o.c == 15) {
#pragma unroll
                        for (int r = 0; r < R; r++) ip_pair_real<15, true>(v[r].e[0], v[r].e[1], m0, m1, m2, m3);
                    } else {
#pragma unroll
                        for (int r = 0; r < R; r++) ip_pair<true>(v[r].e[0], v[r].e[1], m0, m1, m2, m3);
                    }
                }
                break; }
            case PO_ROT_REG: {
                const cx<T> km = c.c0;
                if (o.c == 1) {    // inside the complex64 vector
                    if constexpr (E == 2) {
#pragma unroll
                        for (int r = 0; r < R; r++) ip_shear3(v[r].e[0], v[r].e[1], km.re, km.im);
                    }
                } else {
                    pass_rot_reg_at<T, KH, R, E>(v, o.a, km.re, km.im);
                }
                break; }
            case PO_ROT_LANE: {
                const cx<T> km = c.c0;
                const bool hi = (lane >> o.a) & 1;
                switch (o.a) {
                case 0: pass_rot_lane<T, 1, R, E>(v, hi, o.c, km.re); break;
                case 1: pass_rot_lane<T, 2, R, E>(v, hi, o.c, km.re); break;
                case 2: pass_rot_lane<T, 4, R, E>(v, hi, o.c, km.re); break;
                case 3: pass_rot_lane<T, 8, R, E>(v, hi, o.c, km.re); break;
                case 4: pass_rot_lane_swap<T, 16, R, E>(v, km.re, km.im); break;
                case 5: pass_rot_lane_swap<T, 32, R, E>(v, km.re, km.im); break;
                default: break;
                }
                break; }
            case PO_SIGN: {
                const int t0 = rec.dw(5), t1 = rec.dw(6);
                if (t0 >= 0 && !((tb >> t0) & 1)) break;   // a tile bit clear: no amplitude of the wave
                if (t1 >= 0 && !((tb >> t1) & 1)) break;
                uint32_t mk[E];
#pragma unroll
                for (int e = 0; e < E; e++)
                    mk[e] = (((uint32_t)(base + e) & (uint32_t)o.a) == (uint32_t)o.a) ? 0x80000000u : 0u;
#pragma unroll
                for (int r = 0; r < R; r++)
                    if ((r & o.b) == o.b)
#pragma unroll
                        for (int e = 0; e < E; e++) flip_sign(v[r].e[e], mk[e]);
                break; }
            case PO_PHASES: {
                // L per lane (and vector element): the scalar times z_p for each
                // marked position whose bit is set
                cx<T> L[E];
#pragma unroll
                for (int e = 0; e < E; e++) L[e] = c.c0;
                int x = op + 1;   // the next coefficient record
                if (o.a) {
                    const PassRec<T> zr(a.ops + x);
                    x++;
#pragma unroll
                    for (int p = 0; p < S; p++) {
                        if (!((o.a >> p) & 1)) continue;
                        const cx<T> z = zr.xcoef(p);
#pragma unroll
                        for (int e = 0; e < E; e++) L[e] = csel(((base + e) >> p) & 1, cmul(L[e], z), L[e]);
                    }
                }
                if (o.b) {         // register table: 8 entries a record
#pragma unroll
                    for (int k = 0; k < (R + 7) / 8; k++) {
                        const PassRec<T> tr(a.ops + x + k);
#pragma unroll
                        for (int j = 0; j < 8 && k * 8 + j < R; j++) {
                            const int r = k * 8 + j;
                            const cx<T> t = tr.xcoef(j);
                            if (o.a) {
#pragma unroll
                                for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], cmul(L[e], t));
                            } else {   // the scalar is in the table: a wave-uniform factor per register
#pragma unroll
                                for (int e = 0; e < E; e++) ip_cmul_s(v[r].e[e], t);
                            }
                        }
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < R; r++)
#pragma unroll
                        for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], L[e]);
                }
                op += o.c;   // past the coefficient records
                break; }
            case PO_DIAG_FREE: {
                // the factor depends only on lane, tile and vector bits: one per lane (and element)
                if (o.c) {   // both bits are tile bits: one entry for the whole wave
                    const int l = (int)((tb >> o.a) & 1) | (o.b >= 0 ? (int)((tb >> o.b) & 1) << 1 : 0);
                    if (!((o.active >> l) & 1)) break;
                    const cx<T> f = csel4((l & 2) != 0, (l & 1) != 0, c);
#pragma unroll
                    for (int r = 0; r < R; r++)
#pragma unroll
                        for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], f);
                    break;
                }
                cx<T> f[E];
#pragma unroll
                for (int e = 0; e < E; e++) {
                    const uint64_t idx = base + e;
                    const int x = (int)((idx >> o.a) & 1), y = o.b >= 0 ? (int)((idx >> o.b) & 1) : 0;
                    f[e] = csel4(y != 0, x != 0, c);
                }
#pragma unroll
                for (int r = 0; r < R; r++)
#pragma unroll
                    for (int e = 0; e < E; e++) ip_cmul(v[r].e[e], f[e]);
                break; }
            case PO_DIAG_REG1: {
                // entries e[(register bit << 1) | y]; y is the other bit of the index
                cx<T> ph0[E], ph1[E];
                bool act0, act1;
                if (o.c) {   // y is a tile bit (or absent): the same for the whole wave
                    const int y = o.b >= 0 ? (int)((tb >> o.b) & 1) : 0;
                    act0 = (o.active >> y) & 1;
                    act1 = (o.active >> (2 + y)) & 1;
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        ph0[e] = csel(y != 0, c.c1, c.c0);
                        ph1[e] = csel(y != 0, c.c3, c.c2);
                    }
                } else {     // y is a lane or vector bit
                    act0 = (o.active & 3u) != 0;
                    act1 = (o.active & 12u) != 0;
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        const uint64_t idx = base + e;
                        const int y = (int)((idx >> o.b) & 1);
                        ph0[e] = csel(y != 0, c.c1, c.c0);
                        ph1[e] = csel(y != 0, c.c3, c.c2);
                    }
                }
                switch (o.a) {
                case 0: pass_diag_reg1<T, 0, R, E>(v, ph0, ph1, act0, act1); break;
                case 1: if constexpr (KH > 1) pass_diag_reg1<T, 1, R, E>(v, ph0, ph1, act0, act1); break;
                case 2: if constexpr (KH > 2) pass_diag_reg1<T, 2, R, E>(v, ph0, ph1, act0, act1); break;
                case 3: if constexpr (KH > 3) pass_diag_reg1<T, 3, R, E>(v, ph0, ph1, act0, act1); break;
                case 4: if constexpr (KH > 4) pass_diag_reg1<T, 4, R, E>(v, ph0, ph1, act0, act1); break;
                default: break;
                }
                break; }
            case PO_DIAG_REG2: {
                switch (o.a * 8 + o.b) {
                case 1: if constexpr (KH > 1) pass_diag_reg2<T, 0, 1, R, E>(v, c, o.active); break;
                case 2: if constexpr (KH > 2) pass_diag_reg2<T, 0, 2, R, E>(v, c, o.active); break;
                case 3: if constexpr (KH > 3) pass_diag_reg2<T, 0, 3, R, E>(v, c, o.active); break;
                case 4: if constexpr (KH > 4) pass_diag_reg2<T, 0, 4, R, E>(v, c, o.active); break;
                case 10: if constexpr (KH > 2) pass_diag_reg2<T, 1, 2, R, E>(v, c, o.active); break;
                case 11: if constexpr (KH > 3) pass_diag_reg2<T, 1, 3, R, E>(v, c, o.active); break;
                case 12: if constexpr (KH > 4) pass_diag_reg2<T, 1, 4, R, E>(v, c, o.active); break;
                case 19: if constexpr (KH > 3) pass_diag_reg2<T, 2, 3, R, E>(v, c, o.active); break;
                case 20: if constexpr (KH > 4) pass_diag_reg2<T, 2, 4, R, E>(v, c, o.active); break;
                case 28: if constexpr (KH > 4) pass_diag_reg2<T, 3, 4, R, E>(v, c, o.active); break;
                default: break;
                }
                break; }
            default:
                break;
            }
        }

        // The row addresses again from the tile's bits: kept from the loads, the
        // 2^KH of them would hold 2^(KH+1) scalar registers through the op loop
        // (and push other values out to vector-register lanes).
        uint32_t ptl = (uint32_t)pt, pth = (uint32_t)(pt >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
        asm volatile("" : "+s"(ptl), "+s"(pth));
#endif
        const uint64_t pt2 = (uint64_t)ptl | ((uint64_t)pth << 32);
#pragma unroll
        for (int r = 0; r < R; r++) pass_st<T>(row_ptr<T, KH>(st, pt2, hb, r) + slot, v[r]);
    }
}

}  // namespace sqk
