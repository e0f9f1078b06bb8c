// common.hpp -- types and index helpers shared by the gfx950 kernels.
//
// State layout in HBM: interleaved (re, im), amplitude i at element i, qubit q
// = bit q of i (the reference's layout, simulator.cpp:168-175).
//
// Thread mapping used by every sweep kernel ("wave tile"):
//   * each lane moves one 16-byte vector: 1 amplitude (complex<double>) or
//     2 amplitudes (complex<float>), i.e. index bits [0, VB);
//   * the 64 lanes of a wavefront cover the next 6 index bits [VB, VB+6), so a
//     load instruction reads 1 KiB of contiguous HBM;
//   * gate targets at or above bit S = VB+6 become register indices
//     (bit-inserted addressing), targets inside the lane bits are exchanged
//     across lanes with __shfl_xor (ds_bpermute), and a target inside the
//     vector is handled in-register.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sqk {

template <typename T> struct cx { T re, im; };

template <typename T> __host__ __device__ inline cx<T> cmul(cx<T> a, cx<T> b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ inline cx<T> cadd(cx<T> a, cx<T> b) { return {a.re + b.re, a.im + b.im}; }
// acc + a*b
template <typename T> __host__ __device__ inline cx<T> cfma(cx<T> acc, cx<T> a, cx<T> b) {
    return {acc.re + a.re * b.re - a.im * b.im, acc.im + a.re * b.im + a.im * b.re};
}

template <typename T> struct Traits;
template <> struct Traits<double> { static constexpr int VB = 0, LINE = 3; };  // 16 B = 1 amplitude
template <> struct Traits<float> { static constexpr int VB = 1, LINE = 4; };   // 16 B = 2 amplitudes
// LINE: log2 of amplitudes per 128-byte cache line

// One 16-byte vector of amplitudes.
template <typename T> struct alignas(16) Pack { cx<T> e[1 << Traits<T>::VB]; };

typedef float v4f __attribute__((ext_vector_type(4)));

// 16-byte loads/stores.  Sweeps stream the state once per launch, so they use
// nontemporal accesses (measured +5-10% HBM throughput on MI355X).
template <typename T> __device__ inline Pack<T> ld(const cx<T> *p) {
    Pack<T> r;
    const v4f v = __builtin_nontemporal_load(reinterpret_cast<const v4f *>(p));
    __builtin_memcpy(&r, &v, 16);
    return r;
}
template <typename T> __device__ inline void st(cx<T> *p, const Pack<T> &r) {
    v4f v;
    __builtin_memcpy(&v, &r, 16);
    __builtin_nontemporal_store(v, reinterpret_cast<v4f *>(p));
}

template <typename T> __device__ inline Pack<T> shfl_xor_pack(const Pack<T> &x, int mask) {
    int w[4];
    __builtin_memcpy(w, &x, 16);
#pragma unroll
    for (int i = 0; i < 4; i++) w[i] = __shfl_xor(w[i], mask, 64);
    Pack<T> r;
    __builtin_memcpy(&r, w, 16);
    return r;
}

// Inserts a 0 bit at position pos (bits >= pos move up by one).
__host__ __device__ inline uint64_t insert_zero(uint64_t x, int pos) {
    const uint64_t lo = x & ((1ULL << pos) - 1);
    return ((x >> pos) << (pos + 1)) | lo;
}
__host__ __device__ inline uint64_t insert_one(uint64_t x, int pos) {
    return insert_zero(x, pos) | (1ULL << pos);
}

constexpr int kBlock = 256;   // 4 wavefronts per workgroup

}  // namespace sqk
