// bitperm.hpp -- one out-of-place pass that moves index bit q of the state to
// bit p[q] (any permutation of the index bits).
//
// A workgroup owns the 2^tb amplitudes that share the values of every bit
// outside a set T of tile bits.  T holds the source's low c bits (so each
// wave reads runs of 64 consecutive amplitudes) and the bits that land on
// the destination's low c bits (so each wave writes runs of 64 consecutive
// amplitudes).  The tile goes through LDS: read in source order, written in
// destination order.  Both buffers are read and written once: one pass of
// HBM traffic whatever the permutation.
#pragma once
#include "common.hpp"

namespace sqk {

constexpr int kPermMaxTile = 12;   // tile bits: 2^12 amplitudes, 64 KiB of complex128 in LDS
constexpr int kPermMaxBits = 40;

struct BitPermArgs {
    int tb;                             // tile bits
    int n_other;                        // the other bits: log2 of the workgroup count
    int ld_src[kPermMaxTile];           // source bit of load-order bit k (k < c: bit k itself)
    int st_loc[kPermMaxTile];           // load-order position of the bit that store-order bit k carries
    int st_dst[kPermMaxTile];           // destination bit of store-order bit k (k < c: bit k itself)
    int other_src[kPermMaxBits];        // the bits outside the tile, ascending: where they are read
    int other_dst[kPermMaxBits];        // and where they are written
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_bitperm(const cx<T> *__restrict__ src, cx<T> *__restrict__ dst,
                                                    const BitPermArgs a) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    cx<T> *lds = reinterpret_cast<cx<T> *>(lds_raw);
    const uint64_t w = blockIdx.x;
    uint64_t sb = 0, db = 0;   // the tile's fixed bits, read and written (uniform)
    for (int k = 0; k < a.n_other; k++) {
        const uint64_t b = (w >> k) & 1;
        sb |= b << a.other_src[k];
        db |= b << a.other_dst[k];
    }
    const uint32_t cnt = 1u << a.tb;
    for (uint32_t i = threadIdx.x; i < cnt; i += kBlock) {
        uint64_t off = 0;
        for (int k = 0; k < a.tb; k++) off |= (uint64_t)((i >> k) & 1) << a.ld_src[k];
        lds[i] = src[sb + off];
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < cnt; j += kBlock) {
        uint32_t i = 0;
        uint64_t off = 0;
        for (int k = 0; k < a.tb; k++) {
            const uint32_t b = (j >> k) & 1;
            i |= b << a.st_loc[k];
            off |= (uint64_t)b << a.st_dst[k];
        }
        dst[db + off] = lds[i];
    }
}

}  // namespace sqk
