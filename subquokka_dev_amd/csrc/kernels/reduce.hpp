// reduce.hpp -- FP64 reductions over the state and dtype conversion.
// Two-pass, fixed-order: per-block partial sums go to a scratch array and the
// host adds the partials in block order, so results do not depend on timing.
#pragma once
#include "common.hpp"

namespace sqk {

constexpr int kRedBlocks = 2048;

template <int NV>
__device__ inline void block_sum(double (&v)[NV], double *__restrict__ out, uint64_t stride = 1) {
    __shared__ double sm[NV][kBlock / 64];
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double x = v[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((threadIdx.x & 63) == 0) sm[i][threadIdx.x >> 6] = x;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double t = 0;
            for (int w = 0; w < kBlock / 64; w++) t += sm[i][w];
            out[i * stride] = t;
        }
    }
}

// sum |a_i|^2 and sum |a_i|^2 * zz(i) where zz(i) = sum_q (+1 if bits q and
// (q+1) mod n agree else -1).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_norm_zz(const cx<T> *__restrict__ s, uint64_t n_amp, int n,
                                                    double *__restrict__ part) {
    double v[2] = {0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t ring_mask = (n >= 64) ? ~0ULL : ((1ULL << n) - 1);
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_amp; i += stride) {
        const cx<T> a = s[i];
        const double p = (double)a.re * (double)a.re + (double)a.im * (double)a.im;
        // rotate right by one within n bits: bit q of rot = bit (q+1) mod n of i
        const uint64_t rot = ((i >> 1) | (i << (n - 1))) & ring_mask;
        const int differ = __popcll((i ^ rot) & ring_mask);
        v[0] += p;
        v[1] += p * (double)(n - 2 * differ);
    }
    block_sum<2>(v, part + 2 * blockIdx.x);
}

// <a|b> and max |a_i - b_i| between two states of possibly different dtypes.
template <typename TA, typename TB>
__global__ void __launch_bounds__(kBlock) k_compare(const cx<TA> *__restrict__ a, const cx<TB> *__restrict__ b,
                                                    uint64_t n_amp, double *__restrict__ part) {
    double v[3] = {0, 0, 0};
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n_amp; i += stride) {
        const double ar = a[i].re, ai = a[i].im, br = b[i].re, bi = b[i].im;
        v[0] += ar * br + ai * bi;       // Re conj(a) b
        v[1] += ar * bi - ai * br;       // Im conj(a) b
        const double dr = ar - br, di = ai - bi;
        v[2] = fmax(v[2], sqrt(dr * dr + di * di));
    }
    // max does not sum: reduce it separately
    __shared__ double smax[kBlock / 64];
    double m = v[2];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
    double s2[2] = {v[0], v[1]};
    block_sum<2>(s2, part + 3 * blockIdx.x);
    if (threadIdx.x == 0) {
        double mm = 0;
        for (int w = 0; w < kBlock / 64; w++) mm = fmax(mm, smax[w]);
        part[3 * blockIdx.x + 2] = mm;
    }
}

// dst[i] = (TD) src[i]
template <typename TS, typename TD>
__global__ void __launch_bounds__(kBlock) k_convert(const cx<TS> *__restrict__ src, cx<TD> *__restrict__ dst, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        dst[i] = {(TD)src[i].re, (TD)src[i].im};
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill_zero_one(cx<T> *__restrict__ s, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        s[i] = {i == 0 ? (T)1 : (T)0, (T)0};
}

}  // namespace sqk
