// sample.hpp -- seeded measurement sampling in the computational basis.
//
// The reference has no sampler (only the unused MEASURE gate type,
// circuit.hpp:11).  This build's sampler is defined so that a CPU restatement
// reproduces it bit for bit:
//   p_i   = re_i*re_i + im_i*im_i in FP64, no fused multiply-add;
//   a "segment" is 16 consecutive amplitudes, a block is 64 segments (1024
//   amplitudes, or the whole state below 2^10); s = the sequential FP64 sum of
//   a segment's p, B = the sequential sum of a block's s, C = the sequential
//   exclusive prefix of B over blocks (host), total = C_last + B_last;
//   shot k: u = ((x_k >> 11) * 2^-53) * total with x_k the k-th SplitMix64
//   output for the seed; block b = first block with C_b + B_b > u (else the
//   last block with B > 0); r = u - C_b; segment m = first segment of b whose
//   sequential inclusive prefix L_m + s_m > r (else the last with s > 0);
//   t = r - L_m; amplitude j = first of m whose sequential inclusive sum > t
//   (else the last with p > 0).
#pragma once
#include "common.hpp"

namespace sqk {

constexpr int kSampleSeg = 16;     // amplitudes per segment (per lane)
constexpr int kSampleBlockLog = 10;

template <typename T>
__device__ inline double sample_p(const cx<T> &a) {
    const double re = a.re, im = a.im;
    return __dadd_rn(__dmul_rn(re, re), __dmul_rn(im, im));
}

// Block sums B: one wave per block, lane l sums segment l, lane 0 adds the
// 64 segment sums in order.  seg_per_block < 64 only for states below 2^10.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sample_blocks(const cx<T> *__restrict__ s, uint64_t n_blocks,
                                                          int seg_per_block, double *__restrict__ bsum) {
    __shared__ double seg[kBlock / 64][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t b = (uint64_t)blockIdx.x * (kBlock / 64) + w;
    double acc = 0.0;
    if (b < n_blocks && lane < seg_per_block) {
        const cx<T> *p = s + (b * seg_per_block + lane) * kSampleSeg;
#pragma unroll
        for (int j = 0; j < kSampleSeg; j++) acc = __dadd_rn(acc, sample_p(p[j]));
    }
    seg[w][lane] = acc;
    __syncthreads();
    if (lane == 0 && b < n_blocks) {
        double t = 0.0;
        for (int l = 0; l < seg_per_block; l++) t = __dadd_rn(t, seg[w][l]);
        bsum[b] = t;
    }
}

struct SampleShot {
    uint64_t block;
    double r;          // residual inside the block
};

// One thread per shot: the shot's uniform and its block.  Shot k of a seed
// uses SplitMix64 output k, i.e. the mix of seed + (k + 1) * gamma, so every
// shot is drawn independently of the others and equals the sequential
// generator's.  u = uniform * total (one rounding); the block is the first
// whose inclusive prefix exceeds u (the comparisons of std::upper_bound), or
// the last block of nonzero weight when none does; r = u - exclusive prefix.
struct SampleLocateArgs {
    uint64_t seed, k0, n_shots, n_blocks, last_nz;
    double total;
    const double *incl, *excl;   // inclusive / exclusive block prefixes (sequential FP64 sums)
    SampleShot *shots;
};

__global__ void __launch_bounds__(kBlock) k_sample_locate(const SampleLocateArgs a) {
    const uint64_t k = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= a.n_shots) return;
    uint64_t z = a.seed + (a.k0 + k + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    const double u = __dmul_rn((double)(z >> 11) * 0x1.0p-53, a.total);
    uint64_t first = 0, count = a.n_blocks;
    while (count > 0) {
        const uint64_t step = count / 2, it = first + step;
        if (!(u < a.incl[it])) {
            first = it + 1;
            count -= step + 1;
        } else {
            count = step;
        }
    }
    const uint64_t b = first >= a.n_blocks ? a.last_nz : first;
    a.shots[k] = SampleShot{b, __dsub_rn(u, a.excl[b])};
}

// One wave per shot: segment sums, sequential prefix by lane 0, then a
// sequential scan of the chosen segment.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sample_find(const cx<T> *__restrict__ s, const SampleShot *__restrict__ shots,
                                                        uint64_t n_shots, int seg_per_block, uint64_t *__restrict__ out) {
    __shared__ double seg[kBlock / 64][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t k = (uint64_t)blockIdx.x * (kBlock / 64) + w;
    SampleShot sh{0, 0.0};
    if (k < n_shots) sh = shots[k];
    const cx<T> *base = s + sh.block * (uint64_t)seg_per_block * kSampleSeg;
    double acc = 0.0;
    if (k < n_shots && lane < seg_per_block) {
        const cx<T> *p = base + (uint64_t)lane * kSampleSeg;
#pragma unroll
        for (int j = 0; j < kSampleSeg; j++) acc = __dadd_rn(acc, sample_p(p[j]));
    }
    seg[w][lane] = acc;
    __syncthreads();
    if (lane == 0 && k < n_shots) {
        double L = 0.0;
        int m = -1, last = 0;
        double Lm = 0.0, Llast = 0.0;
        for (int l = 0; l < seg_per_block; l++) {
            const double sl = seg[w][l];
            if (sl > 0.0) { last = l; Llast = L; }
            if (m < 0 && __dadd_rn(L, sl) > sh.r) { m = l; Lm = L; }
            L = __dadd_rn(L, sl);
        }
        if (m < 0) { m = last; Lm = Llast; }
        const double t = __dsub_rn(sh.r, Lm);
        const cx<T> *p = base + (uint64_t)m * kSampleSeg;
        double a = 0.0;
        int j = -1, jlast = 0;
        for (int i = 0; i < kSampleSeg; i++) {
            const double pi = sample_p(p[i]);
            if (pi > 0.0) jlast = i;
            a = __dadd_rn(a, pi);
            if (j < 0 && a > t) j = i;
        }
        if (j < 0) j = jlast;
        out[k] = sh.block * (uint64_t)seg_per_block * kSampleSeg + (uint64_t)m * kSampleSeg + (uint64_t)j;
    }
}

}  // namespace sqk
