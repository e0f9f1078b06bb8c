// sq_sample.hip -- seeded measurement shots from the HBM-resident state
// (kernels/sample.hpp defines the sampling rule exactly).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "kernels/sample.hpp"

using namespace sqk;

namespace {

constexpr uint64_t kShotBatch = 1ULL << 20;

template <typename T>
int run_sample(sq_engine *e, uint64_t seed, uint64_t shots, uint64_t *out) {
    const int n = e->n;
    const int spb = n >= kSampleBlockLog ? 64 : 1 << (n - 4);   // segments per block
    const uint64_t n_blocks = n >= kSampleBlockLog ? 1ULL << (n - kSampleBlockLog) : 1;
    const uint64_t batch = std::min(shots, kShotBatch);
    const size_t blk_b = ((n_blocks * sizeof(double) + 255) / 256) * 256;
    const size_t shot_b = ((batch * sizeof(SampleShot) + 255) / 256) * 256;
    const size_t need = 3 * blk_b + shot_b + batch * sizeof(uint64_t);
    if (hipStreamSynchronize(e->stream) != hipSuccess) return e->fail(SQ_ERR_HIP, "stream sync");
    if (need > e->sample_cap) {
        if (e->d_sample) hipFree(e->d_sample);
        e->d_sample = nullptr;
        e->sample_cap = 0;
        hipError_t he = hipMalloc(&e->d_sample, need);
        if (he != hipSuccess) { hipGetLastError(); return e->fail(SQ_ERR_OOM, "hipMalloc(sample buffers)"); }
        e->sample_cap = need;
    }
    char *base = static_cast<char *>(e->d_sample);
    double *d_bsum = reinterpret_cast<double *>(base);
    double *d_incl = reinterpret_cast<double *>(base + blk_b);
    double *d_excl = reinterpret_cast<double *>(base + 2 * blk_b);
    SampleShot *d_shots = reinterpret_cast<SampleShot *>(base + 3 * blk_b);
    uint64_t *d_out = reinterpret_cast<uint64_t *>(base + 3 * blk_b + shot_b);
    const cx<T> *s = static_cast<const cx<T> *>(e->d_state);

    const unsigned g1 = (unsigned)((n_blocks + kBlock / 64 - 1) / (kBlock / 64));
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sample_blocks<T>), g1, kBlock, 0, e->stream, s, n_blocks, spb, d_bsum);
    int rc = e->check_launch("k_sample_blocks");
    if (rc) return rc;
    std::vector<double> B(n_blocks), I(n_blocks);
    hipError_t he = hipMemcpyAsync(B.data(), d_bsum, n_blocks * sizeof(double), hipMemcpyDeviceToHost, e->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return e->hip_fail(he, "block sums");
    double acc = 0.0;
    uint64_t last_nz = 0;
    std::vector<double> C(n_blocks);
    for (uint64_t b = 0; b < n_blocks; b++) {
        C[b] = acc;
        acc = acc + B[b];
        I[b] = acc;
        if (B[b] > 0.0) last_nz = b;
    }
    const double total = acc;
    if (!(total > 0.0)) return e->fail(SQ_ERR_ARG, "the state has zero norm");

    he = hipMemcpy(d_incl, I.data(), n_blocks * sizeof(double), hipMemcpyHostToDevice);
    if (he == hipSuccess) he = hipMemcpy(d_excl, C.data(), n_blocks * sizeof(double), hipMemcpyHostToDevice);
    if (he != hipSuccess) return e->hip_fail(he, "block prefixes");
    for (uint64_t k0 = 0; k0 < shots; k0 += batch) {
        const uint64_t cnt = std::min(batch, shots - k0);
        SampleLocateArgs la{};
        la.seed = seed;
        la.k0 = k0;
        la.n_shots = cnt;
        la.n_blocks = n_blocks;
        la.last_nz = last_nz;
        la.total = total;
        la.incl = d_incl;
        la.excl = d_excl;
        la.shots = d_shots;
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sample_locate), (unsigned)((cnt + kBlock - 1) / kBlock), kBlock, 0,
                           e->stream, la);
        rc = e->check_launch("k_sample_locate");
        if (rc) return rc;
        const unsigned g2 = (unsigned)((cnt + kBlock / 64 - 1) / (kBlock / 64));
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_sample_find<T>), g2, kBlock, 0, e->stream, s, d_shots, cnt, spb, d_out);
        rc = e->check_launch("k_sample_find");
        if (rc) return rc;
        he = hipMemcpyAsync(out + k0, d_out, cnt * sizeof(uint64_t), hipMemcpyDeviceToHost, e->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(e->stream);
        if (he != hipSuccess) return e->hip_fail(he, "shot results");
    }
    return SQ_OK;
}

}  // namespace

extern "C" int sq_sample(sq_engine *e, uint64_t seed, uint64_t shots, uint64_t *out) {
    if (!e || (!out && shots)) return SQ_ERR_ARG;
    if (shots == 0) return SQ_OK;
    if (e->out_of_core()) return e->fail(SQ_ERR_UNSUPPORTED, "sampling needs an HBM-resident state");
    if (e->n < 4) return e->fail(SQ_ERR_ARG, "sampling needs at least 4 qubits");
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    return e->dtype == SQ_C128 ? run_sample<double>(e, seed, shots, out) : run_sample<float>(e, seed, shots, out);
}
