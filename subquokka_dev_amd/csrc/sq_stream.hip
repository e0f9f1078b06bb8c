// sq_stream.hip -- out-of-core runner: the state in pinned host memory,
// streamed through HBM.  This is the one-GPU form of the reference's
// DIO_Runner (DIO_Runner.cpp:15-106, 314-...; state in 2^file_qbit O_DIRECT
// files, chunks read with pread, gated, written with pwrite, and gates on
// "file" qubits pairing the chunk streams of 2 or 4 files,
// MPI_DIO_Func.cpp:22-157).
//
// Layout: amplitude i of the 2^n state at host index i.  A pass of the plan
// works on a qubit set P of exactly W = resident_qbit qubits: the low W-h
// qubits plus h qubits chosen by the planner (the finder's chunk-set choice,
// finder.cpp:651-698, here with W-h fixed low qubits).  The 2^(n-W) groups of a
// pass (one per value of the qubits outside P) are each 2^h contiguous runs of
// 2^(W-h) amplitudes; a group is copied into an HBM buffer with 2^h
// hipMemcpyAsync calls, the pass's gates run on it as on a W-qubit state
// (qubit P[j] -> bit j) with the ordinary sweep kernels, and it is copied
// back.  Two buffers and three streams overlap upload, compute and download.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "host/planner.hpp"

using sq::LoweredGate;

namespace {

constexpr int kMaxHigh = 6;   // at most 2^6 runs per group

template <typename F>
void parallel_for(uint64_t n, F f) {
    unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned k = 0; k < t; k++)
        th.emplace_back([=] { f(n * k / t, n * (k + 1) / t); });
    for (auto &x : th) x.join();
}

// Remaps a lowered gate from global qubits to bit positions inside P.
LoweredGate remap(const LoweredGate &g, const std::vector<int> &P) {
    LoweredGate r = g;
    const int cnt = g.cls == sq::OpClass::Perm ? 2 * g.n_pairs : g.k;
    for (int j = 0; j < cnt; j++)
        r.t[j] = (int)(std::lower_bound(P.begin(), P.end(), g.t[j]) - P.begin());
    return r;
}

}  // namespace

int sq_engine::stream_create() {
    const size_t host_bytes = (size_t)1 << n;
    hipError_t he = hipHostMalloc(&h_state, host_bytes * amp_bytes(), hipHostMallocDefault);
    if (he != hipSuccess) { h_state = nullptr; hipGetLastError(); return fail(SQ_ERR_OOM, "hipHostMalloc(state)"); }
    const size_t buf_bytes = ((size_t)1 << res_qbit) * amp_bytes();
    for (int b = 0; b < 2; b++) {
        he = hipMalloc(&d_group[b], buf_bytes);
        if (he != hipSuccess) { hipGetLastError(); return fail(SQ_ERR_OOM, "hipMalloc(group buffer)"); }
        hipEventCreate(&ev_loaded[b]);
        hipEventCreate(&ev_computed[b]);
        hipEventCreate(&ev_stored[b]);
    }
    hipEventCreate(&ev_pass);
    if (hipStreamCreateWithFlags(&s_h2d, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s_d2h, hipStreamNonBlocking) != hipSuccess)
        return fail(SQ_ERR_HIP, "stream create");
    return SQ_OK;
}

void sq_engine::stream_destroy() {
    if (s_h2d) hipStreamSynchronize(s_h2d);
    if (s_d2h) hipStreamSynchronize(s_d2h);
    for (int b = 0; b < 2; b++) {
        if (d_group[b]) hipFree(d_group[b]);
        if (ev_loaded[b]) hipEventDestroy(ev_loaded[b]);
        if (ev_computed[b]) hipEventDestroy(ev_computed[b]);
        if (ev_stored[b]) hipEventDestroy(ev_stored[b]);
        d_group[b] = nullptr;
    }
    if (ev_pass) hipEventDestroy(ev_pass);
    ev_pass = nullptr;
    if (h_state) hipHostFree(h_state);
    h_state = nullptr;
    if (s_h2d) hipStreamDestroy(s_h2d);
    if (s_d2h) hipStreamDestroy(s_d2h);
    s_h2d = s_d2h = nullptr;
}

int sq_engine::stream_init_zero() {
    char *p = static_cast<char *>(h_state);
    const size_t total = ((size_t)1 << n) * amp_bytes();
    parallel_for(total, [&](uint64_t a, uint64_t b) { std::memset(p + a, 0, b - a); });
    if (dtype == SQ_C128) static_cast<double *>(h_state)[0] = 1.0;
    else static_cast<float *>(h_state)[0] = 1.0f;
    return SQ_OK;
}

int sq_engine::stream_set_state(const void *host, int32_t host_dtype) {
    const uint64_t na = n_amps();
    if (host_dtype == dtype) {
        const size_t total = na * amp_bytes();
        char *d = static_cast<char *>(h_state);
        const char *s = static_cast<const char *>(host);
        parallel_for(total, [&](uint64_t a, uint64_t b) { std::memcpy(d + a, s + a, b - a); });
    } else if (host_dtype == SQ_C128) {
        const double *s = static_cast<const double *>(host);
        float *d = static_cast<float *>(h_state);
        parallel_for(2 * na, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = (float)s[i]; });
    } else {
        const float *s = static_cast<const float *>(host);
        double *d = static_cast<double *>(h_state);
        parallel_for(2 * na, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = s[i]; });
    }
    return SQ_OK;
}

int sq_engine::stream_get_state(void *host, int32_t host_dtype) {
    const uint64_t na = n_amps();
    if (host_dtype == dtype) {
        const size_t total = na * amp_bytes();
        char *d = static_cast<char *>(host);
        const char *s = static_cast<const char *>(h_state);
        parallel_for(total, [&](uint64_t a, uint64_t b) { std::memcpy(d + a, s + a, b - a); });
    } else if (host_dtype == SQ_C128) {
        const float *s = static_cast<const float *>(h_state);
        double *d = static_cast<double *>(host);
        parallel_for(2 * na, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = s[i]; });
    } else {
        const double *s = static_cast<const double *>(h_state);
        float *d = static_cast<float *>(host);
        parallel_for(2 * na, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = (float)s[i]; });
    }
    return SQ_OK;
}

// FP64 sums of |a|^2 and |a|^2 zz(i) over the host state.
int sq_engine::stream_norm_zz(double *v) {
    const uint64_t na = n_amps();
    const uint64_t mask = (n >= 64) ? ~0ULL : ((1ULL << n) - 1);
    const unsigned t = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<double> s0(t, 0.0), s1(t, 0.0);
    std::vector<std::thread> th;
    for (unsigned k = 0; k < t; k++)
        th.emplace_back([&, k] {
            const uint64_t a = na * k / t, b = na * (k + 1) / t;
            double x0 = 0, x1 = 0;
            for (uint64_t i = a; i < b; i++) {
                double re, im;
                if (dtype == SQ_C128) { re = static_cast<double *>(h_state)[2 * i]; im = static_cast<double *>(h_state)[2 * i + 1]; }
                else { re = static_cast<float *>(h_state)[2 * i]; im = static_cast<float *>(h_state)[2 * i + 1]; }
                const double p = re * re + im * im;
                const uint64_t rot = ((i >> 1) | (i << (n - 1))) & mask;
                x0 += p;
                x1 += p * (double)(n - 2 * __builtin_popcountll((i ^ rot) & mask));
            }
            s0[k] = x0; s1[k] = x1;
        });
    for (auto &x : th) x.join();
    v[0] = v[1] = 0;
    for (unsigned k = 0; k < t; k++) { v[0] += s0[k]; v[1] += s1[k]; }
    return SQ_OK;
}

int sq_engine::stream_apply(const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low) {
    const int W = res_qbit;
    const int h = std::min(kMaxHigh, std::min(W - 1, n - W + 2));   // free slots above the fixed low run
    std::vector<int> fixed;
    for (int q = 0; q < W - h; q++) fixed.push_back(q);
    const auto passes = sq::plan_passes(gates, n, W, fixed);

    const size_t B = amp_bytes();
    const uint64_t run = 1ULL << (W - h);
    const uint64_t n_groups = 1ULL << (n - W);
    void *const saved_state = d_state;
    const int saved_n = n;
    int rc = SQ_OK;
    for (const auto &pass : passes) {
        // groups of different passes overlap in host memory: every upload of
        // this pass waits for every download of the previous one
        hipEventRecord(ev_pass, s_d2h);
        hipStreamWaitEvent(s_h2d, ev_pass, 0);
        // P = fixed low run + exactly h further qubits (pad with unused ones)
        std::vector<char> in(n, 0);
        for (int q : pass.qubits) in[q] = 1;
        for (int q : fixed) in[q] = 1;
        int size = (int)std::count(in.begin(), in.end(), 1);
        for (int q = W - h; q < n && size < W; q++) if (!in[q]) { in[q] = 1; size++; }
        std::vector<int> P, H;
        for (int q = 0; q < n; q++) if (in[q]) { P.push_back(q); if (q >= W - h) H.push_back(q); }
        if ((int)P.size() != W || (int)H.size() != h) { rc = fail(SQ_ERR_ARG, "internal: bad pass qubit set"); break; }
        std::vector<LoweredGate> pg;
        for (int gi : pass.gates) pg.push_back(remap(low[gi], P));
        std::vector<uint64_t> run_off(1ULL << h);   // host offset of run r within a group
        for (uint64_t r = 0; r < run_off.size(); r++) {
            uint64_t o = 0;
            for (int j = 0; j < h; j++) if ((r >> j) & 1) o += 1ULL << H[j];
            run_off[r] = o;
        }
        char *host = static_cast<char *>(h_state);
        for (uint64_t g = 0; g < n_groups; g++) {
            const int b = (int)(g & 1);
            uint64_t base = g;
            for (int q : P) base = ((base >> q) << (q + 1)) | (base & ((1ULL << q) - 1));
            char *dbuf = static_cast<char *>(d_group[b]);
            // the buffer's previous download must finish before it is refilled
            hipStreamWaitEvent(s_h2d, ev_stored[b], 0);
            for (uint64_t r = 0; r < run_off.size(); r++)
                hipMemcpyAsync(dbuf + r * run * B, host + (base + run_off[r]) * B, run * B, hipMemcpyHostToDevice, s_h2d);
            hipEventRecord(ev_loaded[b], s_h2d);
            hipStreamWaitEvent(stream, ev_loaded[b], 0);
            d_state = d_group[b];
            n = W;
            for (const auto &lg : pg) {
                rc = apply_lowered(lg);
                if (rc) break;
            }
            n = saved_n;
            d_state = saved_state;
            if (rc) break;
            hipEventRecord(ev_computed[b], stream);
            hipStreamWaitEvent(s_d2h, ev_computed[b], 0);
            for (uint64_t r = 0; r < run_off.size(); r++)
                hipMemcpyAsync(host + (base + run_off[r]) * B, dbuf + r * run * B, run * B, hipMemcpyDeviceToHost, s_d2h);
            hipEventRecord(ev_stored[b], s_d2h);
        }
        if (rc) break;
        n_stream_passes++;
    }
    d_state = saved_state;
    n = saved_n;
    hipError_t he = hipStreamSynchronize(s_d2h);
    if (he == hipSuccess) he = hipStreamSynchronize(stream);
    if (!rc && he != hipSuccess) rc = hip_fail(he, "out-of-core pass");
    return rc;
}
