// sq_engine.hip -- the MI355X state-vector engine behind include/subquokka_hip.h.
//
// One engine = one state (in HBM) + one HIP stream on one device.  The
// reference's MEM_Runner (MEM_Runner.cpp:57-185) owns a vector<complex<double>>
// and sweeps it with OpenMP threads; here every gate (or subcircuit block) is
// one kernel launch over the HBM-resident state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "host/gate_lowering.hpp"
#include "engine_internal.hpp"
#include "kernels/sweep.hpp"
#include "kernels/reduce.hpp"

using namespace sqk;
using sq::LoweredGate;
using sq::OpClass;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

int sq_engine::fail(int code, const std::string &msg) {
    err = msg;
    return code;
}

int sq_engine::hip_fail(hipError_t e, const char *what) {
    err = std::string(what) + ": " + hipGetErrorString(e);
    return SQ_ERR_HIP;
}

#define SQ_HIP(call)                                      \
    do {                                                  \
        hipError_t e__ = (call);                          \
        if (e__ != hipSuccess) return e->hip_fail(e__, #call); \
    } while (0)

// ---------------------------------------------------------------------------
// launch bookkeeping: optional HIP events around each launch
// ---------------------------------------------------------------------------

int sq_engine::stat_begin(const char *name, double alg_bytes, double alg_flops) {
    if (!stats_on) return -1;
    PendingStat ps;
    std::snprintf(ps.name, sizeof ps.name, "%s", name);
    ps.alg_bytes = alg_bytes;
    ps.alg_flops = alg_flops;
    if (hipEventCreate(&ps.start) != hipSuccess || hipEventCreate(&ps.stop) != hipSuccess) return -1;
    hipEventRecord(ps.start, stream);
    pending.push_back(ps);
    return (int)pending.size() - 1;
}

void sq_engine::stat_end(int slot) {
    if (slot < 0) return;
    hipEventRecord(pending[slot].stop, stream);
}

int sq_engine::check_launch(const char *what) {
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(le, what);
    return SQ_OK;
}

unsigned sq_engine::grid_for(uint64_t work_items, unsigned per_block) const {
    uint64_t b = (work_items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    if (b > max_blocks) b = max_blocks;
    return (unsigned)b;
}

// One work item per wave/thread, no grid-stride iterations: on MI355X a flat
// grid of short-lived workgroups streams HBM faster than a persistent grid
// (measured 5.5-6.0 TB/s against 4.7-5.3 TB/s for the one-qubit sweep).
unsigned sq_engine::grid_flat(uint64_t work_items, unsigned per_block) const {
    uint64_t b = (work_items + per_block - 1) / per_block;
    if (b < 1) b = 1;
    // A launch holds fewer than 2^32 threads (kBlock per workgroup): a
    // one-qubit gate on a lane bit of a 32-qubit complex128 state asks for
    // 2^24 workgroups, exactly 2^32.  Kernels launched here either keep a
    // grid-stride loop or (register and fused passes) stay far below the cap
    // for any state that fits in HBM.
    const uint64_t cap = 0xffffffffULL / kBlock;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// ---------------------------------------------------------------------------
// gate dispatch
// ---------------------------------------------------------------------------

namespace {

template <typename T>
void fill_matrix(cx<T> *dst, const std::complex<double> *m, int n) {
    for (int i = 0; i < n; i++) dst[i] = {(T)m[i].real(), (T)m[i].imag()};
}

// One wave tile per wave, flat grid.  Two adjacent tiles per wave (more
// loads in flight per lane) measured 3.7% slower averaged over the 30 targets
// of a 30-qubit complex128 state (6.02 against 5.80 ms per gate).
template <typename T, int KV, int KL, int KH>
void launch_dense(sq_engine *e, cx<T> *s, const DenseArgs<T> &a, unsigned) {
    constexpr int U = 1;
    const unsigned grid = e->grid_flat(a.n_tiles, (kBlock / 64) * U);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_dense<T, KV, KL, KH, U>), grid, kBlock, 0, e->stream, s, a);
}

// Dense 1q/2q gate on sorted targets t[0..k-1].
template <typename T>
int apply_dense(sq_engine *e, const LoweredGate &g) {
    constexpr int VB = Traits<T>::VB;
    constexpr int S = VB + 6;
    cx<T> *s = reinterpret_cast<cx<T> *>(e->d_state);
    const int n = e->n;
    const int k = g.k;
    char nm[48];
    std::snprintf(nm, sizeof nm, "dense%d_%s", k, e->dtype == SQ_C128 ? "c128" : "c64");
    const double bytes = 2.0 * e->amp_bytes() * std::ldexp(1.0, n) * g.pass_fraction;

    if (n < S) {   // smaller than one wave tile
        SmallArgs<T> a{};
        fill_matrix(a.m, g.m, 1 << (2 * k));
        a.k = k;
        for (int j = 0; j < k; j++) a.t[j] = g.t[j];
        a.n_groups = 1ULL << (n - k);
        int slot = e->stat_begin("dense_small", bytes);
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_dense_small<T>), e->grid_for(a.n_groups, kBlock), kBlock, 0, e->stream, s, a);
        e->stat_end(slot);
        return e->check_launch("k_dense_small");
    }
    DenseArgs<T> a{};
    fill_matrix(a.m, g.m, 1 << (2 * k));
    int kv = 0, kl = 0, kh = 0;
    for (int j = 0; j < k; j++) {
        const int t = g.t[j];
        if (t < VB) kv++;
        else if (t < S) a.lmask[kl++] = 1 << (t - VB);
        else a.hpos[kh++] = t;
    }
    a.n_tiles = 1ULL << (n - S - kh);
    const unsigned grid = e->grid_flat(a.n_tiles, kBlock / 64);
    int slot = e->stat_begin(nm, bytes);
    const int code = kv * 100 + kl * 10 + kh;
    switch (code) {
    case 100: launch_dense<T, 1, 0, 0>(e, s, a, grid); break;
    case 10:  launch_dense<T, 0, 1, 0>(e, s, a, grid); break;
    case 1:   launch_dense<T, 0, 0, 1>(e, s, a, grid); break;
    case 110: launch_dense<T, 1, 1, 0>(e, s, a, grid); break;
    case 101: launch_dense<T, 1, 0, 1>(e, s, a, grid); break;
    case 20:  launch_dense<T, 0, 2, 0>(e, s, a, grid); break;
    case 11:  launch_dense<T, 0, 1, 1>(e, s, a, grid); break;
    case 2:   launch_dense<T, 0, 0, 2>(e, s, a, grid); break;
    default:
        e->stat_end(slot);
        return e->fail(SQ_ERR_ARG, "internal: unsupported dense shape");
    }
    e->stat_end(slot);
    return e->check_launch("k_dense");
}

template <typename T>
int apply_diag(sq_engine *e, const LoweredGate &g) {
    constexpr int VB = Traits<T>::VB;
    cx<T> *s = reinterpret_cast<cx<T> *>(e->d_state);
    const int n = e->n, k = g.k;
    DiagArgs<T> a{};
    fill_matrix(a.d, g.m, 1 << k);
    a.k = k;
    for (int j = 0; j < k; j++) a.t[j] = g.t[j];
    // target j can be fixed to 1 when every entry with bit j = 0 is the identity
    for (int j = 0; j < k; j++) {
        bool all_ones_on_zero = true;
        for (int l = 0; l < (1 << k); l++)
            if (!((l >> j) & 1) && ((g.diag_active >> l) & 1)) all_ones_on_zero = false;
        // Only bits at or above a 128-byte line are skipped: fixing a lower bit
        // would read and write half-used lines (no traffic saved, slower).
        if (all_ones_on_zero && g.t[j] >= Traits<T>::LINE && n - VB - a.nfixed > 0) a.fixed[a.nfixed++] = g.t[j];
    }
    std::sort(a.fixed, a.fixed + a.nfixed);
    a.n_vec = 1ULL << (n - VB - a.nfixed);
    char nm[48];
    std::snprintf(nm, sizeof nm, "diag%d_%s", k, e->dtype == SQ_C128 ? "c128" : "c64");
    // bytes actually rewritten: fraction of the state visited
    const double bytes = 2.0 * e->amp_bytes() * std::ldexp(1.0, n) * g.pass_fraction;
    int slot = e->stat_begin(nm, bytes);
    hipLaunchKernelGGL(HIP_KERNEL_NAME(k_diag<T>), e->grid_flat(a.n_vec, kBlock * 2), kBlock, 0, e->stream, s, a);
    e->stat_end(slot);
    return e->check_launch("k_diag");
}

template <typename T>
int apply_perm(sq_engine *e, const LoweredGate &g) {
    constexpr int VB = Traits<T>::VB;
    cx<T> *s = reinterpret_cast<cx<T> *>(e->d_state);
    const int n = e->n;
    const int q = 2 * g.n_pairs;
    bool vector_ok = n - VB - q >= 0;
    for (int j = 0; j < q; j++) if (g.t[j] < VB) vector_ok = false;
    if (!vector_ok) {
        // A pair touches bit 0 of a complex<float> vector: apply each pair as a
        // dense SWAP matrix (exact: entries are 0 and 1).
        for (int pr = 0; pr < g.n_pairs; pr++) {
            LoweredGate sw{};
            sw.cls = OpClass::Dense;
            sw.k = 2;
            sw.t[0] = std::min(g.t[2 * pr], g.t[2 * pr + 1]);
            sw.t[1] = std::max(g.t[2 * pr], g.t[2 * pr + 1]);
            sw.m[0] = 1; sw.m[6] = 1; sw.m[9] = 1; sw.m[15] = 1;
            sw.pass_fraction = 0.5;
            int rc = apply_dense<T>(e, sw);
            if (rc) return rc;
        }
        return SQ_OK;
    }
    PermArgs a{};
    for (int j = 0; j < q; j++) { a.p[j] = g.t[j]; a.sorted[j] = g.t[j]; }
    std::sort(a.sorted, a.sorted + q);
    a.n_vec = 1ULL << (n - VB - q);
    char nm[48];
    std::snprintf(nm, sizeof nm, "perm%d_%s", g.n_pairs, e->dtype == SQ_C128 ? "c128" : "c64");
    const double bytes = 2.0 * e->amp_bytes() * std::ldexp(1.0, n) * g.pass_fraction;
    int slot = e->stat_begin(nm, bytes);
    if (g.n_pairs == 1) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_perm<T, 1>), e->grid_flat(a.n_vec, kBlock), kBlock, 0, e->stream, s, a);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_perm<T, 2>), e->grid_flat(a.n_vec, kBlock), kBlock, 0, e->stream, s, a);
    e->stat_end(slot);
    return e->check_launch("k_perm");
}

template <typename T>
int apply_lowered_t(sq_engine *e, const LoweredGate &g) {
    switch (g.cls) {
    case OpClass::Dense: return apply_dense<T>(e, g);
    case OpClass::Diag: return apply_diag<T>(e, g);
    case OpClass::Perm: return apply_perm<T>(e, g);
    }
    return e->fail(SQ_ERR_ARG, "internal: bad op class");
}

}  // namespace

int sq_engine::apply_lowered(const LoweredGate &g) {
    return dtype == SQ_C128 ? apply_lowered_t<double>(this, g) : apply_lowered_t<float>(this, g);
}

int sq_engine::lower(const sq_gate &g, LoweredGate &lg) {
    std::string why;
    int rc = sq::lower_gate(g, n, lg, &why);
    if (rc) return fail(rc, why);
    return SQ_OK;
}

// ---------------------------------------------------------------------------
// reductions (partials summed on the host in block order)
// ---------------------------------------------------------------------------

int sq_engine::ensure_partials() {
    if (d_part) return SQ_OK;
    hipError_t he = hipMalloc(&d_part, sizeof(double) * 4 * kRedBlocks);
    if (he != hipSuccess) return hip_fail(he, "hipMalloc(partials)");
    return SQ_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

namespace {

// sq_create / sq_create_file_state: files == nullptr for memory tiers.
int create_engine(const sq_config *cfg, const char *const *files, int n_files, bool direct, bool keep,
                  sq_engine **out) {
    if (!cfg || !out) return SQ_ERR_ARG;
    *out = nullptr;
    if (cfg->total_qbit < 1 || cfg->total_qbit > 36) return SQ_ERR_ARG;
    if (cfg->dtype != SQ_C128 && cfg->dtype != SQ_C64) return SQ_ERR_ARG;
    if (cfg->chunk_qbit < 0 || cfg->chunk_qbit > cfg->total_qbit) return SQ_ERR_ARG;
    if (cfg->sweep_mode < 0 || cfg->sweep_mode > 1 || cfg->pass_qbits < 0 || cfg->pass_qbits > 5) return SQ_ERR_ARG;
    if (cfg->host_tier != SQ_HOST_AUTO && cfg->host_tier != SQ_HOST_PAGEABLE) return SQ_ERR_ARG;
    // a resident chunk at least as large as the state means the state fits in HBM
    const int resident = cfg->resident_qbit >= cfg->total_qbit ? 0 : cfg->resident_qbit;
    if (resident < 0) return SQ_ERR_ARG;
    if (files) {
        if (resident == 0 || n_files < 1 || (n_files & (n_files - 1)) || (1LL << cfg->total_qbit) < n_files)
            return SQ_ERR_ARG;
        for (int i = 0; i < n_files; i++) if (!files[i]) return SQ_ERR_ARG;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= cfg->device || cfg->device < 0) return SQ_ERR_HIP;
    auto *e = new sq_engine();
    e->n = cfg->total_qbit;
    e->chunk = cfg->chunk_qbit > 0 ? cfg->chunk_qbit : std::min(cfg->total_qbit, 12);
    e->dtype = cfg->dtype;
    e->device = cfg->device;
    e->sweep_mode = cfg->sweep_mode;
    // register qubits per pass: five for both state types (2^5 rows of 16-byte
    // vectors per wave fit the register file at three waves per SIMD without
    // spills; fewer, longer passes were measured faster than four)
    e->pass_qbits = cfg->pass_qbits > 0 ? cfg->pass_qbits : 5;
    if (hipSetDevice(e->device) != hipSuccess) { delete e; return SQ_ERR_HIP; }
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) { delete e; return SQ_ERR_HIP; }
    if (resident != 0) {
        // out-of-core runner: resident chunk of 2^resident amplitudes
        if (resident < 8 || resident > 34) { hipStreamDestroy(e->stream); delete e; return SQ_ERR_ARG; }
        e->res_qbit = resident;
        e->host_tier = cfg->host_tier;
        int rc = files ? e->stream_open_files(files, n_files, direct, keep) : SQ_OK;
        if (!rc) rc = e->stream_create();
        if (rc) {
            std::fprintf(stderr, "[sq_create]: %s\n", e->err.c_str());
            e->stream_destroy();
            hipStreamDestroy(e->stream);
            delete e;
            return rc;
        }
    } else if (hipMalloc(&e->d_state, e->state_bytes()) != hipSuccess) {
        hipGetLastError();
        hipStreamDestroy(e->stream);
        delete e;
        return SQ_ERR_OOM;
    }
    // the device address of the no-allocation entry point's scratch array
    // (once per device; every engine on the device shares it)
    if (int rc = e->setup_scratch()) {
        std::fprintf(stderr, "[sq_create]: %s\n", e->err.c_str());
        if (e->out_of_core()) e->stream_destroy();
        if (e->d_state) hipFree(e->d_state);
        hipStreamDestroy(e->stream);
        delete e;
        return rc;
    }
    int cus = 256;
    hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device);
    e->max_blocks = (unsigned)std::max(1, cus) * 8;
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, e->device) == hipSuccess && lds > 0)
        e->lds_per_block = (size_t)lds;
    if (!(files && keep)) {
        const int rc = sq_init_zero(e);
        if (rc) {
            std::fprintf(stderr, "[sq_create]: %s\n", e->err.c_str());
            sq_destroy(e);
            return rc;
        }
    }
    *out = e;
    return SQ_OK;
}

}  // namespace

extern "C" {

int sq_create(const sq_config *cfg, sq_engine **out) { return create_engine(cfg, nullptr, 0, false, false, out); }

int sq_create_file_state(const sq_config *cfg, const char *const *paths, int32_t n_files, int32_t direct_io,
                         int32_t keep_contents, sq_engine **out) {
    if (!paths) return SQ_ERR_ARG;
    return create_engine(cfg, paths, n_files, direct_io != 0, keep_contents != 0, out);
}

int sq_destroy(sq_engine *e) {
    if (!e) return SQ_ERR_ARG;
    hipSetDevice(e->device);
    if (e->stream) hipStreamSynchronize(e->stream);
    sq_stats_reset(e);
    if (e->out_of_core()) e->stream_destroy();
    if (e->d_state) hipFree(e->d_state);
    if (e->d_alt) hipFree(e->d_alt);
    if (e->d_part) hipFree(e->d_part);
    if (e->d_stage) hipFree(e->d_stage);
    e->free_extras();
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return SQ_OK;
}

const char *sq_last_error(const sq_engine *e) { return e ? e->err.c_str() : "null engine"; }

int sq_init_zero(sq_engine *e) {
    if (!e) return SQ_ERR_ARG;
    SQ_HIP(hipSetDevice(e->device));
    if (e->out_of_core()) return e->stream_init_zero();
    const uint64_t n = e->n_amps();
    if (e->dtype == SQ_C128)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fill_zero_one<double>), e->grid_for(n, kBlock), kBlock, 0, e->stream, reinterpret_cast<cx<double> *>(e->d_state), n);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_fill_zero_one<float>), e->grid_for(n, kBlock), kBlock, 0, e->stream, reinterpret_cast<cx<float> *>(e->d_state), n);
    return e->check_launch("k_fill_zero_one");
}

int sq_synchronize(sq_engine *e) {
    if (!e) return SQ_ERR_ARG;
    SQ_HIP(hipStreamSynchronize(e->stream));
    return SQ_OK;
}

int sq_device_state(sq_engine *e, void **dptr, void **stream) {
    if (!e) return SQ_ERR_ARG;
    if (e->out_of_core()) return e->fail(SQ_ERR_ARG, "the out-of-core state is in host memory");
    if (dptr) *dptr = e->d_state;
    if (stream) *stream = (void *)e->stream;
    return SQ_OK;
}

// Host <-> device state transfer.  Same dtype: one hipMemcpy.  Different dtype:
// staged through a device buffer in slices and converted by a kernel.
int sq_set_state(sq_engine *e, const void *host, uint64_t n_amps, int32_t dtype) {
    if (!e || !host) return SQ_ERR_ARG;
    if (n_amps != e->n_amps()) return e->fail(SQ_ERR_ARG, "sq_set_state: n_amps must be 2^total_qbit");
    if (dtype != SQ_C128 && dtype != SQ_C64) return e->fail(SQ_ERR_ARG, "bad dtype");
    SQ_HIP(hipSetDevice(e->device));
    if (e->out_of_core()) return e->stream_set_state(host, dtype);
    if (dtype == e->dtype) {
        SQ_HIP(hipMemcpyAsync(e->d_state, host, e->state_bytes(), hipMemcpyHostToDevice, e->stream));
        SQ_HIP(hipStreamSynchronize(e->stream));
        return SQ_OK;
    }
    const size_t hb = dtype == SQ_C128 ? 16 : 8;
    const uint64_t slice = std::min<uint64_t>(n_amps, 1ULL << 24);
    if (!e->d_stage) SQ_HIP(hipMalloc(&e->d_stage, slice * 16));
    for (uint64_t off = 0; off < n_amps; off += slice) {
        const uint64_t cnt = std::min(slice, n_amps - off);
        SQ_HIP(hipMemcpyAsync(e->d_stage, (const char *)host + off * hb, cnt * hb, hipMemcpyHostToDevice, e->stream));
        if (dtype == SQ_C128)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<double, float>), e->grid_for(cnt, kBlock), kBlock, 0, e->stream, 
                (const cx<double> *)e->d_stage, (cx<float> *)e->d_state + off, cnt);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<float, double>), e->grid_for(cnt, kBlock), kBlock, 0, e->stream, 
                (const cx<float> *)e->d_stage, (cx<double> *)e->d_state + off, cnt);
        SQ_HIP(hipStreamSynchronize(e->stream));
    }
    return e->check_launch("k_convert");
}

int sq_get_state(sq_engine *e, void *host, uint64_t n_amps, int32_t dtype) {
    if (!e || !host) return SQ_ERR_ARG;
    if (n_amps != e->n_amps()) return e->fail(SQ_ERR_ARG, "sq_get_state: n_amps must be 2^total_qbit");
    if (dtype != SQ_C128 && dtype != SQ_C64) return e->fail(SQ_ERR_ARG, "bad dtype");
    SQ_HIP(hipSetDevice(e->device));
    if (e->out_of_core()) return e->stream_get_state(host, dtype);
    if (dtype == e->dtype) {
        SQ_HIP(hipMemcpyAsync(host, e->d_state, e->state_bytes(), hipMemcpyDeviceToHost, e->stream));
        SQ_HIP(hipStreamSynchronize(e->stream));
        return SQ_OK;
    }
    const size_t hb = dtype == SQ_C128 ? 16 : 8;
    const uint64_t slice = std::min<uint64_t>(n_amps, 1ULL << 24);
    if (!e->d_stage) SQ_HIP(hipMalloc(&e->d_stage, slice * 16));
    for (uint64_t off = 0; off < n_amps; off += slice) {
        const uint64_t cnt = std::min(slice, n_amps - off);
        if (dtype == SQ_C128)
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<float, double>), e->grid_for(cnt, kBlock), kBlock, 0, e->stream, 
                (const cx<float> *)e->d_state + off, (cx<double> *)e->d_stage, cnt);
        else
            hipLaunchKernelGGL(HIP_KERNEL_NAME(k_convert<double, float>), e->grid_for(cnt, kBlock), kBlock, 0, e->stream, 
                (const cx<double> *)e->d_state + off, (cx<float> *)e->d_stage, cnt);
        SQ_HIP(hipMemcpyAsync((char *)host + off * hb, e->d_stage, cnt * hb, hipMemcpyDeviceToHost, e->stream));
        SQ_HIP(hipStreamSynchronize(e->stream));
    }
    return e->check_launch("k_convert");
}

int sq_apply_gates(sq_engine *e, const sq_gate *gates, uint64_t n) {
    if (!e || (!gates && n)) return SQ_ERR_ARG;
    SQ_HIP(hipSetDevice(e->device));
    // Validate everything first so a bad line leaves the state untouched.
    std::vector<LoweredGate> low(n);
    for (uint64_t i = 0; i < n; i++) {
        int rc = e->lower(gates[i], low[i]);
        if (rc) { e->err += " (gate " + std::to_string(i) + ")"; return rc; }
    }
    if (e->out_of_core()) return e->stream_apply(std::vector<sq_gate>(gates, gates + n), low);
    return e->apply_plain(std::vector<sq_gate>(gates, gates + n), low);
}

int sq_norm2(sq_engine *e, double *out) {
    double v[2];
    int rc = e ? e->norm_zz(v) : SQ_ERR_ARG;
    if (rc) return rc;
    if (out) *out = v[0];
    return SQ_OK;
}

int sq_expect_zz_ring(sq_engine *e, double *out) {
    double v[2];
    int rc = e ? e->norm_zz(v) : SQ_ERR_ARG;
    if (rc) return rc;
    if (out) *out = v[1];
    return SQ_OK;
}

int sq_inner(sq_engine *a, sq_engine *b, double *re, double *im) {
    double v[3];
    if (!a || !b) return SQ_ERR_ARG;
    int rc = a->compare(b, v);
    if (rc) return rc;
    if (re) *re = v[0];
    if (im) *im = v[1];
    return SQ_OK;
}

int sq_max_abs_diff(sq_engine *a, sq_engine *b, double *out) {
    double v[3];
    if (!a || !b) return SQ_ERR_ARG;
    int rc = a->compare(b, v);
    if (rc) return rc;
    if (out) *out = v[2];
    return SQ_OK;
}

int sq_stats_enable(sq_engine *e, int32_t on) {
    if (!e) return SQ_ERR_ARG;
    e->stats_on = on != 0;
    return SQ_OK;
}

int sq_stats_reset(sq_engine *e) {
    if (!e) return SQ_ERR_ARG;
    hipStreamSynchronize(e->stream);
    for (auto &p : e->pending) { hipEventDestroy(p.start); hipEventDestroy(p.stop); }
    e->pending.clear();
    e->n_stream_passes = 0;
    return SQ_OK;
}

int sq_stream_info(sq_engine *e, uint64_t *passes, double *host_link_bytes, int32_t *host_tier) {
    if (!e) return SQ_ERR_ARG;
    if (host_tier) *host_tier = e->out_of_core() ? e->host_tier : SQ_HOST_AUTO;
    if (passes) *passes = e->n_stream_passes;
    // every pass moves the whole state host -> HBM and back
    if (host_link_bytes) *host_link_bytes = 2.0 * (double)e->n_stream_passes * (double)e->n_amps() * (double)e->amp_bytes();
    return SQ_OK;
}

int sq_stats_get(sq_engine *e, sq_kstat *out, int32_t cap, int32_t *n_out) {
    if (!e) return SQ_ERR_ARG;
    SQ_HIP(hipStreamSynchronize(e->stream));
    std::vector<sq_kstat> agg;
    for (auto &p : e->pending) {
        float ms = 0;
        SQ_HIP(hipEventElapsedTime(&ms, p.start, p.stop));
        auto it = std::find_if(agg.begin(), agg.end(), [&](const sq_kstat &k) { return std::strcmp(k.name, p.name) == 0; });
        if (it == agg.end()) {
            sq_kstat k{};
            std::snprintf(k.name, sizeof k.name, "%s", p.name);
            agg.push_back(k);
            it = agg.end() - 1;
        }
        it->launches++;
        it->total_ms += ms;
        it->alg_bytes += p.alg_bytes;
        it->alg_flops += p.alg_flops;
    }
    if (n_out) *n_out = (int32_t)agg.size();
    for (int i = 0; i < (int)agg.size() && i < cap; i++) out[i] = agg[i];
    return (int)agg.size() > cap ? SQ_ERR_ARG : SQ_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// reduction members
// ---------------------------------------------------------------------------

int sq_engine::norm_zz(double *v) {
    if (out_of_core()) return stream_norm_zz(v);
    int rc = ensure_partials();
    if (rc) return rc;
    hipSetDevice(device);
    const uint64_t na = n_amps();
    const unsigned grid = (unsigned)std::min<uint64_t>(kRedBlocks, (na + kBlock - 1) / kBlock);
    if (dtype == SQ_C128)
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_norm_zz<double>), grid, kBlock, 0, stream, (const cx<double> *)d_state, na, n, d_part);
    else
        hipLaunchKernelGGL(HIP_KERNEL_NAME(k_norm_zz<float>), grid, kBlock, 0, stream, (const cx<float> *)d_state, na, n, d_part);
    rc = check_launch("k_norm_zz");
    if (rc) return rc;
    std::vector<double> h(2 * grid);
    hipError_t he = hipMemcpyAsync(h.data(), d_part, sizeof(double) * 2 * grid, hipMemcpyDeviceToHost, stream);
    if (he == hipSuccess) he = hipStreamSynchronize(stream);
    if (he != hipSuccess) return hip_fail(he, "norm partials");
    v[0] = v[1] = 0;
    for (unsigned b = 0; b < grid; b++) { v[0] += h[2 * b]; v[1] += h[2 * b + 1]; }
    return SQ_OK;
}

int sq_engine::compare(sq_engine *o, double *v) {
    if (o->n != n || o->device != device) return fail(SQ_ERR_ARG, "engines differ in size or device");
    if (out_of_core() || o->out_of_core()) return fail(SQ_ERR_UNSUPPORTED, "compare needs HBM-resident states");
    int rc = ensure_partials();
    if (rc) return rc;
    hipSetDevice(device);
    // order the other engine's pending work before ours
    hipEvent_t ev;
    hipEventCreate(&ev);
    hipEventRecord(ev, o->stream);
    hipStreamWaitEvent(stream, ev, 0);
    const uint64_t na = n_amps();
    const unsigned grid = (unsigned)std::min<uint64_t>(kRedBlocks, (na + kBlock - 1) / kBlock);
    const bool a64 = dtype == SQ_C128, b64 = o->dtype == SQ_C128;
    if (a64 && b64) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compare<double, double>), grid, kBlock, 0, stream, (const cx<double> *)d_state, (const cx<double> *)o->d_state, na, d_part);
    else if (a64) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compare<double, float>), grid, kBlock, 0, stream, (const cx<double> *)d_state, (const cx<float> *)o->d_state, na, d_part);
    else if (b64) hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compare<float, double>), grid, kBlock, 0, stream, (const cx<float> *)d_state, (const cx<double> *)o->d_state, na, d_part);
    else hipLaunchKernelGGL(HIP_KERNEL_NAME(k_compare<float, float>), grid, kBlock, 0, stream, (const cx<float> *)d_state, (const cx<float> *)o->d_state, na, d_part);
    rc = check_launch("k_compare");
    hipEventDestroy(ev);
    if (rc) return rc;
    std::vector<double> h(3 * grid);
    hipError_t he = hipMemcpyAsync(h.data(), d_part, sizeof(double) * 3 * grid, hipMemcpyDeviceToHost, stream);
    if (he == hipSuccess) he = hipStreamSynchronize(stream);
    if (he != hipSuccess) return hip_fail(he, "compare partials");
    v[0] = v[1] = v[2] = 0;
    for (unsigned b = 0; b < grid; b++) { v[0] += h[3 * b]; v[1] += h[3 * b + 1]; v[2] = std::max(v[2], h[3 * b + 2]); }
    return SQ_OK;
}
