// sq_stream.hip -- out-of-core runner: the state outside HBM, streamed through
// it pass by pass.  This is the one-GPU form of the reference's DIO_Runner
// (DIO_Runner.cpp:15-106, 314-...; the state in 2^file_qbit O_DIRECT files,
// chunks read with pread, gated, written with pwrite; gates on "file" qubits
// pair the chunk streams of 2 or 4 files, MPI_DIO_Func.cpp:22-157).
//
// Where the state lives (the host tier):
//   * PINNED:   one pinned host allocation; groups move by DMA straight from
//               and to it (2^h hipMemcpyAsync per group and direction);
//   * PAGEABLE: ordinary host memory (for states larger than the pinnable
//               limit); CPU threads gather a group's runs into a pinned
//               staging buffer, one DMA moves the packed group, and the
//               result is scattered back while the GPU works on the next one;
//   * FILES:    2^f files, file i holding amplitudes [i 2^(n-f), (i+1) 2^(n-f))
//               as raw little-endian (re, im) -- the reference's state-file
//               layout (test_util.py:385-403, setupFiles simulator.cpp:102-129)
//               -- read and written with pread/pwrite (O_DIRECT for DirectIO,
//               create_file simulator.cpp:71-86) into the same staging.
//
// A pass of the plan works on a qubit set P of exactly W = resident_qbit
// qubits: the low W-h qubits plus h qubits chosen by the planner (the
// finder's chunk-set choice, finder.cpp:651-698, here with W-h fixed low
// qubits).  The 2^(n-W) groups of a pass (one per value of the qubits outside
// P) are each 2^h contiguous runs of 2^(W-h) amplitudes; the pass's gates run
// on a group as on a W-qubit state (qubit P[j] -> bit j) with the ordinary
// kernels.  Two HBM buffers and three streams (upload, compute, download)
// overlap the groups.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"
#include "host/planner.hpp"

using sq::LoweredGate;

namespace {

constexpr int kMaxHigh = 6;          // at most 2^6 runs per group
constexpr size_t kIoAlign = 4096;    // O_DIRECT alignment (DIO_Runner.cpp:155-164)

unsigned host_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

template <typename F>
void parallel_for(uint64_t n, F f) {
    const unsigned t = host_threads();
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

// Whole-range pread/pwrite (short transfers continue).  On failure *err holds
// errno (EIO for an unexpected end of file).
bool pread_all(int fd, char *dst, size_t bytes, off_t off, int *err) {
    while (bytes) {
        const ssize_t r = pread(fd, dst, bytes, off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) { *err = r < 0 ? errno : EIO; return false; }
        dst += r; bytes -= (size_t)r; off += r;
    }
    return true;
}
bool pwrite_all(int fd, const char *src, size_t bytes, off_t off, int *err) {
    while (bytes) {
        const ssize_t r = pwrite(fd, src, bytes, off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) { *err = r < 0 ? errno : EIO; return false; }
        src += r; bytes -= (size_t)r; off += r;
    }
    return true;
}

}  // namespace

// Failure message of a state-file transfer, with the system's reason.
int sq_engine::io_fail(const char *what) {
    const int err = io_errno.load();
    std::string m = std::string(what) + ": " + std::strerror(err ? err : EIO);
    if (file_direct) m += " (O_DIRECT)";
    return fail(SQ_ERR_IO, m);
}

// Free slots of a pass above its fixed low run.  At least four, so that a
// VSWAP_2_2 whose four qubits all lie above the run still fits one pass
// (resident_qbit >= 8 keeps a fixed run of at least one qubit); beyond that,
// up to n - W + 2 (fewer, longer runs when few qubits lie outside the chunk).
int sq_engine::stream_high() const {
    return std::min(kMaxHigh, std::min(res_qbit - 1, std::max(4, n - res_qbit + 2)));
}

int sq_engine::stream_create() {
    const size_t state = state_bytes();
    const size_t buf_bytes = ((size_t)1 << res_qbit) * amp_bytes();
    hipError_t he;
    if (host_tier == SQ_HOST_FILES) {
        const int nf = (int)file_fds.size();
        const size_t per_file = state / (size_t)nf;
        if (file_direct && ((((size_t)1 << (res_qbit - stream_high())) * amp_bytes()) % kIoAlign || per_file % kIoAlign))
            return fail(SQ_ERR_ARG, "O_DIRECT needs runs and files in multiples of 4 KiB (raise resident_qbit)");
    } else if (host_tier == SQ_HOST_AUTO) {
        he = hipHostMalloc(&h_state, state, hipHostMallocDefault);
        if (he == hipSuccess) {
            host_tier = SQ_HOST_PINNED;
        } else {
            hipGetLastError();
            h_state = nullptr;
            host_tier = SQ_HOST_PAGEABLE;   // larger than the pinnable limit
        }
    }
    if (host_tier == SQ_HOST_PAGEABLE) {
        h_state = std::aligned_alloc(kIoAlign, (state + kIoAlign - 1) / kIoAlign * kIoAlign);
        if (!h_state) return fail(SQ_ERR_OOM, "host state allocation");
        h_state_pageable = true;
    }
    if (host_tier != SQ_HOST_PINNED) {
        for (int b = 0; b < 2; b++) {
            he = hipHostMalloc(&h_stage[b], buf_bytes, hipHostMallocDefault);
            if (he != hipSuccess) { hipGetLastError(); h_stage[b] = nullptr; return fail(SQ_ERR_OOM, "hipHostMalloc(staging)"); }
        }
    }
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
        if (h_stage[b]) hipHostFree(h_stage[b]);
        if (ev_loaded[b]) hipEventDestroy(ev_loaded[b]);
        if (ev_computed[b]) hipEventDestroy(ev_computed[b]);
        if (ev_stored[b]) hipEventDestroy(ev_stored[b]);
        d_group[b] = nullptr;
        h_stage[b] = nullptr;
    }
    if (ev_pass) hipEventDestroy(ev_pass);
    ev_pass = nullptr;
    if (h_state) {
        if (h_state_pageable) std::free(h_state);
        else hipHostFree(h_state);
    }
    h_state = nullptr;
    for (int fd : file_fds) if (fd >= 0) close(fd);
    file_fds.clear();
    if (s_h2d) hipStreamDestroy(s_h2d);
    if (s_d2h) hipStreamDestroy(s_d2h);
    s_h2d = s_d2h = nullptr;
}

// Opens (creating or truncating to size) the state files of the FILES tier.
int sq_engine::stream_open_files(const char *const *paths, int n_files, bool direct, bool keep) {
    host_tier = SQ_HOST_FILES;
    file_direct = direct;
    const size_t per_file = state_bytes() / (size_t)n_files;
    for (int i = 0; i < n_files; i++) {
        int flags = O_RDWR | O_CREAT;
        if (direct) flags |= O_DIRECT;
        const int fd = open(paths[i], flags, 0666);
        if (fd < 0) {
            return fail(SQ_ERR_IO, std::string("[FILE]: ") + paths[i] + " create failed: " + std::strerror(errno));
        }
        file_fds.push_back(fd);
        struct stat st {};
        if (fstat(fd, &st) != 0) return fail(SQ_ERR_IO, std::string("fstat ") + paths[i]);
        if (keep && (size_t)st.st_size != per_file)
            return fail(SQ_ERR_IO, std::string(paths[i]) + ": size does not match the state");
        if (!keep && ftruncate(fd, (off_t)per_file) != 0)
            return fail(SQ_ERR_IO, std::string("ftruncate ") + paths[i] + ": " + std::strerror(errno));
    }
    return SQ_OK;
}

// Host-tier byte access at global byte offset off (ranges may cross files).
// O_DIRECT transfers DMA straight between the file and the user pages, and
// the kernel refuses (EFAULT) pinned memory from hipHostMalloc as their
// target, so with file_direct every transfer goes through `bounce`, an
// ordinary page-aligned buffer, and a memcpy to or from the staging memory.
bool sq_engine::host_read(uint64_t off, size_t bytes, void *dst, char *bounce) {
    if (host_tier != SQ_HOST_FILES) {
        std::memcpy(dst, static_cast<char *>(h_state) + off, bytes);
        return true;
    }
    if (file_direct && !bounce) { io_errno = EINVAL; return false; }
    const uint64_t per_file = state_bytes() / file_fds.size();
    char *d = static_cast<char *>(dst);
    while (bytes) {
        const uint64_t f = off / per_file, fo = off % per_file;
        const size_t take = (size_t)std::min<uint64_t>(bytes, per_file - fo);
        int err = 0;
        if (!pread_all(file_fds[f], file_direct ? bounce : d, take, (off_t)fo, &err)) { io_errno = err; return false; }
        if (file_direct) std::memcpy(d, bounce, take);
        d += take; off += take; bytes -= take;
    }
    return true;
}

bool sq_engine::host_write(uint64_t off, size_t bytes, const void *src, char *bounce) {
    if (host_tier != SQ_HOST_FILES) {
        std::memcpy(static_cast<char *>(h_state) + off, src, bytes);
        return true;
    }
    if (file_direct && !bounce) { io_errno = EINVAL; return false; }
    const uint64_t per_file = state_bytes() / file_fds.size();
    const char *s = static_cast<const char *>(src);
    while (bytes) {
        const uint64_t f = off / per_file, fo = off % per_file;
        const size_t take = (size_t)std::min<uint64_t>(bytes, per_file - fo);
        if (file_direct) std::memcpy(bounce, s, take);
        int err = 0;
        if (!pwrite_all(file_fds[f], file_direct ? bounce : s, take, (off_t)fo, &err)) { io_errno = err; return false; }
        s += take; off += take; bytes -= take;
    }
    return true;
}

// Copies `runs` runs of run_bytes each between the host tier (at host
// offsets run_off[r] * B + base) and a packed staging buffer, on all host
// threads.  Pieces stay 4 KiB-aligned for O_DIRECT.
bool sq_engine::stage_runs(bool to_stage, uint64_t base_bytes, const std::vector<uint64_t> &run_off_bytes,
                           size_t run_bytes, char *stage) {
    const size_t piece = std::max<size_t>(kIoAlign, std::min<size_t>(run_bytes, (size_t)8 << 20));
    const uint64_t per_run = (run_bytes + piece - 1) / piece;
    const uint64_t total = per_run * run_off_bytes.size();
    std::vector<char> ok(host_threads(), 1);
    const unsigned t = host_threads();
    std::vector<std::thread> th;
    for (unsigned k = 0; k < t; k++)
        th.emplace_back([&, k] {
            // O_DIRECT: one page-aligned bounce buffer per thread (host_read)
            char *bounce = nullptr;
            if (host_tier == SQ_HOST_FILES && file_direct) {
                bounce = static_cast<char *>(std::aligned_alloc(kIoAlign, (piece + kIoAlign - 1) / kIoAlign * kIoAlign));
                if (!bounce) { io_errno = ENOMEM; ok[k] = 0; return; }
            }
            for (uint64_t p = total * k / t; p < total * (k + 1) / t; p++) {
                const uint64_t r = p / per_run, i = p % per_run;
                const size_t o = (size_t)i * piece;
                const size_t len = std::min(piece, run_bytes - o);
                char *sp = stage + r * run_bytes + o;
                const uint64_t hp = base_bytes + run_off_bytes[r] + o;
                const bool good = to_stage ? host_read(hp, len, sp, bounce) : host_write(hp, len, sp, bounce);
                if (!good) ok[k] = 0;
            }
            std::free(bounce);
        });
    for (auto &x : th) x.join();
    return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

// Applies fn(global amplitude offset, count, pointer) over the state in
// chunks of at most one staging buffer (FILES tier) or in place.
template <typename F>
int sq_engine::host_chunks(bool write_back, F fn) {
    const uint64_t na = n_amps();
    if (host_tier != SQ_HOST_FILES) {
        fn(0, na, static_cast<char *>(h_state));
        return SQ_OK;
    }
    const uint64_t chunk = std::min<uint64_t>(na, 1ULL << res_qbit);
    char *buf = static_cast<char *>(h_stage[0]);
    const size_t B = amp_bytes();
    for (uint64_t a = 0; a < na; a += chunk) {
        std::vector<uint64_t> one{0};
        if (!stage_runs(true, a * B, one, chunk * B, buf)) return io_fail("state file read");
        fn(a, chunk, buf);
        if (write_back && !stage_runs(false, a * B, one, chunk * B, buf)) return io_fail("state file write");
    }
    return SQ_OK;
}

int sq_engine::stream_init_zero() {
    const size_t B = amp_bytes();
    if (host_tier != SQ_HOST_FILES) {
        char *p = static_cast<char *>(h_state);
        parallel_for(state_bytes(), [&](uint64_t a, uint64_t b) { std::memset(p + a, 0, b - a); });
    } else {
        // zero every file through the staging buffer (setupStateFile, simulator.cpp:638-658)
        const uint64_t chunk = std::min<uint64_t>(n_amps(), 1ULL << res_qbit);
        char *buf = static_cast<char *>(h_stage[0]);
        parallel_for(chunk * B, [&](uint64_t a, uint64_t b) { std::memset(buf + a, 0, b - a); });
        std::vector<uint64_t> one{0};
        for (uint64_t a = 0; a < n_amps(); a += chunk)
            if (!stage_runs(false, a * B, one, chunk * B, buf)) return io_fail("state file write");
    }
    // amplitude 0 = 1 (MEM_Runner.cpp:79-81)
    alignas(16) unsigned char one[16] = {};
    if (dtype == SQ_C128) { const double v = 1.0; std::memcpy(one, &v, 8); }
    else { const float v = 1.0f; std::memcpy(one, &v, 4); }
    if (host_tier != SQ_HOST_FILES) {
        std::memcpy(h_state, one, B);
    } else {
        char *buf = static_cast<char *>(h_stage[0]);
        std::memset(buf, 0, kIoAlign);
        std::memcpy(buf, one, B);
        alignas(kIoAlign) char bounce[kIoAlign];
        if (!host_write(0, std::min<size_t>(kIoAlign, state_bytes()), buf, bounce)) return io_fail("state file write");
    }
    return SQ_OK;
}

int sq_engine::stream_set_state(const void *host, int32_t host_dtype) {
    const size_t hb = host_dtype == SQ_C128 ? 16 : 8;
    return host_chunks(true, [&](uint64_t a0, uint64_t cnt, char *dst) {
        const char *src = static_cast<const char *>(host) + a0 * hb;
        if (host_dtype == dtype) {
            parallel_for(cnt * hb, [&](uint64_t a, uint64_t b) { std::memcpy(dst + a, src + a, b - a); });
        } else if (host_dtype == SQ_C128) {
            const double *s = reinterpret_cast<const double *>(src);
            float *d = reinterpret_cast<float *>(dst);
            parallel_for(2 * cnt, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = (float)s[i]; });
        } else {
            const float *s = reinterpret_cast<const float *>(src);
            double *d = reinterpret_cast<double *>(dst);
            parallel_for(2 * cnt, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = s[i]; });
        }
    });
}

int sq_engine::stream_get_state(void *host, int32_t host_dtype) {
    const size_t hb = host_dtype == SQ_C128 ? 16 : 8;
    return host_chunks(false, [&](uint64_t a0, uint64_t cnt, char *src) {
        char *dst = static_cast<char *>(host) + a0 * hb;
        if (host_dtype == dtype) {
            parallel_for(cnt * hb, [&](uint64_t a, uint64_t b) { std::memcpy(dst + a, src + a, b - a); });
        } else if (host_dtype == SQ_C128) {
            const float *s = reinterpret_cast<const float *>(src);
            double *d = reinterpret_cast<double *>(dst);
            parallel_for(2 * cnt, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = s[i]; });
        } else {
            const double *s = reinterpret_cast<const double *>(src);
            float *d = reinterpret_cast<float *>(dst);
            parallel_for(2 * cnt, [&](uint64_t a, uint64_t b) { for (uint64_t i = a; i < b; i++) d[i] = (float)s[i]; });
        }
    });
}

// FP64 sums of |a|^2 and |a|^2 zz(i) over the host state.
int sq_engine::stream_norm_zz(double *v) {
    const uint64_t mask = (n >= 64) ? ~0ULL : ((1ULL << n) - 1);
    const unsigned t = host_threads();
    double acc0 = 0, acc1 = 0;
    int rc = host_chunks(false, [&](uint64_t a0, uint64_t cnt, char *p) {
        std::vector<double> s0(t, 0.0), s1(t, 0.0);
        std::vector<std::thread> th;
        for (unsigned k = 0; k < t; k++)
            th.emplace_back([&, k] {
                const uint64_t a = cnt * k / t, b = cnt * (k + 1) / t;
                // two-level sums (4096-term blocks): the rounding error stays
                // near one block's instead of growing with 2^n / threads terms
                double x0 = 0, x1 = 0;
                for (uint64_t j0 = a; j0 < b; j0 += 4096) {
                    double y0 = 0, y1 = 0;
                    const uint64_t j1 = std::min<uint64_t>(b, j0 + 4096);
                    for (uint64_t j = j0; j < j1; j++) {
                        double re, im;
                        if (dtype == SQ_C128) { re = reinterpret_cast<double *>(p)[2 * j]; im = reinterpret_cast<double *>(p)[2 * j + 1]; }
                        else { re = reinterpret_cast<float *>(p)[2 * j]; im = reinterpret_cast<float *>(p)[2 * j + 1]; }
                        const double pr = re * re + im * im;
                        const uint64_t i = a0 + j;
                        const uint64_t rot = ((i >> 1) | (i << (n - 1))) & mask;
                        y0 += pr;
                        y1 += pr * (double)(n - 2 * __builtin_popcountll((i ^ rot) & mask));
                    }
                    x0 += y0;
                    x1 += y1;
                }
                s0[k] = x0; s1[k] = x1;
            });
        for (auto &x : th) x.join();
        for (unsigned k = 0; k < t; k++) { acc0 += s0[k]; acc1 += s1[k]; }
    });
    if (rc) return rc;
    v[0] = acc0;
    v[1] = acc1;
    return SQ_OK;
}

int sq_engine::stream_apply(const std::vector<sq_gate> &gates, const std::vector<LoweredGate> &low) {
    const int W = res_qbit;
    const int h = stream_high();   // free slots above the fixed low run
    std::vector<int> fixed;
    for (int q = 0; q < W - h; q++) fixed.push_back(q);
    const auto passes = sq::plan_passes(gates, n, W, fixed);

    const size_t B = amp_bytes();
    const uint64_t run = 1ULL << (W - h);
    const uint64_t n_groups = 1ULL << (n - W);
    const bool staged = host_tier != SQ_HOST_PINNED;
    void *const saved_state = d_state;
    const int saved_n = n;
    int rc = SQ_OK;
    for (const auto &pass : passes) {
        // groups of different passes overlap in host memory: every upload of
        // this pass waits for every download of the previous one (staged
        // tiers scatter every group before the pass ends)
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
        std::vector<sq_gate> pgates;
        for (int gi : pass.gates) {
            pg.push_back(remap(low[gi], P));
            sq_gate g = gates[gi];
            for (int j = 0; j < g.nq; j++) g.q[j] = (int)(std::lower_bound(P.begin(), P.end(), g.q[j]) - P.begin());
            pgates.push_back(g);
        }
        // the pass's gates as register passes over a W-qubit group: planned and
        // uploaded once, run on every group
        PlainProgram prog;
        n = W;
        rc = build_plain(pgates, pg, prog);
        n = saved_n;
        if (rc) break;
        std::vector<uint64_t> run_off(1ULL << h), run_off_bytes(1ULL << h);   // host offset of run r within a group
        for (uint64_t r = 0; r < run_off.size(); r++) {
            uint64_t o = 0;
            for (int j = 0; j < h; j++) if ((r >> j) & 1) o += 1ULL << H[j];
            run_off[r] = o;
            run_off_bytes[r] = o * B;
        }
        char *host = static_cast<char *>(h_state);
        int64_t pending_scatter[2] = {-1, -1};
        uint64_t pending_base[2] = {0, 0};
        auto scatter = [&](int b) -> bool {
            if (pending_scatter[b] < 0) return true;
            if (hipEventSynchronize(ev_stored[b]) != hipSuccess) return false;
            const bool ok = stage_runs(false, pending_base[b] * B, run_off_bytes, run * B, static_cast<char *>(h_stage[b]));
            pending_scatter[b] = -1;
            return ok;
        };
        for (uint64_t g = 0; g < n_groups; g++) {
            const int b = (int)(g & 1);
            uint64_t base = g;
            for (int q : P) base = ((base >> q) << (q + 1)) | (base & ((1ULL << q) - 1));
            char *dbuf = static_cast<char *>(d_group[b]);
            if (staged) {
                // the staging buffer's previous group goes back first, then this group is packed into it
                if (!scatter(b)) { rc = io_fail("out-of-core scatter"); break; }
                if (!stage_runs(true, base * B, run_off_bytes, run * B, static_cast<char *>(h_stage[b]))) {
                    rc = io_fail("out-of-core gather");
                    break;
                }
            }
            // the buffer's previous download must finish before it is refilled
            hipStreamWaitEvent(s_h2d, ev_stored[b], 0);
            if (staged) {
                hipMemcpyAsync(dbuf, h_stage[b], (run << h) * B, hipMemcpyHostToDevice, s_h2d);
            } else {
                for (uint64_t r = 0; r < run_off.size(); r++)
                    hipMemcpyAsync(dbuf + r * run * B, host + (base + run_off[r]) * B, run * B, hipMemcpyHostToDevice, s_h2d);
            }
            hipEventRecord(ev_loaded[b], s_h2d);
            hipStreamWaitEvent(stream, ev_loaded[b], 0);
            d_state = d_group[b];
            n = W;
            rc = run_plain_program(prog, pg);
            n = saved_n;
            d_state = saved_state;
            if (rc) break;
            hipEventRecord(ev_computed[b], stream);
            hipStreamWaitEvent(s_d2h, ev_computed[b], 0);
            if (staged) {
                hipMemcpyAsync(h_stage[b], dbuf, (run << h) * B, hipMemcpyDeviceToHost, s_d2h);
                pending_scatter[b] = (int64_t)g;
                pending_base[b] = base;
            } else {
                for (uint64_t r = 0; r < run_off.size(); r++)
                    hipMemcpyAsync(host + (base + run_off[r]) * B, dbuf + r * run * B, run * B, hipMemcpyDeviceToHost, s_d2h);
            }
            hipEventRecord(ev_stored[b], s_d2h);
        }
        if (!rc && staged && (!scatter(0) || !scatter(1))) rc = io_fail("out-of-core scatter");
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
