// engine_internal.hpp -- the engine object shared by the .hip translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "host/gate_lowering.hpp"
#include "host/planner.hpp"

struct PendingStat {
    char name[48];
    double alg_bytes;
    double alg_flops;
    hipEvent_t start, stop;
};

// A plain gate list as register passes, planned and its op program uploaded
// once; it runs on any state buffer of the size (n qubits) it was planned for.
struct PlainLaunch {
    bool merged = false;
    int gate = -1;                 // single-gate step: index into the lowered list
    int op_begin = 0, n_ops = 0;   // merged step: slice of the uploaded ops
    double flops_per_amp = 0;      // FP64 (complex64: FP32) operations executed per amplitude
    int hpos[5] = {};              // register qubits of the pass (virtual positions)
    std::vector<int> v2l;          // virtual position -> qubit (its index bit)
};
struct FusedPassPlan;   // fused passes of a gate list (sq_fused.hip)

struct PlainProgram {
    int n = 0, kh = 0;
    bool per_gate = false;         // one launch per gate (sweep_mode 1, small states)
    std::vector<PlainLaunch> launches;
    const void *d_ops = nullptr;   // device copy of the op records the launches slice
};

struct sq_engine {
    int n = 0;                  // total_qbit
    int chunk = 12;             // chunk_qbit (tile width of the subcircuit sweep)
    int dtype = SQ_C128;
    int device = 0;
    void *d_state = nullptr;    // 2^n amplitudes, interleaved (re, im)
    double *d_part = nullptr;   // reduction partials
    void *d_stage = nullptr;    // dtype-conversion staging (16 MiB amplitudes max)
    hipStream_t stream = nullptr;
    unsigned max_blocks = 2048;
    size_t lds_per_block = 65536;   // hipDeviceAttributeMaxSharedMemoryPerBlock
    std::string err;

    bool stats_on = false;
    std::vector<PendingStat> pending;

    // extension buffers owned by other translation units
    void *d_blockprog = nullptr;    // subcircuit program (device copy)
    size_t blockprog_cap = 0;
    void *d_fused = nullptr;        // fused-block matrices
    size_t fused_cap = 0;
    void *d_aux = nullptr;          // adjoint: second state (lambda)
    size_t aux_cap = 0;
    void *d_ckpt = nullptr;         // adjoint: FP64 checkpoints
    size_t ckpt_cap = 0;
    void *d_gradpart = nullptr;     // adjoint: per-parameter block partials
    size_t gradpart_cap = 0;
    uint64_t n_tile_passes = 0;     // tile passes launched (diagnostics)
    void *d_sample = nullptr;       // sampler: block sums, shots, results
    size_t sample_cap = 0;
    void *d_passprog = nullptr;     // register-pass program of the plain sweep
    size_t passprog_cap = 0;
    void *d_scratch = nullptr;      // this device's static scratch array (sq_apply_gates_noalloc)
    int sweep_mode = 0;             // 0: register passes, 1: one launch per gate
    int pass_qbits = 5;             // register qubits per pass (1..5; set by sq_create)
    int lane_moves = 3;             // top lane positions a pass may give to other qubits (0..3)
    int plan_beam = sq::kDefaultPlanBeam;   // register-pass planner: sets kept per size (1: greedy)
    bool phase_split = true;        // register passes: one-qubit dense gates as phase, real rotation, phase
    void *d_state2 = nullptr;       // run_plain_program: a second state the same program runs on (set per call)
    // Register-pass schedules of recent gate lists, keyed by their structure
    // (qubits and gate classes: the schedule does not depend on angles or
    // coefficients), most recent first.
    struct PlanCacheEntry {
        std::vector<int32_t> key;
        std::vector<sq::SweepStep> steps;
    };
    std::vector<PlanCacheEntry> plan_cache;
    // fused-pass plans of recent gate lists, keyed the same way, most recent first
    std::vector<std::pair<std::vector<int32_t>, std::shared_ptr<FusedPassPlan>>> fpass_cache;
    std::array<int64_t, 20> pass_census{};   // op kinds of the last register-pass program (sq_pass_census)
    bool bitperm_ok = true;         // the blocked sweep's final qubit order by one permutation pass
    void *d_alt = nullptr;          // second state buffer (out-of-place permutation pass)

    // out-of-core runner (resident_qbit > 0): state outside HBM
    int res_qbit = 0;
    int host_tier = SQ_HOST_AUTO;   // where the state lives (sq_host_tier)
    void *h_state = nullptr;        // PINNED / PAGEABLE tiers
    bool h_state_pageable = false;
    void *h_stage[2] = {nullptr, nullptr};   // pinned staging (PAGEABLE / FILES tiers)
    std::vector<int> file_fds;      // FILES tier: file i = amplitudes [i 2^n/nf, (i+1) 2^n/nf)
    bool file_direct = false;
    std::atomic<int> io_errno{0};   // errno of the last failed state-file transfer (any host thread)
    int io_fail(const char *what);
    void *d_group[2] = {nullptr, nullptr};
    hipEvent_t ev_loaded[2] = {nullptr, nullptr}, ev_computed[2] = {nullptr, nullptr},
               ev_stored[2] = {nullptr, nullptr}, ev_pass = nullptr;
    hipStream_t s_h2d = nullptr, s_d2h = nullptr;
    uint64_t n_stream_passes = 0;
    bool out_of_core() const { return res_qbit > 0; }
    int stream_create();
    int stream_open_files(const char *const *paths, int n_files, bool direct, bool keep);
    int stream_high() const;
    // bounce: an ordinary page-aligned buffer of at least `bytes`, through which
    // O_DIRECT file transfers go (required when file_direct)
    bool host_read(uint64_t off, size_t bytes, void *dst, char *bounce);
    bool host_write(uint64_t off, size_t bytes, const void *src, char *bounce);
    bool stage_runs(bool to_stage, uint64_t base_bytes, const std::vector<uint64_t> &run_off_bytes, size_t run_bytes,
                    char *stage);
    template <typename F> int host_chunks(bool write_back, F fn);
    void stream_destroy();
    int stream_init_zero();
    int stream_set_state(const void *host, int32_t host_dtype);
    int stream_get_state(void *host, int32_t host_dtype);
    int stream_norm_zz(double *v);
    int stream_apply(const std::vector<sq_gate> &gates, const std::vector<sq::LoweredGate> &low);

    uint64_t n_amps() const { return 1ULL << n; }
    size_t amp_bytes() const { return dtype == SQ_C128 ? 16 : 8; }
    size_t state_bytes() const { return (size_t)n_amps() * amp_bytes(); }

    int fail(int code, const std::string &msg);
    int hip_fail(hipError_t e, const char *what);
    int check_launch(const char *what);
    unsigned grid_for(uint64_t work_items, unsigned per_block) const;
    unsigned grid_flat(uint64_t work_items, unsigned per_block) const;
    int stat_begin(const char *name, double alg_bytes, double alg_flops = 0.0);
    void stat_end(int slot);

    int lower(const sq_gate &g, sq::LoweredGate &lg);
    int apply_lowered(const sq::LoweredGate &g);
    int apply_plain(const std::vector<sq_gate> &gates, const std::vector<sq::LoweredGate> &low);
    // CNOTs (exact U2 permutation matrices) as H CZ H when register passes run;
    // false (g2, l2 untouched) when nothing changes
    bool expand_cnots(const std::vector<sq_gate> &gates, const std::vector<sq::LoweredGate> &low,
                      std::vector<sq_gate> &g2, std::vector<sq::LoweredGate> &l2) const;
    // dst: where the op records go (nullptr: the engine's own growable buffer;
    // otherwise a caller-owned device buffer of dst_cap bytes, never grown)
    int build_plain(const std::vector<sq_gate> &gates, const std::vector<sq::LoweredGate> &low, PlainProgram &prog,
                    void *dst = nullptr, size_t dst_cap = 0);
    int ensure_alt();
    // Moves index bit q of the state to bit perm[q] (a permutation of [0, n)
    // keeping the vector bits), out of place through d_alt in one pass.
    int bit_permute(const std::vector<int> &perm);
    int setup_scratch();
    int run_plain_program(const PlainProgram &prog, const std::vector<sq::LoweredGate> &low);
    int ensure_partials();
    int norm_zz(double *v);
    int compare(sq_engine *o, double *v);
    void free_extras();
};
