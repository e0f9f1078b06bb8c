/*
 * subquokka_hip.h -- C ABI of the MI355X state-vector engine.
 *
 * This is the boundary a SubQuokka runner binds.  The reference selects a
 * runner by the INI key runner_type (simulator.cpp:202-228) and calls its
 * circuitRunner interface (simulator.hpp:85-97):
 *     virtual void run(std::vector<Gate*>&);               plain gate-by-gate
 *     virtual void run(std::vector<std::vector<Gate*>>&);  subcircuit blocks
 * MEM_Runner owns a 2^N complex<double> state in RAM (MEM_Runner.cpp:57-88).
 * Here the engine owns the state in HBM (or, for the out-of-core runner, in
 * pinned host memory) and each entry point below replaces one piece of that
 * path; the replaced reference code is cited on each declaration.
 *
 * Conventions:
 *  - Qubit q is bit q of the amplitude index (simulator.cpp:168-175).
 *  - Host state buffers are interleaved (re, im): complex<double> for
 *    SQ_C128, complex<float> for SQ_C64.
 *  - Every call returns SQ_OK (0) or a negative sq_status; no call exits the
 *    process.  sq_last_error() returns the message of the last failure on that
 *    engine.  Calls on one engine are serialised by the caller; separate
 *    engines are independent (no global state).
 *  - No torch types: plain pointers and sizes only.
 */
#ifndef SUBQUOKKA_HIP_H
#define SUBQUOKKA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum sq_status {
    SQ_OK = 0,
    SQ_ERR_ARG = -1,          /* bad argument / size mismatch */
    SQ_ERR_UNKNOWN_GATE = -2, /* simulator.cpp:593-596 "[setGate]: Not implemented yet." */
    SQ_ERR_QUBIT_INDEX = -3,  /* simulator.cpp:764-774 "[setGate]: Qubit Index Error" */
    SQ_ERR_CONTRACT = -4,     /* input outside the reference's valid layout
                                 (U2 targ[0] > targ[1], circuit.cpp:2809; repeated targets) */
    SQ_ERR_PARSE = -5,        /* malformed circuit/INI text */
    SQ_ERR_OOM = -6,          /* device or pinned-host allocation failed */
    SQ_ERR_HIP = -7,          /* a HIP runtime call failed */
    SQ_ERR_UNSUPPORTED = -8,  /* runner_type RDMA, VSWAP_3_3/4_4/6_6, U3, MPI_* */
    SQ_ERR_IO = -9
} sq_status;

typedef enum sq_dtype {
    SQ_C128 = 0,   /* complex<double>, the reference's only state type */
    SQ_C64 = 1     /* complex<float> */
} sq_dtype;

/* Gate kinds: the MEM gate names of setGate_MEM (simulator.cpp:542-597). */
typedef enum sq_gate_kind {
    SQ_GATE_H = 1, SQ_GATE_X, SQ_GATE_Y, SQ_GATE_Z, SQ_GATE_P,
    SQ_GATE_RX, SQ_GATE_RY, SQ_GATE_RZ, SQ_GATE_U1, SQ_GATE_U2,
    SQ_GATE_SWAP, SQ_GATE_CP, SQ_GATE_RZZ, SQ_GATE_VSWAP_1_1, SQ_GATE_VSWAP_2_2
} sq_gate_kind;

/* One circuit line, parsed.  Replaces the Gate* objects built by setGate_MEM
 * (simulator.cpp:259-356, 542-597; circuit.hpp:19-47).
 *   q[]   targets as written (symmetric 2q gates may be in either order)
 *   param angle of P, RX, RY, RZ, CP, RZZ
 *   mat   U1: 8 reals, row-major (re,im) of [[c0,c1],[c2,c3]] (circuit.cpp:274-277)
 *         U2: 32 reals, 4x4 row-major in basis |q0 q1> with q0 as the most
 *             significant bit (the layout circuit.cpp:2811-2819 reverses). */
typedef struct sq_gate {
    int32_t kind;
    int32_t nq;
    int32_t q[4];
    double param;
    double mat[32];
} sq_gate;

/* A dense k-qubit block (1 <= k <= 6) with a 2^k x 2^k complex matrix in FP64,
 * row-major interleaved (re,im); local index bit j <-> qubit q[j].
 * No reference counterpart: it is the MI355X form of the subcircuit blocks
 * that MEM_Runner::run(subcircuits) applies tile by tile (MEM_Runner.cpp:173-182). */
typedef struct sq_fused {
    int32_t k;
    int32_t q[6];
    const double *u;
} sq_fused;

/* Where the out-of-core runner keeps the state (resident_qbit > 0). */
typedef enum sq_host_tier {
    SQ_HOST_AUTO = 0,       /* pinned host memory when the whole state can be pinned,
                               else ordinary memory + pinned staging */
    SQ_HOST_PAGEABLE = 1,   /* ordinary host memory + pinned staging buffers */
    SQ_HOST_PINNED = 2,     /* (reported by sq_stream_info) */
    SQ_HOST_FILES = 3       /* state files: sq_create_file_state */
} sq_host_tier;

typedef struct sq_config {
    int32_t total_qbit;       /* [system] total_qbit (simulator.cpp:90) */
    int32_t chunk_qbit;       /* [system] chunk_qbit (simulator.cpp:94); blocked-mode tile */
    int32_t dtype;            /* sq_dtype */
    int32_t device;           /* HIP device ordinal */
    int32_t resident_qbit;    /* 0: whole state in HBM (MEM runner).  >0: out-of-core
                                 runner, state in pinned host memory, 2^resident_qbit
                                 amplitudes per streamed chunk (DIO_Runner role) */
    int32_t sweep_mode;       /* 0 = register passes (a run of one-qubit and diagonal
                                 gates per read/write of the state); subcircuits with
                                 VSWAPs relabelled and one permutation pass at the end.
                                 1 = one launch per gate; subcircuits as VSWAP sweeps
                                 and one LDS tile pass per gate block */
    int32_t pass_qbits;       /* register qubits per pass, 1..5 (0: 5) */
    int32_t host_tier;        /* out-of-core runner: SQ_HOST_AUTO or SQ_HOST_PAGEABLE */
} sq_config;

typedef struct sq_engine sq_engine;

/* ---- lifetime: replaces MEM_Runner::MEM_Runner (MEM_Runner.cpp:57-88) ---- */
int sq_create(const sq_config *cfg, sq_engine **out);
/* Out-of-core runner whose state lives in n_files files (a power of two, at
 * most 2^total_qbit): file i holds amplitudes [i 2^N/n_files, (i+1) 2^N/n_files)
 * as raw little-endian (re, im) of the engine's dtype -- the reference's
 * state files (setupFiles simulator.cpp:102-129, loader_IO test_util.py:385-403).
 * direct_io opens them O_DIRECT (runner_type DirectIO, create_file
 * simulator.cpp:71-86; runs and files must then be multiples of 4 KiB).
 * keep_contents = 0 creates or truncates the files and sets |0...0>;
 * 1 opens existing files of the right size as the current state (resume).
 * cfg->resident_qbit must be > 0. */
int sq_create_file_state(const sq_config *cfg, const char *const *paths, int32_t n_files, int32_t direct_io,
                         int32_t keep_contents, sq_engine **out);
int sq_destroy(sq_engine *e);
/* |0...0>: buffer[0] = 1 (MEM_Runner.cpp:79-81; setupStateFile simulator.cpp:639-658) */
int sq_init_zero(sq_engine *e);
/* Copies a host state in/out, converting between dtypes if they differ.
 * n_amps must be 2^total_qbit.  (No reference counterpart beyond the public
 * MEM_Runner::buffer, MEM_Runner.hpp:27, and the text dump simulator.cpp:696-704.) */
int sq_set_state(sq_engine *e, const void *host, uint64_t n_amps, int32_t dtype);
int sq_get_state(sq_engine *e, void *host, uint64_t n_amps, int32_t dtype);
/* Device pointer and stream of the resident state (SQ_ERR_ARG for the
 * out-of-core runner).  Lets a caller time or read the state in place. */
int sq_device_state(sq_engine *e, void **dptr, void **hip_stream);
int sq_synchronize(sq_engine *e);
const char *sq_last_error(const sq_engine *e);

/* ---- the hot path ---- */
/* Plain gate-by-gate sweep: MEM_Runner::run(vector<Gate*>&) (MEM_Runner.cpp:101-134). */
int sq_apply_gates(sq_engine *e, const sq_gate *gates, uint64_t n);
/* sq_apply_gates with no allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Same kernels and result
 * as sq_apply_gates on an HBM-resident engine; the register-pass program goes
 * into a fixed 4 MiB device array (one per GPU, set up by sq_create, held until
 * the process ends) instead of a buffer the engine grows: a list whose program
 * exceeds it (about 32768 merged gates in complex128, 65536 in complex64) fails
 * with SQ_ERR_ARG before anything runs.  The call still synchronises the
 * engine's stream before it writes the array, as sq_apply_gates does before it
 * uploads.  Every engine on a GPU shares that GPU's array, which keeps the last
 * call's program: calls that overlap on one GPU (from two host threads, or two
 * engines whose streams run at once) race on it.  SQ_ERR_UNSUPPORTED for the
 * out-of-core runner. */
int sq_apply_gates_noalloc(sq_engine *e, const sq_gate *gates, uint64_t n);
/* Subcircuit sweep: MEM_Runner::run(vector<vector<Gate*>>&) (MEM_Runner.cpp:136-185).
 * block_sizes[i] gates form block i; blocks whose first gate is VSWAP_* are
 * qubit-swap passes, the others gate blocks.  The final state, in the layout the
 * VSWAPs leave, is the reference's; how it gets there follows sweep_mode. */
int sq_apply_blocks(sq_engine *e, const sq_gate *gates, const uint32_t *block_sizes,
                    uint64_t n_blocks);
/* Dense fused blocks on MFMA (no reference counterpart, see sq_fused).
 * Matrices are supplied by the caller in FP64. */
int sq_apply_fused(sq_engine *e, const sq_fused *blocks, uint64_t n);
/* Fused blocks from gates: block i is block_sizes[i] consecutive gates acting on
 * at most five qubits (the planner's mode 1 output).  Each block is composed in
 * FP64 on the GPU, then applied on MFMA (FP32 accumulation for SQ_C64, FP64 for
 * SQ_C128).  The result equals applying the gates in order. */
int sq_run_fused(sq_engine *e, const sq_gate *gates, const uint32_t *block_sizes, uint64_t n_blocks);
/* A plain gate list (the sq_apply_gates / MEM_Runner::run(vector<Gate*>&) input,
 * MEM_Runner.cpp:101-134) as fused passes on a complex64 state: the engine cuts
 * the list into passes over tiles of tile_qbit qubits (0: 13; else 9..13, at
 * most N), groups each pass's gates into four-qubit blocks, composes each block
 * in FP64 on the GPU and applies it to the tile in LDS on MFMA with split bf16
 * products accumulated in FP32 (about 1e-5 relative per block; the reference's
 * cache-resident subcircuits, MEM_Runner.cpp:173-182, in one read and write of
 * the state per pass).  Diagonal gates may act on any qubits.  No reference
 * counterpart for the arithmetic.  SQ_ERR_UNSUPPORTED for SQ_C128 engines and
 * the out-of-core runner; SQ_ERR_ARG below 9 qubits. */
int sq_apply_gates_fused(sq_engine *e, const sq_gate *gates, uint64_t n, int32_t tile_qbit);
/* Composes gates (all targets inside q[0..k-1]) into one 2^k x 2^k FP64 matrix on
 * the GPU; u_out receives 2*4^k doubles. */
int sq_compose_block(sq_engine *e, const sq_gate *gates, uint64_t n, int32_t k,
                     const int32_t *q, double *u_out);

/* ---- reductions (FP64 accumulation on the GPU) ---- */
int sq_norm2(sq_engine *e, double *out);
/* <psi| sum_i Z_i Z_{(i+1) mod N} |psi> */
int sq_expect_zz_ring(sq_engine *e, double *out);
/* <a|b> over two engines of the same size (any dtypes). */
int sq_inner(sq_engine *a, sq_engine *b, double *re, double *im);
/* max_i |a_i - b_i| over two engines of the same size. */
int sq_max_abs_diff(sq_engine *a, sq_engine *b, double *out);

/* ---- measurement (no reference counterpart: only the unused MEASURE gate
 * type, circuit.hpp:11) ----
 * shots samples of the computational-basis distribution |a_i|^2 / sum |a|^2
 * of the HBM-resident state, driven by SplitMix64(seed); out[k] = index of
 * shot k.  The rule is exact and reproducible (FP64 sums in a fixed order;
 * see kernels/sample.hpp), so equal seeds give equal shots. */
int sq_sample(sq_engine *e, uint64_t seed, uint64_t shots, uint64_t *out);

/* ---- adjoint gradient (no reference counterpart) ----
 * Hardware-efficient ansatz: layers x (RY(theta) on every qubit, RZ(theta) on
 * every qubit, CNOT ring i -> (i+1) mod N); theta has 2*N*layers entries in
 * gate order.  Energy E = <psi| sum_i Z_i Z_{i+1} |psi> from |0...0>.
 * n_theta must equal 2*N*layers (SQ_ERR_ARG otherwise); grad receives
 * n_theta values.  checkpoint_every: layers between re-anchoring checkpoints
 * (0 = none). */
int sq_hea_energy(sq_engine *e, int32_t layers, const double *theta, uint64_t n_theta, double *energy);
int sq_adjoint_grad_hea(sq_engine *e, int32_t layers, const double *theta, uint64_t n_theta,
                        int32_t checkpoint_every, double *energy, double *grad);

/* ---- adjoint gradient of parameterised circuits (no reference counterpart;
 * the rotation and phase matrices are the reference's, circuit.cpp:2655-2795,
 * 2947-3041) ----
 * A diagonal observable is a sum of Z strings: h(i) = sum_t coef_t *
 * (-1)^popcount(i & mask_t) (mask 0: a constant term). */
typedef struct sq_zterm {
    double coef;
    uint64_t mask;
} sq_zterm;
/* <psi|H|psi> of the current state, accumulated in FP64. */
int sq_expect_z(sq_engine *e, const sq_zterm *terms, uint64_t n_terms, double *out);
/* Energy E = <psi|H|psi> of the circuit applied to the engine's current state,
 * and dE/dtheta.  Gate i uses the angle theta[param[i]] when param[i] >= 0
 * (RX, RY, RZ, P, CP or RZZ; its own param field is ignored; several gates may
 * share a parameter), else it is a fixed gate of any kind but VSWAP.  The
 * forward runs as the plain sweep's register passes and keeps an FP64 copy of
 * the state every checkpoint_every parameterised gates (0: none); the reverse
 * sweep un-applies every gate to psi and lambda = H psi (fixed gates as
 * register passes, each parameterised gate in one pass that also takes its
 * gradient term, accumulated in FP64) and re-anchors psi to the copies.  grad
 * receives n_theta values.  On return the state is the one the call started
 * from (exactly, from the first copy, when checkpoint_every > 0; else up to
 * the rounding of the reverse sweep). */
int sq_adjoint_grad(sq_engine *e, const sq_gate *gates, const int32_t *param, uint64_t n_gates,
                    const double *theta, uint64_t n_theta, const sq_zterm *terms, uint64_t n_terms,
                    int32_t checkpoint_every, double *energy, double *grad);

/* ---- per-kernel timing (HIP events on the engine's stream) ---- */
typedef struct sq_kstat {
    char name[48];
    uint64_t launches;
    double total_ms;
    double alg_bytes;   /* algorithmic bytes: 2*B*2^N*f_g summed over launches */
    double alg_flops;   /* algorithmic real flops (0 where not counted) */
} sq_kstat;
int sq_stats_enable(sq_engine *e, int32_t on);
int sq_stats_get(sq_engine *e, sq_kstat *out, int32_t cap, int32_t *n_out);
int sq_stats_reset(sq_engine *e);
/* Op census of the engine's last register-pass program (diagnostics; no
 * reference counterpart): how many ops of each form its passes run.
 *   0 dense on a register slot, 1 dense on lane bits 0-3, 2 dense on lane bits
 *   4-5, 3 dense inside the complex64 vector, 4 diagonal without register
 *   bits, 5 diagonal with one register bit, 6 diagonal with two, 7 products
 *   of one-qubit phases (dense gates run as phase, real rotation, phase);
 *   8 + 4 t + c: a controlled phase folded into a dense op (complex128), target
 *   t = 0 register slot, 1 lane bits 0-3, 2 lane bits 4-5; control c = 0
 *   register slot 0, 1 another register slot, 2 a lane bit, 3 a tile bit.
 * *n_out = 20; SQ_ERR_ARG when cap < 20 (the first cap counts are written). */
int sq_pass_census(sq_engine *e, int64_t *counts, int32_t cap, int32_t *n_out);
/* Out-of-core runner telemetry (no reference counterpart; DIO_Runner.cpp has
 * only the "<t>s" line): streamed passes since the last sq_stats_reset, the
 * bytes those passes moved over the host link (both directions) and the host
 * tier in use (sq_host_tier; SQ_HOST_AUTO for an HBM-resident engine). */
int sq_stream_info(sq_engine *e, uint64_t *passes, double *host_link_bytes, int32_t *host_tier);

/* ---- host-side formats (no device needed) ---- */
/* Parses one gate line of the MEM grammar (setGate_MEM, simulator.cpp:542-597;
 * check_targ_in_N simulator.cpp:764-774). */
int sq_parse_gate(const char *line, int32_t total_qbit, sq_gate *out);
/* Parses a whole circuit file text.  Plain: one gate per line.  Subcircuit
 * (simulator.cpp:619-636): "<count>" then count gate lines, repeated; block
 * sizes go to block_sizes (may be NULL for plain).  On overflow returns
 * SQ_ERR_ARG with the needed counts in *n_gates / *n_blocks. */
int sq_parse_circuit(const char *text, int32_t total_qbit, int32_t is_subcircuit,
                     sq_gate *gates, uint64_t gate_cap, uint64_t *n_gates,
                     uint32_t *block_sizes, uint64_t block_cap, uint64_t *n_blocks,
                     int32_t *err_line);

/* Seeded synthetic inputs (SplitMix64; identical on every host).
 * kind: 0 = QFT in the reference's order (python_script/cir_qft.py:27-35),
 *       1 = random layers (U3 on every qubit, emitted as U1, then CZ on a
 *           random perfect matching, emitted as CP pi), 2 = HEA (RY, RZ, CNOT
 *           ring as U2).  Writes circuit text into buf (needs *len bytes). */
int sq_gen_circuit(int32_t kind, int32_t n_qubits, int32_t layers, uint64_t seed,
                   const double *theta, char *buf, uint64_t cap, uint64_t *len);
/* Gaussian (re, im) i.i.d., normalised in FP64: 2^n complex<double>. */
int sq_gen_state(int32_t n_qubits, uint64_t seed, double *out);

/* Planner: groups a plain circuit into blocks (finder role, finder.cpp:629-877).
 * Output is subcircuit text ("<count>" then count gate lines, repeated) in buf.
 *   mode 0: blocks with relabelled targets and VSWAP passes for a chunk of
 *           param low qubits (finder.cpp:69-333 format);
 *   mode 3: mode 0 plus swaps that restore the qubit order at the end;
 *   mode 1: fused blocks of at most param qubits (fused_k);
 *   mode 2: the plain sweep's register-pass schedule for complex128 (qubits
 *           0-2 resident plus at most param + 3 chosen qubits per step: param
 *           register qubits and three lane positions). */
int sq_plan(const char *circuit_text, int32_t total_qbit, int32_t mode, int32_t param,
            char *buf, uint64_t cap, uint64_t *len);

#ifdef __cplusplus
}
#endif
#endif /* SUBQUOKKA_HIP_H */
