/*
 * sq_oracle.c -- FP64 CPU oracle for SubQuokka's MEM gate-application path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of bench.py,
 * and by nothing else.  The product library (subquokka_dev_amd) never links,
 * loads or calls it, and there is no CPU fallback built on it.
 *
 * It restates, in plain C with OpenMP, what the reference computes for
 * runner_type=MEM on one rank (mpi_qbit = 0):
 *   - text -> gate parsing              simulator.cpp:259-356, 542-636, 764-774
 *   - gate constants                    circuit.cpp:2542-3041, 3330-3347
 *   - bit_string zero-bit insertion     MEM_Runner.cpp:90-99
 *   - plain gate-by-gate task loop      MEM_Runner.cpp:101-134
 *   - blocked (subcircuit) task loop    MEM_Runner.cpp:136-185
 *   - per-variant loop shapes           circuit.cpp:143-218 (chunk_gate_mem ...
 *                                       nonchunk_nonchunk_gate_mem), 3457-3556
 *   - gate arithmetic                   circuit.cpp:245-375 (HGATE ... RZZGATE)
 * Every amplitude update uses the same formula, operand order and constants
 * as the reference macros.  Arithmetic is plain IEEE double with contraction
 * disabled at build time (-ffp-contract=off), so results are deterministic.
 *
 * Parity pins (all in tests/): closed forms of the reference's own fixture
 * circuits (h/rx/rzz/qft/qaoa from |0..0>), brute-force dense 2^n x 2^n
 * matrices built from textbook gate definitions, the VSWAP_2_2 identity of
 * gate_check_MEM.py:339-341, the QFT closed form R.F.R, and fixtures produced
 * by the reference's own circuit.cpp gate code (oracle/_ref).
 *
 * Deliberate restriction (the reference's contract, not an extension):
 *   - 2q symmetric gates (SWAP, CP, RZZ, VSWAP_1_1) are applied with sorted
 *     targets.  The reference binds loop variants from the unsorted file order
 *     (circuit.cpp:77-83) and mis-indexes unsorted input; its generators always
 *     sort (circuit_generator.py:37-58).
 *   - U2 requires targ[0] < targ[1] (assert at circuit.cpp:2809).
 *   - VSWAP_2_2 requires t0 < t1 < chunk <= t2 < t3, the only layout its loop
 *     (circuit.cpp:3457-3556) enumerates correctly and the only one the finder
 *     emits (finder.cpp:157-167 sorts by physical position).
 *   - VSWAP_3_3/4_4/6_6 swap raw pattern indices (circuit.cpp:4179-4181) and
 *     U3 parsing is disabled (simulator.cpp:577-578): both are rejected.
 *   - A non-VSWAP block of a subcircuit file must have every target < chunk
 *     (MEM_Runner.cpp:173-182 runs each gate tile-locally).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef struct { double re, im; } cplx;

enum {
    SQO_H = 1, SQO_X, SQO_Y, SQO_Z, SQO_P, SQO_RX, SQO_RY, SQO_RZ, SQO_U1,
    SQO_U2, SQO_SWAP, SQO_CP, SQO_RZZ, SQO_VSWAP_1_1, SQO_VSWAP_2_2
};

enum {
    SQO_OK = 0,
    SQO_ERR_UNKNOWN_GATE = -1,   /* simulator.cpp:593-596 "Not implemented yet." */
    SQO_ERR_QUBIT_INDEX = -2,    /* simulator.cpp:764-774 "Qubit Index Error" */
    SQO_ERR_PARSE = -3,
    SQO_ERR_CONTRACT = -4,       /* input outside the reference's valid layout */
    SQO_ERR_ARG = -5
};

typedef struct {
    int kind;
    int nq;
    int q[4];          /* targets; symmetric 2q gates stored sorted */
    double phi;
    cplx c[16];        /* gate constants (matrix or phases), reference layout */
} sqo_gate;

/* ---- complex helpers: the libstdc++ complex<double> formulas ---- */
static inline cplx cmul(cplx a, cplx b) {
    cplx r = { a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re };
    return r;
}
static inline cplx cadd(cplx a, cplx b) { cplx r = { a.re + b.re, a.im + b.im }; return r; }
static inline cplx csub(cplx a, cplx b) { cplx r = { a.re - b.re, a.im - b.im }; return r; }
static inline cplx rmul(double s, cplx a) { cplx r = { s * a.re, s * a.im }; return r; }

/* H literal, circuit.cpp:247 (not 1/sqrt(2)). */
#define SQO_H_COEFF 0.70710678118

/* ------------------------------------------------------------------ */
/* Parsing: setGate_MEM (simulator.cpp:542-597)                       */
/* ------------------------------------------------------------------ */

static int read_ints(const char **p, int n, int *out) {
    for (int i = 0; i < n; i++) {
        char *end;
        long v = strtol(*p, &end, 10);
        if (end == *p) return SQO_ERR_PARSE;
        out[i] = (int)v;
        *p = end;
    }
    return SQO_OK;
}
static int read_doubles(const char **p, int n, double *out) {
    for (int i = 0; i < n; i++) {
        char *end;
        double v = strtod(*p, &end);
        if (end == *p) return SQO_ERR_PARSE;
        out[i] = v;
        *p = end;
    }
    return SQO_OK;
}

static void sort2(int *q) { if (q[0] > q[1]) { int t = q[0]; q[0] = q[1]; q[1] = t; } }

/* Gate constants exactly as the MEM gate constructors build them. */
static void set_constants(sqo_gate *g, const double *raw) {
    double phi = g->phi;
    switch (g->kind) {
    case SQO_P:   /* Phase_Gate_MEM: phi(cos(phi), sin(phi)) circuit.cpp:2655 */
    case SQO_CP:  /* CPhase_Gate_MEM: exp_iPhi circuit.cpp:2949 */
        g->c[0].re = cos(phi); g->c[0].im = sin(phi);
        break;
    case SQO_RX:  /* circuit.cpp:2708-2709 */
        g->c[0].re = cos(phi / 2); g->c[0].im = 0;
        g->c[1].re = 0;            g->c[1].im = sin(-phi / 2);
        break;
    case SQO_RY:  /* circuit.cpp:2736-2737 */
        g->c[0].re = cos(phi / 2); g->c[0].im = 0;
        g->c[1].re = sin(phi / 2); g->c[1].im = 0;
        break;
    case SQO_RZ:  /* circuit.cpp:2764-2765 */
    case SQO_RZZ: /* circuit.cpp:2990-2991 */
        g->c[0].re = cos(phi / 2);  g->c[0].im = sin(phi / 2);   /* exp_p_iPhi_2 */
        g->c[1].re = cos(-phi / 2); g->c[1].im = sin(-phi / 2);  /* exp_n_iPhi_2 */
        break;
    case SQO_U1:  /* row-major [c0 c1; c2 c3], circuit.cpp:274-277 */
        for (int i = 0; i < 4; i++) { g->c[i].re = raw[2 * i]; g->c[i].im = raw[2 * i + 1]; }
        break;
    case SQO_U2:  /* coefficient bit reversal, circuit.cpp:2811-2819 */
        for (int i = 0; i < 16; i++) {
            int b1 = i & 1, b2 = (i >> 1) & 1, b3 = (i >> 2) & 1, b4 = (i >> 3) & 1;
            int res = (b3 << 3) | (b4 << 2) | (b1 << 1) | b2;
            g->c[i].re = raw[2 * res]; g->c[i].im = raw[2 * res + 1];
        }
        break;
    default:
        break;
    }
}

int sqo_parse_gate(const char *line, int N, int chunk, sqo_gate *g) {
    char name[32];
    int consumed = 0;
    memset(g, 0, sizeof(*g));
    if (sscanf(line, " %31s%n", name, &consumed) != 1) return SQO_ERR_PARSE;
    const char *p = line + consumed;
    double raw[32];
    int rc = SQO_OK;
    struct { const char *n; int kind, nq, has_phi, ncoef; } tab[] = {
        {"H", SQO_H, 1, 0, 0},   {"X", SQO_X, 1, 0, 0},   {"Y", SQO_Y, 1, 0, 0},
        {"Z", SQO_Z, 1, 0, 0},   {"P", SQO_P, 1, 1, 0},   {"RX", SQO_RX, 1, 1, 0},
        {"RY", SQO_RY, 1, 1, 0}, {"RZ", SQO_RZ, 1, 1, 0}, {"U1", SQO_U1, 1, 0, 8},
        {"U2", SQO_U2, 2, 0, 32}, {"SWAP", SQO_SWAP, 2, 0, 0}, {"CP", SQO_CP, 2, 1, 0},
        {"RZZ", SQO_RZZ, 2, 1, 0}, {"VSWAP_1_1", SQO_VSWAP_1_1, 2, 0, 0},
        {"VSWAP_2_2", SQO_VSWAP_2_2, 4, 0, 0},
    };
    int found = -1;
    for (unsigned i = 0; i < sizeof(tab) / sizeof(tab[0]); i++)
        if (strcmp(name, tab[i].n) == 0) { found = (int)i; break; }
    if (found < 0) return SQO_ERR_UNKNOWN_GATE;
    g->kind = tab[found].kind;
    g->nq = tab[found].nq;
    if ((rc = read_ints(&p, g->nq, g->q)) != SQO_OK) return rc;
    for (int i = 0; i < g->nq; i++)
        if (g->q[i] >= N || g->q[i] < 0) return SQO_ERR_QUBIT_INDEX;
    for (int i = 0; i < g->nq; i++)
        for (int j = i + 1; j < g->nq; j++)
            if (g->q[i] == g->q[j]) return SQO_ERR_CONTRACT;
    if (tab[found].has_phi && (rc = read_doubles(&p, 1, &g->phi)) != SQO_OK) return rc;
    if (tab[found].ncoef && (rc = read_doubles(&p, tab[found].ncoef, raw)) != SQO_OK) return rc;
    switch (g->kind) {
    case SQO_SWAP: case SQO_CP: case SQO_RZZ: case SQO_VSWAP_1_1:
        sort2(g->q);
        break;
    case SQO_U2:
        if (g->q[0] > g->q[1]) return SQO_ERR_CONTRACT;   /* circuit.cpp:2809 */
        break;
    case SQO_VSWAP_2_2:
        if (!(g->q[0] < g->q[1] && g->q[1] < chunk && chunk <= g->q[2] && g->q[2] < g->q[3]))
            return SQO_ERR_CONTRACT;
        break;
    default:
        break;
    }
    set_constants(g, raw);
    return SQO_OK;
}

/* ------------------------------------------------------------------ */
/* Index helpers                                                       */
/* ------------------------------------------------------------------ */

/* bit_string, MEM_Runner.cpp:90-99: insert a 0 bit at each (ascending) target. */
long long sqo_bit_string(long long task, const int *targ, int n) {
    long long res = task;
    for (int i = 0; i < n; i++) {
        int x = targ[i];
        long long mask = (1LL << x) - 1;
        res = ((res >> x) << (x + 1)) | (res & mask);
    }
    return res;
}

/* ------------------------------------------------------------------ */
/* Gate arithmetic on one tuple (circuit.cpp:245-375)                  */
/* ------------------------------------------------------------------ */

/* One-qubit update on the pair (o0 = |0>, o1 = |1>). */
static inline void one_qubit(cplx *b, const sqo_gate *g, long long o0, long long o1) {
    cplx q0, q1;
    switch (g->kind) {
    case SQO_H:   /* HGATE */
        q0 = b[o0]; q1 = b[o1];
        b[o0] = rmul(SQO_H_COEFF, cadd(q0, q1));
        b[o1] = rmul(SQO_H_COEFF, csub(q0, q1));
        break;
    case SQO_X:   /* XGATE */
        q0 = b[o0]; q1 = b[o1];
        b[o0] = q1; b[o1] = q0;
        break;
    case SQO_Y: { /* YGATE: q1 * (-1i), q0 * 1i */
        cplx mi = {0.0, -1.0}, pi = {0.0, 1.0};
        q0 = b[o0]; q1 = b[o1];
        b[o0] = cmul(q1, mi); b[o1] = cmul(q0, pi);
        break; }
    case SQO_Z: { /* ZGATE on |1>: *= -1 */
        b[o1].re *= -1; b[o1].im *= -1;
        break; }
    case SQO_P:   /* PGATE on |1> */
        b[o1] = cmul(b[o1], g->c[0]);
        break;
    case SQO_U1:  /* U1GATE */
        q0 = b[o0]; q1 = b[o1];
        b[o0] = cadd(cmul(g->c[0], q0), cmul(g->c[1], q1));
        b[o1] = cadd(cmul(g->c[2], q0), cmul(g->c[3], q1));
        break;
    case SQO_RX:  /* RXGATE */
        q0 = b[o0]; q1 = b[o1];
        b[o0] = cadd(cmul(q0, g->c[0]), cmul(q1, g->c[1]));
        b[o1] = cadd(cmul(q0, g->c[1]), cmul(q1, g->c[0]));
        break;
    case SQO_RY:  /* RYGATE */
        q0 = b[o0]; q1 = b[o1];
        b[o0] = csub(cmul(q0, g->c[0]), cmul(q1, g->c[1]));
        b[o1] = cadd(cmul(q0, g->c[1]), cmul(q1, g->c[0]));
        break;
    case SQO_RZ:  /* RZGATE: |0> *= exp(-i phi/2), |1> *= exp(+i phi/2) */
        b[o0] = cmul(b[o0], g->c[1]);
        b[o1] = cmul(b[o1], g->c[0]);
        break;
    default:
        break;
    }
}

/* Two-qubit update on the quad (o0=|00>, o1=|t0>, o2=|t1>, o3=|t0 t1>). */
static inline void two_qubit(cplx *b, const sqo_gate *g, long long o0, long long o1,
                             long long o2, long long o3) {
    switch (g->kind) {
    case SQO_U2: { /* U2Gate */
        cplx q0 = b[o0], q1 = b[o1], q2 = b[o2], q3 = b[o3];
        const cplx *c = g->c;
        cplx r[4];
        for (int row = 0; row < 4; row++) {
            cplx acc = cmul(c[4 * row + 0], q0);
            acc = cadd(acc, cmul(c[4 * row + 1], q1));
            acc = cadd(acc, cmul(c[4 * row + 2], q2));
            acc = cadd(acc, cmul(c[4 * row + 3], q3));
            r[row] = acc;
        }
        b[o0] = r[0]; b[o1] = r[1]; b[o2] = r[2]; b[o3] = r[3];
        break; }
    case SQO_SWAP: case SQO_VSWAP_1_1: { /* SWAPGate on |01> <-> |10> */
        cplx q0 = b[o1], q1 = b[o2];
        b[o1] = q1; b[o2] = q0;
        break; }
    case SQO_CP:   /* CPGATE on |11> */
        b[o3] = cmul(b[o3], g->c[0]);
        break;
    case SQO_RZZ:  /* RZZGATE */
        b[o0] = cmul(b[o0], g->c[1]);
        b[o1] = cmul(b[o1], g->c[0]);
        b[o2] = cmul(b[o2], g->c[0]);
        b[o3] = cmul(b[o3], g->c[1]);
        break;
    default:
        break;
    }
}

/* ------------------------------------------------------------------ */
/* Per-task loop shapes (circuit.cpp:143-218 and 3457-3556)            */
/* ------------------------------------------------------------------ */

static inline int is_chunk(int t, int chunk) { return t < chunk && t >= 0; }

/* Runs the loop variant bound to g for the task that starts at idx. */
static void run_task(cplx *b, const sqo_gate *g, long long idx, long long chunk_state, int chunk) {
    if (g->nq == 1) {
        long long half_off = 1LL << g->q[0], off = half_off * 2;
        if (is_chunk(g->q[0], chunk)) {
            /* chunk_gate_mem */
            long long o0 = idx, o1 = idx + half_off;
            for (long long i = 0; i < chunk_state; i += off) {
                for (long long j = 0; j < half_off; j++) {
                    one_qubit(b, g, o0, o1);
                    o0++; o1++;
                }
                o0 += half_off; o1 += half_off;
            }
        } else {
            /* nonchunk_gate_mem */
            long long o0 = idx, o1 = idx + half_off;
            for (long long i = 0; i < chunk_state; i++) {
                one_qubit(b, g, o0, o1);
                o0++; o1++;
            }
        }
        return;
    }
    if (g->kind == SQO_VSWAP_2_2) {
        /* VSWAP_Gate_2_2_MEM::run_nonchunk_chunk */
        long long h0 = 1LL << g->q[0], h1 = 1LL << g->q[1];
        long long h2 = 1LL << g->q[2], h3 = 1LL << g->q[3];
        long long off1 = 2 * h1, off0 = 2 * h0;
        long long o0001 = idx + h0, o0010 = idx + h1, o0011 = idx + h0 + h1;
        long long o0100 = idx + h2, o0110 = idx + h1 + h2, o0111 = idx + h0 + h1 + h2;
        long long o1000 = idx + h3, o1001 = idx + h0 + h3, o1011 = idx + h0 + h1 + h3;
        long long o1100 = idx + h2 + h3, o1101 = idx + h0 + h2 + h3, o1110 = idx + h1 + h2 + h3;
        cplx t;
#define SQO_SW(a, c) t = b[a]; b[a] = b[c]; b[c] = t;
        for (long long i = 0; i < chunk_state; i += off1) {
            for (long long j = 0; j < h1; j += off0) {
                for (long long k = 0; k < h0; k++) {
                    SQO_SW(o0001, o0100) SQO_SW(o0010, o1000) SQO_SW(o0011, o1100)
                    SQO_SW(o0110, o1001) SQO_SW(o0111, o1101) SQO_SW(o1011, o1110)
                    o0001++; o0010++; o0011++; o0100++; o0110++; o0111++;
                    o1000++; o1001++; o1011++; o1100++; o1101++; o1110++;
                }
                o0001 += h0; o0010 += h0; o0011 += h0; o0100 += h0; o0110 += h0; o0111 += h0;
                o1000 += h0; o1001 += h0; o1011 += h0; o1100 += h0; o1101 += h0; o1110 += h0;
            }
            o0001 += h1; o0010 += h1; o0011 += h1; o0100 += h1; o0110 += h1; o0111 += h1;
            o1000 += h1; o1001 += h1; o1011 += h1; o1100 += h1; o1101 += h1; o1110 += h1;
        }
#undef SQO_SW
        return;
    }
    /* two-qubit gates, targets sorted (TWO_QUBIT_GATE_MEM circuit.cpp:2798-2805) */
    long long h0 = 1LL << g->q[0], h1 = 1LL << g->q[1];
    long long off_0 = 2 * h0, off_1 = 2 * h1;
    long long o0 = idx, o1 = idx + h0, o2 = idx + h1, o3 = idx + h0 + h1;
    if (is_chunk(g->q[1], chunk)) {
        /* chunk_chunk_gate_mem */
        for (long long i = 0; i < chunk_state; i += off_1) {
            for (long long j = 0; j < h1; j += off_0) {
                for (long long k = 0; k < h0; k++) {
                    two_qubit(b, g, o0, o1, o2, o3);
                    o0++; o1++; o2++; o3++;
                }
                o0 += h0; o1 += h0; o2 += h0; o3 += h0;
            }
            o0 += h1; o1 += h1; o2 += h1; o3 += h1;
        }
    } else if (is_chunk(g->q[0], chunk)) {
        /* nonchunk_chunk_gate_mem */
        for (long long j = 0; j < chunk_state; j += off_0) {
            for (long long k = 0; k < h0; k++) {
                two_qubit(b, g, o0, o1, o2, o3);
                o0++; o1++; o2++; o3++;
            }
            o0 += h0; o1 += h0; o2 += h0; o3 += h0;
        }
    } else {
        /* nonchunk_nonchunk_gate_mem */
        for (long long k = 0; k < chunk_state; k++) {
            two_qubit(b, g, o0, o1, o2, o3);
            o0++; o1++; o2++; o3++;
        }
    }
}

static int chunk_count_of(const sqo_gate *g, int chunk) {
    int c = 0;
    for (int i = 0; i < g->nq; i++) c += is_chunk(g->q[i], chunk);
    return c;
}

/* One gate, task loop of MEM_Runner::run(vector<Gate*>&), MEM_Runner.cpp:114-131. */
static void apply_gate_tasks(cplx *b, int N, int chunk, const sqo_gate *g) {
    long long chunk_state = 1LL << chunk;
    long long n_task = 1LL << (N - g->nq);
    long long incre = chunk_state / (1LL << chunk_count_of(g, chunk));
    long long n_iter = (n_task + incre - 1) / incre;
    #pragma omp parallel for schedule(static)
    for (long long it = 0; it < n_iter; it++) {
        long long task = it * incre;
        long long idx = sqo_bit_string(task, g->q, g->nq);
        run_task(b, g, idx, chunk_state, chunk);
    }
}

/* ------------------------------------------------------------------ */
/* Public entry points                                                 */
/* ------------------------------------------------------------------ */

static int check_shape(int N, int chunk) {
    /* setSEG asserts, simulator.cpp:98-99 (file_qbit = 0 here) */
    if (N < 1 || N > 40 || chunk < 1 || chunk > N) return SQO_ERR_ARG;
    return SQO_OK;
}

void sqo_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int sqo_get_threads(void) { return omp_get_max_threads(); }

/* Applies one parsed gate in plain mode. */
int sqo_apply_gate(double *state, int N, int chunk, const sqo_gate *g) {
    int rc = check_shape(N, chunk);
    if (rc) return rc;
    apply_gate_tasks((cplx *)state, N, chunk, g);
    return SQO_OK;
}

/* Splits text into lines; calls fn(line) for each non-empty line.  Returns the
 * first error, with the 1-based line number in *err_line. */
static const char *next_line(const char *p, char *buf, size_t cap, int *len_out) {
    if (!*p) return NULL;
    const char *e = strchr(p, '\n');
    size_t n = e ? (size_t)(e - p) : strlen(p);
    if (n >= cap) n = cap - 1;
    memcpy(buf, p, n);
    buf[n] = 0;
    *len_out = (int)n;
    return e ? e + 1 : p + strlen(p);
}

static int blank(const char *s) {
    for (; *s; s++) if (*s != ' ' && *s != '\t' && *s != '\r') return 0;
    return 1;
}

/* Plain circuit (is_subcircuit=0): setupCircuit_MEM (simulator.cpp:601-615) then
 * MEM_Runner::run(vector<Gate*>&).  Gates are applied as they are parsed, in
 * file order. */
int sqo_run_plain_text(double *state, int N, int chunk, const char *text, int *err_line) {
    int rc = check_shape(N, chunk);
    if (rc) return rc;
    char buf[4096];
    int len, line_no = 0;
    const char *p = text;
    while ((p = next_line(p, buf, sizeof buf, &len)) != NULL) {
        line_no++;
        if (blank(buf)) continue;
        sqo_gate g;
        rc = sqo_parse_gate(buf, N, chunk, &g);
        if (rc) { if (err_line) *err_line = line_no; return rc; }
        apply_gate_tasks((cplx *)state, N, chunk, &g);
    }
    return SQO_OK;
}

/* Subcircuit file (is_subcircuit=1): setupSubCircuits_MEM (simulator.cpp:619-636)
 * then MEM_Runner::run(vector<vector<Gate*>>&) (MEM_Runner.cpp:136-185). */
int sqo_run_sub_text(double *state, int N, int chunk, const char *text, int *err_line) {
    int rc = check_shape(N, chunk);
    if (rc) return rc;
    cplx *b = (cplx *)state;
    long long chunk_state = 1LL << chunk;
    char buf[4096];
    int len, line_no = 0;
    const char *p = text;
    sqo_gate *blk = NULL;
    int cap = 0;
    while ((p = next_line(p, buf, sizeof buf, &len)) != NULL) {
        line_no++;
        if (blank(buf)) continue;
        char *end;
        long cnt = strtol(buf, &end, 10);
        if (end == buf || cnt < 0) { rc = SQO_ERR_PARSE; goto fail; }
        if (cnt > cap) {
            cap = (int)cnt;
            sqo_gate *nb = (sqo_gate *)realloc(blk, sizeof(sqo_gate) * (size_t)cap);
            if (!nb) { rc = SQO_ERR_ARG; goto fail; }
            blk = nb;
        }
        for (long i = 0; i < cnt; i++) {
            p = next_line(p, buf, sizeof buf, &len);
            line_no++;
            if (!p) { rc = SQO_ERR_PARSE; goto fail; }
            rc = sqo_parse_gate(buf, N, chunk, &blk[i]);
            if (rc) goto fail;
        }
        if (cnt == 0) continue;
        if (blk[0].kind == SQO_VSWAP_1_1 || blk[0].kind == SQO_VSWAP_2_2) {
            /* VSWAP block: bit_string task loop over the first gate's targets
             * (MEM_Runner.cpp:146-171).  Each gate of the block runs its own loop;
             * the finder always emits one gate per VSWAP block. */
            for (long i = 0; i < cnt; i++) apply_gate_tasks(b, N, chunk, &blk[i]);
        } else {
            for (long i = 0; i < cnt; i++)
                for (int k = 0; k < blk[i].nq; k++)
                    if (blk[i].q[k] >= chunk) { rc = SQO_ERR_CONTRACT; goto fail; }
            long long n_tiles = 1LL << (N - chunk);
            #pragma omp parallel for schedule(static)
            for (long long t = 0; t < n_tiles; t++) {
                long long idx = t * chunk_state;
                for (long i = 0; i < cnt; i++) run_task(b, &blk[i], idx, chunk_state, chunk);
            }
        }
    }
    free(blk);
    return SQO_OK;
fail:
    if (err_line) *err_line = line_no;
    free(blk);
    return rc;
}

/* <psi| sum_i Z_i Z_{i+1 mod N} |psi>, pairwise-summed in FP64 (no reference
 * counterpart; pinned by basis-state closed forms and brute force in tests). */
double sqo_expect_zz_ring(const double *state, int N) {
    const cplx *b = (const cplx *)state;
    long long n = 1LL << N;
    double total = 0.0;
    #pragma omp parallel for schedule(static) reduction(+:total)
    for (long long i = 0; i < n; i++) {
        double p = b[i].re * b[i].re + b[i].im * b[i].im;
        int s = 0;
        for (int q = 0; q < N; q++) {
            int a = (int)((i >> q) & 1), c = (int)((i >> ((q + 1) % N)) & 1);
            s += (a == c) ? 1 : -1;
        }
        total += p * s;
    }
    return total;
}
