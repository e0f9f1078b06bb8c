// ref_driver.cpp -- harness that runs the reference's own MEM gate classes.
//
// TEST INFRASTRUCTURE ONLY (see sq_oracle.c for who may load it).
//
// Built by oracle/Makefile together with circuit.cpp taken from the reference
// tree where it lies; the output goes to oracle/_ref/ only.  circuit.cpp needs
// only the globals and segment predicates that simulator.cpp defines
// (simulator.cpp:28-49); simulator.cpp and MEM_Runner.cpp themselves include
// <mpi.h>, which this image lacks, so this harness defines those few globals
// and restates the MEM_Runner task loops (MEM_Runner.cpp:90-185) around the
// reference's Gate::_run bindings.  The gate arithmetic, constants and
// per-variant loop shapes are the reference's own compiled code.
#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>
#include <complex>
#include <omp.h>

#include "circuit.hpp"
#include "simulator.hpp"

struct SEG seg;
struct ENV env;

// Segment predicates, same definitions as simulator.cpp:31-49.
bool isMpi(int t) { return t >= seg.N; }
bool isFile(int t) { return (t >= (seg.middle + seg.chunk)) && (t < seg.N); }
bool isMiddle(int t) { return (t < (seg.middle + seg.chunk)) && (t >= seg.chunk); }
bool isChunk(int t) { return t < seg.chunk && t >= 0; }

namespace {

// setSEG + the parts of setENV the MEM gates read (simulator.cpp:88-100, 140-175),
// with mpi_qbit = file_qbit = 0.
void setup_env(int N, int chunk) {
    seg.N = N; seg.mpi = 0; seg.file = 0; seg.chunk = chunk;
    seg.middle = N - chunk;
    env.runner_type = "MEM";
    env.is_MPI = 0; env.rank = 0; env.num_file = 1; env.num_thread = omp_get_max_threads();
    env.file_state = 1LL << N; env.thread_state = env.file_state;
    env.chunk_state = 1LL << chunk;
    env.chunk_size = env.chunk_state * (long long)sizeof(std::complex<double>);
    env.qubit_offset.clear(); env.qubit_size.clear();
    long long off = 1, sz = (long long)sizeof(std::complex<double>);
    for (int i = 0; i <= N; i++) { env.qubit_offset.push_back(off); env.qubit_size.push_back(sz); off *= 2; sz *= 2; }
}

// Gate construction by name, the MEM branch of setGate_MEM (simulator.cpp:542-597).
Gate *make_gate(const std::string &line, int &rc) {
    std::stringstream ss(line);
    std::string op;
    ss >> op;
    auto ints = [&](int n) { std::vector<int> t(n); for (auto &x : t) ss >> x; return t; };
    rc = 0;
    if (op == "H") return new H_Gate_MEM(ints(1));
    if (op == "X") return new X_Gate_MEM(ints(1));
    if (op == "Y") return new Y_Gate_MEM(ints(1));
    if (op == "Z") return new Z_Gate_MEM(ints(1));
    if (op == "P" || op == "RX" || op == "RY" || op == "RZ") {
        auto t = ints(1); double phi; ss >> phi;
        if (op == "P") return new Phase_Gate_MEM(t, phi);
        if (op == "RX") return new RX_Gate_MEM(t, phi);
        if (op == "RY") return new RY_Gate_MEM(t, phi);
        return new RZ_Gate_MEM(t, phi);
    }
    if (op == "CP" || op == "RZZ") {
        auto t = ints(2); double phi; ss >> phi;
        if (op == "CP") return new CPhase_Gate_MEM(t, phi);
        return new RZZ_Gate_MEM(t, phi);
    }
    if (op == "U1" || op == "U2") {
        int nq = op == "U1" ? 1 : 2;
        auto t = ints(nq);
        int n = 1 << (2 * nq);
        std::vector<std::complex<double>> c(n);
        for (auto &z : c) { double re, im; ss >> re >> im; z = {re, im}; }
        if (nq == 1) return new U1_Gate_MEM(t, c);
        return new U2_Gate_MEM(t, c);
    }
    if (op == "SWAP") return new SWAP_Gate_MEM(ints(2));
    if (op == "VSWAP_1_1") return new VSWAP_Gate_1_1_MEM(ints(2));
    if (op == "VSWAP_2_2") return new VSWAP_Gate_2_2_MEM(ints(4));
    rc = -1;
    return nullptr;
}

inline long long bit_string(long long task, const std::vector<int> &targ) {
    long long res = task;
    for (auto x : targ) {
        long long mask = (1LL << x) - 1;
        res = ((res >> x) << (x + 1)) | (res & mask);
    }
    return res;
}

// One gate (or the gates of one VSWAP block) over bit_string tasks, as in
// MEM_Runner.cpp:114-131 and 158-170.
void run_tasks(std::vector<std::complex<double>> &buf, std::vector<Gate *> &gs) {
    std::vector<int> targ = gs[0]->targs;
    long long n_task = 1LL << (seg.N - (int)targ.size());
    long long incre = env.chunk_state / (1LL << gs[0]->chunk_count);
    #pragma omp parallel for schedule(static)
    for (long long task = 0; task < n_task; task += incre) {
        long long idx = bit_string(task, targ);
        for (auto &g : gs) g->_run(buf, idx);
    }
}

std::vector<std::string> lines_of(const char *text) {
    std::vector<std::string> out;
    std::stringstream ss(text);
    std::string line;
    while (std::getline(ss, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        out.push_back(line);
    }
    return out;
}

}  // namespace

extern "C" {

void sqref_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

// state: 2^N interleaved (re, im) doubles, updated in place.  If seconds is
// not NULL it receives the wall time of the gate loop alone (omp_get_wtime,
// as Simulator::run times Runner->run, simulator.cpp:679-689).
int sqref_time_plain_text(double *state, int N, int chunk, const char *text, double *seconds) {
    setup_env(N, chunk);
    std::vector<std::complex<double>> buf(1ULL << N);
    std::memcpy(buf.data(), state, sizeof(double) * 2 * buf.size());
    std::vector<Gate *> gates;
    for (auto &line : lines_of(text)) {
        int rc;
        Gate *g = make_gate(line, rc);
        if (!g) { for (auto x : gates) delete x; return rc; }
        gates.push_back(g);
    }
    double t0 = omp_get_wtime();
    for (auto g : gates) {
        std::vector<Gate *> one{g};
        run_tasks(buf, one);
    }
    double t1 = omp_get_wtime();
    if (seconds) *seconds = t1 - t0;
    for (auto g : gates) delete g;
    std::memcpy(state, buf.data(), sizeof(double) * 2 * buf.size());
    return 0;
}

int sqref_run_plain_text(double *state, int N, int chunk, const char *text) {
    return sqref_time_plain_text(state, N, chunk, text, nullptr);
}

// Subcircuit text: "<count>" then count gate lines, repeated.
int sqref_run_sub_text(double *state, int N, int chunk, const char *text) {
    setup_env(N, chunk);
    std::vector<std::complex<double>> buf(1ULL << N);
    std::memcpy(buf.data(), state, sizeof(double) * 2 * buf.size());
    auto lines = lines_of(text);
    size_t i = 0;
    while (i < lines.size()) {
        int cnt = std::stoi(lines[i++]);
        std::vector<Gate *> blk;
        for (int k = 0; k < cnt && i < lines.size(); k++) {
            int rc;
            Gate *g = make_gate(lines[i++], rc);
            if (!g) { for (auto x : blk) delete x; return rc; }
            blk.push_back(g);
        }
        if (blk.empty()) continue;
        if (blk[0]->type == VSWAP) {
            run_tasks(buf, blk);
        } else {
            long long n = 1LL << seg.N;
            #pragma omp parallel for schedule(static)
            for (long long idx = 0; idx < n; idx += env.chunk_state)
                for (auto &g : blk) g->_run(buf, idx);
        }
        for (auto g : blk) delete g;
    }
    std::memcpy(state, buf.data(), sizeof(double) * 2 * buf.size());
    return 0;
}

}  // extern "C"
