// simulator.hpp -- C++ host that keeps SubQuokka's inputs and runner selection.
//
// Mirrors the reference's Simulator (simulator.hpp:56-97, simulator.cpp:88-748):
//   Simulator(ini, circuit)  setupIni + setupCir (simulator.cpp:708-724)
//   run()                    times Runner->run and prints "<t>s" (671-706)
//   circuitRunner            run(plain circuit) / run(subcircuits) (85-97)
// Runner selection by [system] runner_type (simulator.cpp:202-228):
//   MEM, MPI_MEM, GPU   -> HBM_Runner: state resident in HBM
//   DirectIO, MPI_DIO,
//   IO, MPI_IO          -> Stream_Runner: state in pinned host memory, streamed
//                          through HBM in 2^resident_qbit-amplitude chunks
//   RDMA                -> error, as in the reference
// The reference's errors (exit(1) after a message) become SimError exceptions
// carrying an sq_status; the quokka CLI turns them back into exit(1).
#pragma once
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "subquokka_hip.h"

namespace sq {

struct SimError : std::runtime_error {
    int code;
    SimError(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// INI file reader with inih semantics: [section], key=value or key: value,
// ';' / '#' comments, case-insensitive section and key names.
class IniReader {
public:
    explicit IniReader(const std::string &path);
    static IniReader from_text(const std::string &text);
    int parse_error() const { return error_; }   // -1: cannot open; >0: first bad line
    std::string get(const std::string &section, const std::string &key, const std::string &dflt) const;
    long get_int(const std::string &section, const std::string &key, long dflt) const;
    bool has(const std::string &section, const std::string &key) const;

private:
    IniReader() = default;
    void parse(const std::string &text);
    int error_ = 0;
    std::map<std::string, std::string> values_;
};

// setSEG (simulator.cpp:88-100)
struct Segments {
    int N = 0, mpi = 0, file = 0, middle = 0, chunk = 0;
};

// setENV (simulator.cpp:140-186) plus this engine's keys.
struct Env {
    std::string runner_type = "IO";
    int is_subcircuit = 1;
    std::string dump_file;
    int MPI_buffer_size = 16;
    int MPI_testing = 0;
    std::string state_paths;
    // engine keys (defaults keep the reference's behaviour)
    std::string precision = "c128";   // c128 | c64
    int fused_k = 0;                   // >0: plan the plain circuit into fused k-qubit blocks
    int fused_passes = 0;              // 1: the plain circuit as fused passes (complex64)
    int resident_qbit = 0;             // Stream_Runner chunk width (0 = min(N, 30))
    uint64_t init_seed = 0;            // 0: |0...0> (reference); >0: seeded Gaussian state
    std::string dump_binary;           // raw little-endian (re, im) doubles of the final state
    int device = 0;
};

class Simulator {
public:
    class circuitRunner {
    public:
        virtual ~circuitRunner() = default;
        virtual void run(std::vector<sq_gate> &circuit) = 0;                    // is_subcircuit=0
        virtual void run(std::vector<std::vector<sq_gate>> &subcircuits) = 0;   // is_subcircuit=1
        virtual sq_engine *engine() = 0;
    };

    Simulator(const std::string &ini, const std::string &cir);
    ~Simulator();
    Simulator(const Simulator &) = delete;
    Simulator &operator=(const Simulator &) = delete;

    double run();   // returns the Runner->run wall time in seconds and prints "<t>s"

    const Segments &seg() const { return seg_; }
    const Env &env() const { return env_; }
    circuitRunner *runner() { return runner_; }
    size_t n_gates() const;

private:
    void setup_ini(const std::string &ini);
    void setup_cir(const std::string &cir);
    void dump();

    Segments seg_;
    Env env_;
    std::vector<sq_gate> circuit_;
    std::vector<std::vector<sq_gate>> subcircuits_;
    circuitRunner *runner_ = nullptr;
};

class HBM_Runner : public Simulator::circuitRunner {
public:
    HBM_Runner(const Segments &seg, const Env &env);
    ~HBM_Runner() override;
    void run(std::vector<sq_gate> &circuit) override;
    void run(std::vector<std::vector<sq_gate>> &subcircuits) override;
    sq_engine *engine() override { return e_; }

protected:
    HBM_Runner() = default;
    void check(int rc, const char *what);
    sq_engine *e_ = nullptr;
    Env env_;
    int n_ = 0;
    bool streamed_ = false;   // out-of-core runner: fused blocks need the state in HBM
};

class Stream_Runner : public HBM_Runner {
public:
    Stream_Runner(const Segments &seg, const Env &env);
};

}  // namespace sq
