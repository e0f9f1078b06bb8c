// simulator.cpp -- see simulator.hpp.
#include "simulator.hpp"
#include "planner.hpp"

#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <vector>

namespace sq {

// ---------------------------------------------------------------------------
// IniReader
// ---------------------------------------------------------------------------

namespace {
std::string lower(std::string s) {
    for (auto &c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}
std::string trim(const std::string &s) {
    size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}
std::string read_file(const std::string &path, bool &ok) {
    std::ifstream f(path, std::ios::binary);
    ok = (bool)f;
    std::stringstream ss;
    if (ok) ss << f.rdbuf();
    return ss.str();
}
}  // namespace

IniReader::IniReader(const std::string &path) {
    bool ok;
    std::string text = read_file(path, ok);
    if (!ok) { error_ = -1; return; }
    parse(text);
}

IniReader IniReader::from_text(const std::string &text) {
    IniReader r;
    r.parse(text);
    return r;
}

void IniReader::parse(const std::string &text) {
    std::stringstream ss(text);
    std::string line, section;
    int ln = 0;
    while (std::getline(ss, line)) {
        ln++;
        std::string t = trim(line);
        if (t.empty() || t[0] == ';' || t[0] == '#') continue;
        if (t[0] == '[') {
            size_t e = t.find(']');
            if (e == std::string::npos) { if (!error_) error_ = ln; continue; }
            section = lower(trim(t.substr(1, e - 1)));
            continue;
        }
        size_t eq = t.find_first_of("=:");
        if (eq == std::string::npos) { if (!error_) error_ = ln; continue; }
        std::string key = lower(trim(t.substr(0, eq)));
        std::string val = trim(t.substr(eq + 1));
        size_t c = val.find(" ;");   // inline comment
        if (c != std::string::npos) val = trim(val.substr(0, c));
        std::string k = section + "=" + key;
        auto it = values_.find(k);
        if (it != values_.end()) it->second += "\n" + val;   // inih joins repeated keys
        else values_[k] = val;
    }
}

std::string IniReader::get(const std::string &section, const std::string &key, const std::string &dflt) const {
    auto it = values_.find(lower(section) + "=" + lower(key));
    return it == values_.end() ? dflt : it->second;
}

long IniReader::get_int(const std::string &section, const std::string &key, long dflt) const {
    std::string v = get(section, key, "");
    const char *p = v.c_str();
    char *end;
    long n = std::strtol(p, &end, 0);
    return end > p ? n : dflt;
}

bool IniReader::has(const std::string &section, const std::string &key) const {
    return values_.count(lower(section) + "=" + lower(key)) != 0;
}

// ---------------------------------------------------------------------------
// Simulator
// ---------------------------------------------------------------------------

Simulator::Simulator(const std::string &ini, const std::string &cir) {
    setup_ini(ini);
    std::cerr << "[Simulator]: Ini setup." << std::endl;
    setup_cir(cir);
    std::cerr << "[Simulator]: Circuit setup." << std::endl;
    std::cerr << "[Simulator]: StateFile setup." << std::endl;
    std::cerr << "[Simulator]: Ready to go." << std::endl;
}

Simulator::~Simulator() { delete runner_; }

size_t Simulator::n_gates() const {
    if (!env_.is_subcircuit) return circuit_.size();
    size_t n = 0;
    for (auto &b : subcircuits_) n += b.size();
    return n;
}

void Simulator::setup_ini(const std::string &ini) {
    IniReader reader(ini);
    if (reader.parse_error() < 0) throw SimError(SQ_ERR_IO, "Error parsing config.ini!");
    const std::string s = "system";
    // setSEG
    seg_.N = (int)reader.get_int(s, "total_qbit", 0);
    seg_.mpi = (int)reader.get_int(s, "mpi_qbit", 0);
    seg_.N -= seg_.mpi;
    seg_.file = (int)reader.get_int(s, "file_qbit", 0);
    seg_.chunk = (int)reader.get_int(s, "chunk_qbit", 0);
    seg_.middle = seg_.N - seg_.file - seg_.chunk;
    if (seg_.chunk < 1) throw SimError(SQ_ERR_ARG, "chunk_qbit must be >= 1 (simulator.cpp:98)");
    if (seg_.N < seg_.file + seg_.chunk) throw SimError(SQ_ERR_ARG, "total_qbit < file_qbit + chunk_qbit (simulator.cpp:99)");
    // setENV
    env_.dump_file = reader.get(s, "dump_file", "");
    env_.runner_type = reader.get(s, "runner_type", "IO");
    env_.is_subcircuit = (int)reader.get_int(s, "is_subcircuit", 1);
    env_.MPI_buffer_size = (int)reader.get_int(s, "MPI_buffe_size", 16);
    env_.MPI_testing = (int)reader.get_int(s, "MPI_testing", 0);
    env_.state_paths = reader.get(s, "state_paths", "");
    env_.precision = reader.get(s, "precision", "c128");
    env_.fused_k = (int)reader.get_int(s, "fused_k", 0);
    env_.fused_passes = (int)reader.get_int(s, "fused_passes", 0);
    env_.resident_qbit = (int)reader.get_int(s, "resident_qbit", 0);
    env_.init_seed = (uint64_t)reader.get_int(s, "init_seed", 0);
    env_.dump_binary = reader.get(s, "dump_binary", "");
    env_.device = (int)reader.get_int(s, "device", 0);
    if (env_.precision != "c128" && env_.precision != "c64")
        throw SimError(SQ_ERR_ARG, "[Config File]: precision must be c128 or c64");
    if (env_.fused_k < 0 || env_.fused_k > 5)
        throw SimError(SQ_ERR_ARG, "[Config File]: fused_k must be 0 (off) or 1..5");
    if (env_.fused_passes != 0 && env_.fused_passes != 1)
        throw SimError(SQ_ERR_ARG, "[Config File]: fused_passes must be 0 or 1");
    if (env_.fused_passes && (env_.precision != "c64" || env_.fused_k))
        throw SimError(SQ_ERR_ARG, "[Config File]: fused_passes needs precision = c64 and fused_k = 0");
    if (seg_.mpi > 0)
        throw SimError(SQ_ERR_UNSUPPORTED, "[Config File]: mpi_qbit > 0 needs several GPUs; this engine runs on one");

    const std::string &rt = env_.runner_type;
    if (rt == "IO" || rt == "MPI_IO" || rt == "DirectIO" || rt == "MPI_DIO") {
        std::cout << (rt == "DirectIO" || rt == "MPI_DIO" ? "DIO Mode\n" : "IO Mode\n");
        runner_ = new Stream_Runner(seg_, env_);
    } else if (rt == "MEM" || rt == "MPI_MEM" || rt == "GPU") {
        std::cout << (rt == "GPU" ? "GPU Mode\n" : "MEM Mode\n");
        runner_ = new HBM_Runner(seg_, env_);
    } else if (rt == "RDMA") {
        throw SimError(SQ_ERR_UNSUPPORTED, "[Config File]: RDMA runner not found.");
    } else {
        throw SimError(SQ_ERR_ARG, "[Config File]: unknown runner_type " + rt);
    }
}

void Simulator::setup_cir(const std::string &cir) {
    bool ok;
    std::string text = read_file(cir, ok);
    if (!ok) throw SimError(SQ_ERR_IO, "[Circuit File]: cannot open " + cir);
    const int total = seg_.N + seg_.mpi;
    uint64_t ng = 0, nb = 0;
    int32_t line = 0;
    int rc = sq_parse_circuit(text.c_str(), total, env_.is_subcircuit, nullptr, 0, &ng, nullptr, 0, &nb, &line);
    if (rc != SQ_OK && rc != SQ_ERR_ARG) {
        const char *why = rc == SQ_ERR_QUBIT_INDEX ? "[setGate]: Qubit Index Error"
                          : rc == SQ_ERR_UNKNOWN_GATE || rc == SQ_ERR_UNSUPPORTED ? "[setGate]: Not implemented yet."
                          : "[setGate]: malformed gate line";
        throw SimError(rc, std::string(why) + " (line " + std::to_string(line) + ")");
    }
    std::vector<sq_gate> gates(std::max<uint64_t>(ng, 1));
    std::vector<uint32_t> sizes(std::max<uint64_t>(nb, 1));
    rc = sq_parse_circuit(text.c_str(), total, env_.is_subcircuit, gates.data(), ng, &ng,
                          env_.is_subcircuit ? sizes.data() : nullptr, env_.is_subcircuit ? nb : 0, &nb, &line);
    if (rc != SQ_OK) throw SimError(rc, "[setGate]: parse failed at line " + std::to_string(line));
    gates.resize(ng);
    if (!env_.is_subcircuit) {
        circuit_ = std::move(gates);
    } else {
        size_t k = 0;
        for (uint64_t b = 0; b < nb; b++) {
            subcircuits_.emplace_back(gates.begin() + k, gates.begin() + k + sizes[b]);
            k += sizes[b];
        }
    }
}

double Simulator::run() {
    std::cerr << "[Simulator]: RUN!!!!!!\n" << std::flush;
    std::cerr << env_.is_subcircuit << std::endl;
    // state preparation is outside the timed region, as the reference's
    // constructor-time allocation is (MEM_Runner.cpp:77-81)
    if (env_.init_seed) {
        const uint64_t n = 1ULL << seg_.N;
        std::vector<double> st(2 * n);
        int rc = sq_gen_state(seg_.N, env_.init_seed, st.data());
        if (rc == SQ_OK) rc = sq_set_state(runner_->engine(), st.data(), n, SQ_C128);
        if (rc != SQ_OK) throw SimError(rc, std::string("initial state: ") + sq_last_error(runner_->engine()));
    }
    int prc = sq_synchronize(runner_->engine());
    if (prc != SQ_OK) throw SimError(prc, std::string("setup: ") + sq_last_error(runner_->engine()));
    auto t0 = std::chrono::steady_clock::now();
    if (env_.is_subcircuit) runner_->run(subcircuits_);
    else runner_->run(circuit_);
    int src = sq_synchronize(runner_->engine());
    if (src != SQ_OK) throw SimError(src, std::string("run: ") + sq_last_error(runner_->engine()));
    auto t1 = std::chrono::steady_clock::now();
    double secs = std::chrono::duration<double>(t1 - t0).count();
    std::cout << secs << "s" << std::endl;
    dump();
    return secs;
}

// Text dump as simulator.cpp:696-704 (11 decimals, re and im on separate
// lines, file name <first three chars of dump_file>0.txt) when MPI_testing is
// set; binary dump when dump_binary is set.
void Simulator::dump() {
    const bool text = !env_.dump_file.empty() && env_.MPI_testing &&
                      (env_.runner_type == "MEM" || env_.runner_type == "MPI_MEM" || env_.runner_type == "GPU");
    if (!text && env_.dump_binary.empty()) return;
    const uint64_t n = 1ULL << seg_.N;
    std::vector<double> st(2 * n);
    int rc = sq_get_state(runner_->engine(), st.data(), n, SQ_C128);
    if (rc != SQ_OK) throw SimError(rc, std::string("dump: ") + sq_last_error(runner_->engine()));
    if (text) {
        std::string name = env_.dump_file.substr(0, 3) + "0.txt";
        std::ofstream f(name);
        f << std::fixed << std::setprecision(11);
        for (uint64_t i = 0; i < n; i++) f << st[2 * i] << "\n" << st[2 * i + 1] << "\n";
    }
    if (!env_.dump_binary.empty()) {
        std::ofstream f(env_.dump_binary, std::ios::binary);
        f.write(reinterpret_cast<const char *>(st.data()), (std::streamsize)(st.size() * sizeof(double)));
        if (!f) throw SimError(SQ_ERR_IO, "dump_binary: write failed");
    }
}

// ---------------------------------------------------------------------------
// Runners
// ---------------------------------------------------------------------------

HBM_Runner::HBM_Runner(const Segments &seg, const Env &env) : env_(env), n_(seg.N) {
    sq_config cfg{};
    cfg.total_qbit = seg.N;
    cfg.chunk_qbit = seg.chunk;
    cfg.dtype = env.precision == "c64" ? SQ_C64 : SQ_C128;
    cfg.device = env.device;
    cfg.resident_qbit = 0;
    int rc = sq_create(&cfg, &e_);
    if (rc != SQ_OK) throw SimError(rc, "sq_create failed (state of " + std::to_string(seg.N) + " qubits)");
}

HBM_Runner::~HBM_Runner() {
    if (e_) sq_destroy(e_);
}

void HBM_Runner::check(int rc, const char *what) {
    if (rc != SQ_OK) throw SimError(rc, std::string(what) + ": " + sq_last_error(e_));
}

// MEM_Runner::run(vector<Gate*>&) (MEM_Runner.cpp:101-134).  With fused_k > 0
// (state in HBM) the planner groups the gates into blocks of at most fused_k
// qubits, each composed in FP64 and applied in one MFMA pass; with
// fused_passes = 1 (complex64) the engine runs the list as fused passes
// (four-qubit blocks in LDS-resident tiles).  The planning is inside the timed
// region, like every other host step of the run.
void HBM_Runner::run(std::vector<sq_gate> &circuit) {
    if (env_.fused_passes && !streamed_) {
        check(sq_apply_gates_fused(e_, circuit.data(), circuit.size(), 0), "sq_apply_gates_fused");
        return;
    }
    if (env_.fused_k > 0 && !streamed_) {
        std::vector<sq_gate> order;
        std::vector<uint32_t> sizes;
        for (const auto &p : sq::plan_passes(circuit, n_, env_.fused_k, {})) {
            for (int gi : p.gates) order.push_back(circuit[gi]);
            sizes.push_back((uint32_t)p.gates.size());
        }
        check(sq_run_fused(e_, order.data(), sizes.data(), sizes.size()), "sq_run_fused");
        return;
    }
    check(sq_apply_gates(e_, circuit.data(), circuit.size()), "sq_apply_gates");
}

// MEM_Runner::run(vector<vector<Gate*>>&) (MEM_Runner.cpp:136-185)
void HBM_Runner::run(std::vector<std::vector<sq_gate>> &subcircuits) {
    std::vector<sq_gate> flat;
    std::vector<uint32_t> sizes;
    for (auto &b : subcircuits) {
        flat.insert(flat.end(), b.begin(), b.end());
        sizes.push_back((uint32_t)b.size());
    }
    check(sq_apply_blocks(e_, flat.data(), sizes.data(), sizes.size()), "sq_apply_blocks");
}

Stream_Runner::Stream_Runner(const Segments &seg, const Env &env) {
    env_ = env;
    streamed_ = true;
    n_ = seg.N;
    sq_config cfg{};
    cfg.total_qbit = seg.N;
    cfg.chunk_qbit = seg.chunk;
    cfg.dtype = env.precision == "c64" ? SQ_C64 : SQ_C128;
    cfg.device = env.device;
    if (env.state_paths.empty()) {
        // state in host memory (no state files named)
        cfg.resident_qbit = env.resident_qbit > 0 ? env.resident_qbit : std::min(seg.N, 30);
        int rc = sq_create(&cfg, &e_);
        if (rc != SQ_OK) throw SimError(rc, "sq_create (out-of-core) failed");
        return;
    }
    // state files: 2^file_qbit comma-separated paths (setupFiles,
    // simulator.cpp:102-129); their directories are created as needed
    std::vector<std::string> paths;
    std::stringstream ss(env.state_paths);
    std::string tok;
    while (std::getline(ss, tok, ',')) if (!tok.empty()) paths.push_back(tok);
    const size_t nf = (size_t)1 << seg.file;
    if (paths.size() < nf) throw SimError(SQ_ERR_ARG, "[Config File]: state_paths lists fewer than 2^file_qbit files");
    paths.resize(nf);
    for (const auto &p : paths) {
        const size_t cut = p.find_last_of('/');
        if (cut == std::string::npos || cut == 0) continue;
        std::string dir;
        for (size_t i = 0; i <= cut; i++) {
            if (p[i] == '/' && !dir.empty()) mkdir(dir.c_str(), 0777);
            dir += p[i];
        }
    }
    std::vector<const char *> cp;
    for (const auto &p : paths) cp.push_back(p.c_str());
    cfg.resident_qbit = env.resident_qbit > 0 ? env.resident_qbit : std::min(seg.N - 1, 30);
    const bool direct = env.runner_type == "DirectIO" || env.runner_type == "MPI_DIO";
    int rc = sq_create_file_state(&cfg, cp.data(), (int32_t)cp.size(), direct ? 1 : 0, 0, &e_);
    if (rc != SQ_OK) throw SimError(rc, "sq_create_file_state failed (see the message above)");
}

}  // namespace sq
