// Driver of the CPU emulation build: runs one circuit through the C ABI with
// the engine compiled for the host (see hip/hip_runtime.h in this folder).
//   emul <plain|sub|fused|fpass|ooc|oocsub|adjoint> <n> <chunk> <c128|c64> <circuit file> [out]
// (fused: subcircuit-format file whose blocks act on at most five qubits;
// fpass: a plain gate list as fused passes, chunk = tile_qbit;
// ooc / oocsub: the out-of-core runner with resident_qbit = chunk, on a plain
// gate list / a subcircuit file)
// Exit code 0 means every launch stayed in bounds (AddressSanitizer aborts
// the process on the first bad access).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <sstream>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"

namespace sqk { alignas(16) unsigned char lds_raw[1 << 16]; }

int main(int argc, char **argv) {
    if (argc != 6 && argc != 7) { std::fprintf(stderr, "usage: emul plain|sub|fused n chunk c128|c64 file [out]\n"); return 2; }
    const std::string mode = argv[1];
    const bool sub = mode == "sub" || mode == "fused" || mode == "oocsub";
    const bool ooc = mode == "ooc" || mode == "oocsub";   // out-of-core: chunk argument = resident_qbit
    const int n = std::atoi(argv[2]), chunk = std::atoi(argv[3]);
    const int dtype = std::string(argv[4]) == "c64" ? SQ_C64 : SQ_C128;
    std::ifstream f(argv[5]);
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    uint64_t ng = 0, nb = 0;
    int32_t line = 0;
    sq_parse_circuit(text.c_str(), n, sub, nullptr, 0, &ng, nullptr, 0, &nb, &line);
    std::vector<sq_gate> gates(ng + 1);
    std::vector<uint32_t> sizes(nb + 1);
    int rc = sq_parse_circuit(text.c_str(), n, sub, gates.data(), ng, &ng, sub ? sizes.data() : nullptr,
                              sub ? nb : 0, &nb, &line);
    if (rc) { std::fprintf(stderr, "parse rc=%d line %d\n", rc, line); return 3; }
    sq_config cfg{};
    cfg.total_qbit = n; cfg.chunk_qbit = (ooc || mode == "adjgen" || mode == "fpass") ? 0 : chunk; cfg.dtype = dtype;
    if (ooc) cfg.resident_qbit = chunk;
    if (const char *kh = std::getenv("EMUL_PASS_QBITS")) cfg.pass_qbits = std::atoi(kh);
    if (const char *sm = std::getenv("EMUL_SWEEP_MODE")) cfg.sweep_mode = std::atoi(sm);
    sq_engine *e = nullptr;
    const char *seed_env = std::getenv("EMUL_INIT_SEED");
    // out-of-core host tier: EMUL_HOST_TIER = pageable | files (files in
    // EMUL_FILES_DIR, opened O_DIRECT when EMUL_DIRECT_IO=1)
    const char *tier = std::getenv("EMUL_HOST_TIER");
    std::vector<std::string> paths;
    if (ooc && tier && std::string(tier) == "pageable") cfg.host_tier = SQ_HOST_PAGEABLE;
    if (ooc && tier && std::string(tier) == "files") {
        const char *dir = std::getenv("EMUL_FILES_DIR");
        for (int i = 0; i < 4; i++) paths.push_back(std::string(dir ? dir : ".") + "/state" + std::to_string(i) + ".bin");
        std::vector<const char *> cp;
        for (auto &p : paths) cp.push_back(p.c_str());
        const char *dio = std::getenv("EMUL_DIRECT_IO");
        rc = sq_create_file_state(&cfg, cp.data(), 4, dio && std::atoi(dio) ? 1 : 0, 0, &e);
    } else {
        rc = sq_create(&cfg, &e);
    }
    if (rc) { std::fprintf(stderr, "create rc=%d\n", rc); return 4; }
    if (const char *want = std::getenv("EMUL_EXPECT_TIER")) {   // the tier the engine reports
        uint64_t passes = 0;
        double bytes = 0;
        int32_t t = -1;
        if (sq_stream_info(e, &passes, &bytes, &t) || t != std::atoi(want)) { std::fprintf(stderr, "tier %d\n", t); return 14; }
    }
    // Cross-lane exchanges are not modelled here (one lane runs at a time), so
    // register passes keep the lane qubits in place unless EMUL_LANE_MOVES asks
    // otherwise (then only the addressing is checked, not the values).
    e->lane_moves = 0;
    if (const char *lm = std::getenv("EMUL_LANE_MOVES")) e->lane_moves = std::atoi(lm);
    if (const char *bp = std::getenv("EMUL_BITPERM")) e->bitperm_ok = std::atoi(bp) != 0;
    if (const char *ps = std::getenv("EMUL_PHASE_SPLIT")) e->phase_split = std::atoi(ps) != 0;
    if (mode == "adjoint") {   // chunk argument = layers; theta from a fixed formula
        const int layers = chunk;
        std::vector<double> theta((size_t)2 * n * layers), grad(theta.size());
        for (size_t i = 0; i < theta.size(); i++) theta[i] = 0.1 + 0.37 * (double)i;
        double energy = 0;
        rc = sq_adjoint_grad_hea(e, layers, theta.data(), theta.size(), 1, &energy, grad.data());
        if (rc) { std::fprintf(stderr, "adjoint rc=%d %s\n", rc, sq_last_error(e)); return 10; }
        rc = sq_hea_energy(e, layers, theta.data(), theta.size(), &energy);
        if (!rc && sq_hea_energy(e, layers, theta.data(), theta.size() - 1, &energy) != SQ_ERR_ARG) return 12;
        if (rc) return 11;
        sq_destroy(e);
        std::printf("ok adjoint params=%zu\n", theta.size());
        return 0;
    }
    if (mode == "adjgen") {   // chunk argument = checkpoint_every; every RX RY RZ P CP RZZ takes an angle
        std::vector<int32_t> param(ng, -1);
        std::vector<double> theta;
        for (uint64_t i = 0; i < ng; i++) {
            const int k = gates[i].kind;
            if (k == SQ_GATE_RX || k == SQ_GATE_RY || k == SQ_GATE_RZ || k == SQ_GATE_P || k == SQ_GATE_CP || k == SQ_GATE_RZZ) {
                param[i] = (int32_t)theta.size();
                theta.push_back(0.1 + 0.37 * (double)i);
            }
        }
        std::vector<sq_zterm> terms;
        for (int q = 0; q < n; q++) terms.push_back({1.0 + q, (1ULL << q) | (1ULL << ((q + 1) % n))});
        std::vector<double> grad(theta.size() + 1);
        double energy = 0;
        rc = sq_adjoint_grad(e, gates.data(), param.data(), ng, theta.data(), theta.size(), terms.data(), terms.size(),
                             chunk, &energy, grad.data());
        if (rc) { std::fprintf(stderr, "adjgen rc=%d %s\n", rc, sq_last_error(e)); return 15; }
        if ((rc = sq_expect_z(e, terms.data(), terms.size(), &energy))) return 16;
        sq_destroy(e);
        std::printf("ok adjgen params=%zu\n", theta.size());
        return 0;
    }
    if (seed_env) {   // seeded Gaussian initial state instead of |0...0>
        std::vector<double> st(2ULL << n);
        sq_gen_state(n, std::strtoull(seed_env, nullptr, 10), st.data());
        if ((rc = sq_set_state(e, st.data(), 1ULL << n, SQ_C128))) return 9;
    }
    if (mode == "fused") rc = sq_run_fused(e, gates.data(), sizes.data(), nb);
    else if (mode == "fpass") rc = sq_apply_gates_fused(e, gates.data(), ng, chunk);
    else rc = sub ? sq_apply_blocks(e, gates.data(), sizes.data(), nb) : sq_apply_gates(e, gates.data(), ng);
    if (rc) { std::fprintf(stderr, "apply rc=%d %s\n", rc, sq_last_error(e)); return 5; }
    if (std::getenv("EMUL_CENSUS")) {   // op census of the last register-pass program
        int64_t cnt[20];
        int32_t nc = 0;
        if (sq_pass_census(e, cnt, 20, &nc)) return 15;
        std::printf("census");
        for (int i = 0; i < nc; i++) std::printf(" %lld", (long long)cnt[i]);
        std::printf("\n");
    }
    std::vector<double> out(2ULL << n);
    if ((rc = sq_get_state(e, out.data(), 1ULL << n, SQ_C128))) { std::fprintf(stderr, "get rc=%d\n", rc); return 6; }
    if (argc == 7) {   // optional state dump (interleaved doubles)
        FILE *fo = std::fopen(argv[6], "wb");
        if (!fo || std::fwrite(out.data(), sizeof(double), out.size(), fo) != out.size()) return 8;
        std::fclose(fo);
    }
    double nrm = 0;
    if ((rc = sq_norm2(e, &nrm))) return 7;
    // measurement shots (addresses only: wave-shared values are not emulated;
    // after fused passes the state is zero here, the matrix cores being unmodelled)
    if (!ooc && mode != "fpass" && nrm > 0) {
        std::vector<uint64_t> shots(777);
        if ((rc = sq_sample(e, 5, shots.size(), shots.data()))) { std::fprintf(stderr, "sample rc=%d %s\n", rc, sq_last_error(e)); return 12; }
        for (uint64_t v : shots) if (v >> n) { std::fprintf(stderr, "shot out of range\n"); return 13; }
    }
    const int direct = e->file_direct ? 1 : 0;
    sq_destroy(e);
    std::printf("ok gates=%llu blocks=%llu direct_io=%d\n", (unsigned long long)ng, (unsigned long long)nb, direct);
    return 0;
}
