// Driver of the CPU emulation build: runs one circuit through the C ABI with
// the engine compiled for the host (see hip/hip_runtime.h in this folder).
//   emul <plain|sub> <n> <chunk> <c128|c64> <circuit file>
// Exit code 0 means every launch stayed in bounds (AddressSanitizer aborts
// the process on the first bad access).
#include <cstdio>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "subquokka_hip.h"

namespace sqk { alignas(16) unsigned char lds_raw[1 << 16]; }

int main(int argc, char **argv) {
    if (argc != 6) { std::fprintf(stderr, "usage: emul plain|sub n chunk c128|c64 file\n"); return 2; }
    const bool sub = std::string(argv[1]) == "sub";
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
    cfg.total_qbit = n; cfg.chunk_qbit = chunk; cfg.dtype = dtype;
    sq_engine *e = nullptr;
    if ((rc = sq_create(&cfg, &e))) { std::fprintf(stderr, "create rc=%d\n", rc); return 4; }
    rc = sub ? sq_apply_blocks(e, gates.data(), sizes.data(), nb) : sq_apply_gates(e, gates.data(), ng);
    if (rc) { std::fprintf(stderr, "apply rc=%d %s\n", rc, sq_last_error(e)); return 5; }
    std::vector<double> out(2ULL << n);
    if ((rc = sq_get_state(e, out.data(), 1ULL << n, SQ_C128))) { std::fprintf(stderr, "get rc=%d\n", rc); return 6; }
    double nrm = 0;
    if ((rc = sq_norm2(e, &nrm))) return 7;
    sq_destroy(e);
    std::printf("ok gates=%llu blocks=%llu\n", (unsigned long long)ng, (unsigned long long)nb);
    return 0;
}
