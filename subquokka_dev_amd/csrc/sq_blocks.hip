// sq_blocks.hip -- subcircuit (blocked) sweep: MEM_Runner::run(subcircuits)
// (MEM_Runner.cpp:136-185).
//
// A subcircuit file alternates VSWAP blocks (qubit-swap passes) and gate
// blocks whose targets all lie below chunk_qbit.  The reference applies a
// gate block tile by tile: for every 2^chunk-amplitude tile, all gates of the
// block in order (MEM_Runner.cpp:173-182), so the tile stays in cache.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "subquokka_hip.h"
#include "engine_internal.hpp"

using sq::LoweredGate;

extern "C" int sq_apply_blocks(sq_engine *e, const sq_gate *gates, const uint32_t *block_sizes,
                               uint64_t n_blocks) {
    if (!e || (!gates && n_blocks) || (!block_sizes && n_blocks)) return SQ_ERR_ARG;
    if (hipSetDevice(e->device) != hipSuccess) return SQ_ERR_HIP;
    uint64_t total = 0;
    for (uint64_t b = 0; b < n_blocks; b++) total += block_sizes[b];
    std::vector<LoweredGate> low(total);
    for (uint64_t i = 0; i < total; i++) {
        int rc = e->lower(gates[i], low[i]);
        if (rc) { e->err += " (gate " + std::to_string(i) + ")"; return rc; }
    }
    // Gate blocks: every gate is tile-local, so applying the block's gates in
    // order over the whole state is the same map as applying them tile by tile.
    for (uint64_t i = 0; i < total; i++) {
        int rc = e->apply_lowered(low[i]);
        if (rc) return rc;
    }
    return SQ_OK;
}

void sq_engine::free_extras() {
    if (d_blockprog) hipFree(d_blockprog);
    if (d_fused) hipFree(d_fused);
    if (d_aux) hipFree(d_aux);
    if (d_ckpt) hipFree(d_ckpt);
    d_blockprog = d_fused = d_aux = d_ckpt = nullptr;
}

extern "C" int sq_apply_fused(sq_engine *e, const sq_fused *, uint64_t) {
    return e ? e->fail(SQ_ERR_UNSUPPORTED, "fused blocks not built yet") : SQ_ERR_ARG;
}
extern "C" int sq_compose_block(sq_engine *e, const sq_gate *, uint64_t, int32_t, const int32_t *, double *) {
    return e ? e->fail(SQ_ERR_UNSUPPORTED, "compose not built yet") : SQ_ERR_ARG;
}
extern "C" int sq_hea_energy(sq_engine *e, int32_t, const double *, double *) {
    return e ? e->fail(SQ_ERR_UNSUPPORTED, "adjoint not built yet") : SQ_ERR_ARG;
}
extern "C" int sq_adjoint_grad_hea(sq_engine *e, int32_t, const double *, int32_t, double *, double *) {
    return e ? e->fail(SQ_ERR_UNSUPPORTED, "adjoint not built yet") : SQ_ERR_ARG;
}
extern "C" int sq_plan(const char *, int32_t, int32_t, int32_t, char *, uint64_t, uint64_t *) {
    return SQ_ERR_UNSUPPORTED;
}
