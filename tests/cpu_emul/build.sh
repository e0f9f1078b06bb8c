#!/bin/bash
# CPU emulation build of the engine with AddressSanitizer (host code only).
# Output: tests/cpu_emul/emul (git-ignored), or $EMUL_OUT.  SQ_TILE_PASSES=1 turns on the
# LDS tile path so its addressing is checked too.
set -e
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
CSRC="$ROOT/subquokka_dev_amd/csrc"
/opt/rocm/llvm/bin/clang++ -std=c++17 -O1 -g -fsanitize=address -fno-omit-frame-pointer \
  -DSQ_TILE_PASSES=1 -I"$HERE" -I"$ROOT/include" -I"$CSRC" \
  -x c++ "$CSRC/sq_engine.hip" -x c++ "$CSRC/sq_blocks.hip" -x c++ "$CSRC/sq_fused.hip" -x c++ "$CSRC/sq_adjoint.hip" -x c++ "$CSRC/sq_stream.hip" -x c++ "$CSRC/sq_pass.hip" -x c++ "$CSRC/sq_sample.hip" \
  "$CSRC/host/formats.cpp" "$CSRC/host/planner.cpp" "$HERE/emul_main.cpp" \
  -o "${EMUL_OUT:-$HERE/emul}"
