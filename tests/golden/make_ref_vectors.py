"""Writes tests/golden/ref_vectors/*: inputs and outputs of the reference's own
compiled gate code (oracle/_ref/libsqref.so: src/circuit.cpp built where it
lies, driven by oracle/ref_driver.cpp) on seeded inputs.

Each case is <name>.txt (the circuit text, MEM grammar) and <name>.npz with
arrays n, chunk, sub (0 plain / 1 subcircuit), state_in, state_out
(complex128).  Run from the repository root after `make -C oracle`:

    python3 tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import subquokka_dev_amd as sq  # noqa: E402
from helpers import RefHarness, random_circuit, random_state, ref_available  # noqa: E402

OUT = os.path.join(HERE, "ref_vectors")

# (name, n, chunk, kind, seed): kind "random" = all gate kinds, "qft" = the
# cir_qft.py order, "sub" = a planner subcircuit (VSWAP blocks) run through the
# reference's blocked loop (MEM_Runner::run(vector<vector<Gate*>>&) restated)
CASES = [
    ("random_n10_c3", 10, 3, "random", 1),
    ("random_n10_c7", 10, 7, "random", 2),
    ("random_n12_c5", 12, 5, "random", 3),
    ("layers_n11_c4", 11, 4, "layers", 4),
    ("qft_n10_c4", 10, 4, "qft", 5),
    ("sub_n12_c6", 12, 6, "sub", 6),
]


def circuit(kind, n, chunk, seed):
    if kind == "random":
        return random_circuit(n, 240, seed=seed)
    if kind == "layers":
        return sq.gen_circuit("random", n, 6, seed=seed)
    if kind == "qft":
        return sq.gen_circuit("qft", n)
    if kind == "sub":
        return sq.plan(random_circuit(n, 200, seed=seed), n, 0, chunk)
    raise ValueError(kind)


def main():
    if not ref_available():
        sys.exit("oracle/_ref/libsqref.so is not built (needs the reference tree)")
    ref = RefHarness()
    os.makedirs(OUT, exist_ok=True)
    for name, n, chunk, kind, seed in CASES:
        text = circuit(kind, n, chunk, seed)
        s0 = random_state(n, 100 + seed)
        out = ref.run(s0, n, text, chunk=chunk, sub=(kind == "sub"))
        with open(os.path.join(OUT, name + ".txt"), "w") as f:
            f.write(text)
        np.savez_compressed(os.path.join(OUT, name + ".npz"), n=np.int64(n), chunk=np.int64(chunk),
                            sub=np.int64(kind == "sub"), state_in=s0, state_out=out)
        print(name, n, chunk, kind, float(np.linalg.norm(out)))


if __name__ == "__main__":
    main()
