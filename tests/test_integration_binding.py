"""The reference-side binding documented for maintainers compiles and links.

The GPU_Runner class and the runner_type=GPU branch that the integration notes
show (the code a maintainer adds to the reference tree) are extracted from the
notes, compiled against the reference's own headers (src/simulator.hpp:85-97,
the circuitRunner interface; src/circuit.hpp, the Gate type) and
include/subquokka_hip.h, and linked against libsubquokka_hip.so.  A change to
either side that breaks the drop-in fails here.  The reference tree is only
read (its headers); nothing of it is built or run.  Skipped when the tree is
absent.
"""
import os
import re
import subprocess

import pytest

import subquokka_dev_amd as sq
from helpers import ROOT

REF_SRC = "/root/reference/src"
NOTES = os.path.join(ROOT, "INTEGRATION.md")

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(REF_SRC, "simulator.hpp")),
                                reason="reference tree not present")


def cpp_blocks():
    text = open(NOTES).read()
    sec = text[text.index("## 2. Binding from the reference's C++ host"):]
    sec = sec[:sec.index("\n## 3.")]
    return re.findall(r"```cpp\n(.*?)```", sec, flags=re.S)


def test_binding_compiles_against_reference_headers_and_links(tmp_path):
    blocks = cpp_blocks()
    assert len(blocks) == 2, "expected the GPU_Runner class and the setupIni branch"
    runner, branch = blocks
    assert "class GPU_Runner : public Simulator::circuitRunner" in runner
    (tmp_path / "GPU_Runner.hpp").write_text(runner)
    # the setupIni branch (simulator.cpp:225-227) inside a stand-in for the
    # if / else-if chain it joins
    main = (
        '#include "GPU_Runner.hpp"\n'
        "struct SEG seg;\n"
        "struct ENV env;\n"
        "static Simulator::circuitRunner *select_runner(const std::string &cir_path) {\n"
        "    Simulator::circuitRunner *Runner = nullptr;\n"
        '    if (env.runner_type == "MEM") {\n'
        "        Runner = nullptr;\n"
        + branch +
        "    return Runner;\n"
        "}\n"
        "int main(int argc, char **argv) {\n"
        "    // never taken on a host without a GPU: the link is what is checked\n"
        "    if (argc > 2) {\n"
        "        seg.N = 4; seg.chunk = 2; env.runner_type = argv[1]; env.is_subcircuit = 0;\n"
        "        Simulator::circuitRunner *r = select_runner(argv[2]);\n"
        "        std::vector<Gate *> plain;\n"
        "        std::vector<std::vector<Gate *>> blocks;\n"
        "        if (r) { r->run(plain); r->run(blocks); delete r; }\n"
        "    }\n"
        '    std::cout << "binding ok" << std::endl;\n'
        "    return 0;\n"
        "}\n")
    (tmp_path / "main.cpp").write_text(main)
    libdir = os.path.dirname(sq.native.LIB_PATH)
    exe = tmp_path / "gpu_runner_binding"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", str(tmp_path), "-I", REF_SRC,
           "-I", os.path.join(ROOT, "include"), str(tmp_path / "main.cpp"), "-o", str(exe),
           "-L", libdir, "-lsubquokka_hip", f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    # the executable loads the engine library and resolves every entry point
    # the binding calls (no device is touched: the runner is not constructed)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "binding ok" in r.stdout, r.stderr[-2000:]
