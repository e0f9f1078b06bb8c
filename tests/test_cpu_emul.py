"""Memory safety of the engine's launches, checked on the CPU.

tests/cpu_emul/build.sh compiles the engine sources for the host with a CPU
stand-in for the HIP runtime and AddressSanitizer; every launch then runs
block by block and any out-of-bounds access aborts the process.  This covers
the subcircuit sweep (with the LDS tile pass turned on) and the VSWAP
permutations, whose GPU tests are skipped (see test_gpu_parity.py).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import subquokka_dev_amd as sq
from helpers import GOLDEN, ROOT, random_circuit

EMUL_DIR = os.path.join(ROOT, "tests", "cpu_emul")
EMUL = os.path.join(EMUL_DIR, "emul")
CLANG = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.exists(CLANG), reason="clang++ for the emulation build not found")


@pytest.fixture(scope="module")
def emul():
    # parallel test workers each build once; the lock and the rename keep one
    # worker from running the binary while another rewrites it
    import fcntl
    with open(os.path.join(EMUL_DIR, ".build.lock"), "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        tmp = f"{EMUL}.tmp{os.getpid()}"
        subprocess.run(["bash", os.path.join(EMUL_DIR, "build.sh")], check=True, env=dict(os.environ, EMUL_OUT=tmp))
        os.replace(tmp, EMUL)
    return EMUL


def run_case(emul, tmp_path, mode, n, chunk, dtype, text, seed=None, tier=None, pass_qbits=None, sweep_mode=None,
             lane_moves=None):
    f = tmp_path / "c.txt"
    f.write_text(text)
    out = tmp_path / f"state_{tier or 'default'}.bin"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    if seed is not None:
        env["EMUL_INIT_SEED"] = str(seed)
    if pass_qbits is not None:
        env["EMUL_PASS_QBITS"] = str(pass_qbits)
    if sweep_mode is not None:
        env["EMUL_SWEEP_MODE"] = str(sweep_mode)
    if lane_moves is not None:
        env["EMUL_LANE_MOVES"] = str(lane_moves)
    if tier is not None:
        env["EMUL_HOST_TIER"] = tier
        env["EMUL_FILES_DIR"] = str(tmp_path)
    r = subprocess.run([emul, mode, str(n), str(chunk), dtype, str(f), str(out)], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert r.stdout.startswith("ok")
    raw = np.fromfile(out, dtype=np.float64)
    return raw[0::2] + 1j * raw[1::2]


@pytest.mark.parametrize("sweep_mode", [0, 1])   # 1: every gate block as an LDS tile
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_sub_fixture_qft21_tiles(emul, tmp_path, dtype, sweep_mode):
    text = open(os.path.join(GOLDEN, "ref_circuits", "sub_qft21.txt")).read()
    run_case(emul, tmp_path, "sub", 21, 12, dtype, text, sweep_mode=sweep_mode)


@pytest.mark.parametrize("sweep_mode", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_planned_blocks_tiles(emul, tmp_path, dtype, sweep_mode):
    n = 14
    text = sq.plan(random_circuit(n, 200, seed=3), n, 1, 5)
    run_case(emul, tmp_path, "sub", n, 12, dtype, text, sweep_mode=sweep_mode)


def test_wide_block_split(emul, tmp_path):
    n = 16
    text = random_circuit(n, 150, seed=5)
    run_case(emul, tmp_path, "sub", n, 12, "c128", f"{len(text.strip().splitlines())}\n" + text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_finder_plan_with_vswaps(emul, tmp_path, dtype):
    n = 16
    text = sq.plan(random_circuit(n, 150, seed=7), n, 3, 10)
    assert "VSWAP_2_2" in text
    run_case(emul, tmp_path, "sub", n, 10, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [4, 9, 13])
def test_plain_all_gate_kinds(emul, tmp_path, dtype, n):
    text = random_circuit(n, 120, seed=n) + "VSWAP_1_1 0 %d\nVSWAP_2_2 0 1 %d %d\n" % (n - 1, n - 2, n - 1)
    run_case(emul, tmp_path, "plain", n, n, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [6, 9, 12])
def test_fused_blocks(emul, tmp_path, dtype, n):
    text = sq.plan(random_circuit(n, 150, seed=n + 1), n, 1, 5)
    run_case(emul, tmp_path, "fused", n, n, dtype, text)


def test_fused_small_state_numerics(emul, tmp_path, oracle):
    # Below 9 qubits the fused path is scalar code end to end (FP64 composition,
    # one thread per column), so the emulation computes it exactly.
    n = 7
    plain = random_circuit(n, 120, seed=17)
    text = sq.plan(plain, n, 1, 5)
    got = run_case(emul, tmp_path, "fused", n, n, "c128", text, seed=3)
    want = oracle.run(sq.gen_state(n, 3), n, plain)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_cnot_and_diagonal_u2_in_register_passes(emul, tmp_path, oracle, dtype):
    # CNOTs written as U2 run as H CZ H inside register passes, diagonal U2 as
    # diagonal ops.  Every gate acts on qubits 7-12 (register and tile bits,
    # no lane exchange), so the emulation computes the values exactly: they
    # must equal the oracle's dense 4x4 application.
    import random as _random
    from helpers import gate_line
    from test_gpu_parity import cnot_line
    n = 13
    rng = _random.Random(5)
    nrng = np.random.default_rng(5)
    lines = []
    for _ in range(160):
        a, b = rng.sample(range(7, n), 2)
        r = rng.random()
        if r < 0.4:
            lines.append(cnot_line(a, b))
        elif r < 0.5:
            ph = rng.uniform(0, 6.28)
            d = [1, 1, 1, complex(np.cos(ph), np.sin(ph))]
            m = []
            for i in range(4):
                for j in range(4):
                    v = complex(d[i] if i == j else 0)
                    m += [v.real, v.imag]
            lines.append(f"U2 {min(a, b)} {max(a, b)} " + " ".join(repr(float(x)) for x in m))
        else:
            lines.append(gate_line(rng.choice(["U1", "H", "RX"]), [a], rng, nrng))
    text = "\n".join(lines) + "\n"
    got = run_case(emul, tmp_path, "plain", n, n, dtype, text, seed=9)
    want = oracle.run(sq.gen_state(n, 9), n, text)
    assert np.max(np.abs(got - want)) <= (1e-12 if dtype == "c128" else 2e-6)


def test_fold_census_covers_every_form(emul, tmp_path):
    # The op census of the complex128 register-pass program (host side, so the
    # emulation computes it exactly, with the engine's default three lane
    # moves): the circuit the GPU fold test runs holds every folded form --
    # targets on register slots, lane bits 0-3 (DPP) and 4-5 (permlane swap),
    # controls in register slot 0 (swapped into lane bit 4/5 by the permlane
    # path), another slot, a lane bit and a tile bit.
    from helpers import fold_pairs_circuit
    n = 14
    f = tmp_path / "c.txt"
    f.write_text(fold_pairs_circuit(n, 400, seed=0))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", EMUL_CENSUS="1", EMUL_LANE_MOVES="3")
    r = subprocess.run([emul, "plain", str(n), str(n), "c128", str(f)], capture_output=True, text=True, timeout=600,
                       env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    census = [int(x) for x in next(l for l in r.stdout.splitlines() if l.startswith("census")).split()[1:]]
    assert len(census) == 20
    fold = np.array(census[8:20]).reshape(3, 4)
    assert (fold > 0).all(), fold
    assert census[0] > 0 and census[1] > 0 and census[2] > 0 and census[4] > 0 and census[5] > 0


@pytest.mark.parametrize("n,tile", [(9, 0), (11, 9), (13, 12), (14, 13), (16, 12)])
def test_fused_pass_launches(emul, tmp_path, n, tile):
    # sq_apply_gates_fused: the tile loads and stores, every block's LDS
    # reads and writes, the variant selection and the A-operand reads stay in
    # bounds (the matrix cores and LDS sharing are not modelled: values are
    # checked on the GPU).  Every gate kind, so diagonal gates with qubits
    # outside the tile (variants) and tile scalars occur.
    text = random_circuit(n, 300, seed=5 * n + tile)
    run_case(emul, tmp_path, "fpass", n, tile, "c64", text)


def test_fused_pass_entry_checks(emul, tmp_path):
    # complex128 engines and states below 9 qubits are refused
    text = random_circuit(9, 20, seed=1)
    f = tmp_path / "c.txt"
    f.write_text(text)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    for n, dtype in ((9, "c128"), (8, "c64")):
        t = random_circuit(n, 20, seed=1)
        f.write_text(t)
        r = subprocess.run([emul, "fpass", str(n), "0", dtype, str(f)], capture_output=True, text=True, timeout=300,
                           env=env)
        assert r.returncode == 5 and "apply rc=" in r.stderr, (r.stdout + r.stderr)[-2000:]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,layers", [(2, 1), (6, 3), (12, 2)])
def test_adjoint_launches(emul, tmp_path, dtype, n, layers):
    f = tmp_path / "empty.txt"
    f.write_text("")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([emul, "adjoint", str(n), str(layers), dtype, str(f)], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0 and r.stdout.startswith("ok"), (r.stdout + r.stderr)[-3000:]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,w", [(12, 9), (14, 10), (13, 12)])
def test_out_of_core_launches(emul, tmp_path, dtype, n, w):
    text = sq.gen_circuit("random", n, 3, seed=n) + random_circuit(n, 60, seed=w)
    run_case(emul, tmp_path, "ooc", n, w, dtype, text)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,w", [(12, 9), (14, 10), (15, 12)])
def test_out_of_core_staged_tiers_match_pinned(emul, tmp_path, dtype, n, w):
    # the emulated kernels are deterministic, so every host tier (pinned
    # state, pageable state + staging, state files + staging) must leave the
    # same bits: this checks the gather/scatter and file addressing on the CPU
    text = sq.gen_circuit("random", n, 3, seed=n) + random_circuit(n, 60, seed=w)
    base = run_case(emul, tmp_path, "ooc", n, w, dtype, text, seed=7)
    for tier in ("pageable", "files"):
        got = run_case(emul, tmp_path, "ooc", n, w, dtype, text, seed=7, tier=tier)
        assert np.array_equal(got, base), tier
    files = sorted(tmp_path.glob("state[0-9].bin"))
    assert len(files) == 4 and all(f.stat().st_size == (16 if dtype == "c128" else 8) << (n - 2) for f in files)


def register_pass_circuit(n, n_gates, seed, first_dense):
    """Dense one-qubit gates on qubits >= first_dense only (no lane exchange),
    diagonal one- and two-qubit gates anywhere: every register-pass op kind
    except the lane exchange, which the emulation does not model."""
    import random as _random
    from helpers import gate_line
    rng = _random.Random(seed)
    nrng = np.random.default_rng(seed)
    lines = []
    for _ in range(n_gates):
        k = rng.choice(["H", "X", "Y", "U1", "U3", "RX", "RY", "Z", "P", "RZ", "CP", "RZZ", "CP", "RZZ"])
        if k in ("CP", "RZZ"):
            qs = sorted(rng.sample(range(n), 2))
        elif k in ("Z", "P", "RZ"):
            qs = [rng.randrange(n)]
        else:
            qs = [rng.randrange(first_dense, n)]
        lines.append(gate_line(k, qs, rng, nrng))
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("kh", [3, 5])
@pytest.mark.parametrize("n", [12, 14])
def test_register_pass_folded_phases(emul, tmp_path, oracle, n, kh):
    # complex128 passes fold a controlled phase into the next dense gate on a
    # register qubit; the control may sit in a register slot, a lane bit or a
    # tile bit.  Each CP below is followed by a U3-form gate on one of its
    # qubits (dense targets >= 6 only: the emulation has no lane exchange).
    import random as _random
    from helpers import gate_line
    rng = _random.Random(7 * n + kh)
    nrng = np.random.default_rng(n)
    lines = []
    for _ in range(120):
        a = rng.randrange(6, n)
        b = rng.choice([rng.randrange(0, 6), rng.randrange(6, n)])
        if a == b:
            continue
        lines.append(f"CP {min(a, b)} {max(a, b)} {rng.uniform(0, 6.28)!r}")
        lines.append(gate_line(rng.choice(["U3", "U1", "H", "RY"]), [a], rng, nrng))
    text = "\n".join(lines) + "\n"
    got = run_case(emul, tmp_path, "plain", n, 4, "c128", text, seed=n, pass_qbits=kh)
    want = oracle.run(sq.gen_state(n, n), n, text)
    assert np.max(np.abs(got - want)) <= 1e-12


@pytest.mark.parametrize("moves", [1, 3])
@pytest.mark.parametrize("kh", [2, 5])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [12, 14])
def test_register_pass_lane_moves_in_bounds(emul, tmp_path, dtype, n, kh, moves):
    # passes that hand the top lane positions to other qubits: every row and
    # lane offset stays inside the state (AddressSanitizer).  The values are
    # not checked here: lane exchanges are not modelled on the CPU; the GPU
    # parity tests check them.
    text = register_pass_circuit(n, 250, 3 * n, 0)
    run_case(emul, tmp_path, "plain", n, 4, dtype, text, seed=n, pass_qbits=kh, lane_moves=moves)


@pytest.mark.parametrize("kh", [4, 5])
@pytest.mark.parametrize("dtype,first_dense", [("c128", 6), ("c64", 7), ("c64", 0)])
@pytest.mark.parametrize("n", [11, 13])
def test_register_pass_numerics(emul, tmp_path, oracle, dtype, n, first_dense, kh):
    # register passes computed exactly by the emulation (no wave-shared values):
    # diagonal ops on lane, tile and register bits (free / one-register /
    # two-register forms), dense ops on register bits and, for complex<float>
    # with first_dense = 0, inside the vector -- against the oracle
    if first_dense == 0:
        text = register_pass_circuit(n, 250, 3 * n, 7).replace("\n", "\n", 1)
        extra = "".join(f"H 0\nRX 0 {0.3 + 0.1 * i!r}\n" for i in range(5))
        text = extra + text + extra
    else:
        text = register_pass_circuit(n, 250, 3 * n, first_dense)
    got = run_case(emul, tmp_path, "plain", n, 4, dtype, text, seed=n, pass_qbits=kh)
    want = oracle.run(sq.gen_state(n, n), n, text)
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(got - want)) <= tol


@pytest.mark.parametrize("dtype,first_dense", [("c128", 6), ("c64", 7)])
def test_register_pass_phase_split(emul, tmp_path, oracle, dtype, first_dense):
    # Dense gates with complex coefficients run as phase, real rotation, phase;
    # the phases of all resident qubits are applied together (PO_PHASES, census
    # slot 7).  Split and unsplit programs against the oracle and each other,
    # with the edge cases of the split written out as matrices: a phase-only
    # gate (the rotation is the identity and no op is left), an anti-diagonal
    # one (cos 0), a matrix that is not unitary (kept whole), half a unitary,
    # and one-qubit diagonals on lane and tile qubits between them.
    import math
    n = 13
    text = register_pass_circuit(n, 200, 77, first_dense)
    a, b = 0.7, -1.9
    extra = []
    for q in range(first_dense, n):
        extra.append(f"U1 {q} 1 0 0 0 0 0 {math.cos(a)!r} {math.sin(a)!r}")
        extra.append(f"U1 {q} 0 0 {math.cos(b)!r} {math.sin(b)!r} {math.cos(a)!r} {math.sin(a)!r} 0 0")
        extra.append(f"P {q % first_dense} {0.3 + q!r}")
        extra.append(f"U1 {q} 0.5 0.1 0.2 0.3 0.4 0.5 0.6 0.7")
        # half a unitary: it splits, but its rotation part is not a rotation
        # (c^2 + s^2 = 1/4), so it runs as the real 2x2 product
        t, ph, la = 0.3 + 0.1 * q, 1.1, -0.4
        u = [0.5 * math.cos(t), -0.5 * complex(math.cos(la), math.sin(la)) * math.sin(t),
             0.5 * complex(math.cos(ph), math.sin(ph)) * math.sin(t),
             0.5 * complex(math.cos(ph + la), math.sin(ph + la)) * math.cos(t)]
        extra.append(f"U1 {q} " + " ".join(f"{complex(z).real!r} {complex(z).imag!r}" for z in u))
        extra.append(f"RZ {n - 1 - (q % 3)} {0.2 * q!r}")
    text = text + "\n".join(extra) + "\n" + text
    want = oracle.run(sq.gen_state(n, n), n, text)
    tol = 1e-12 if dtype == "c128" else 2e-5   # the unitary-breaking matrix scales the state up
    got = {}
    for split in (1, 0):
        f = tmp_path / "c.txt"
        f.write_text(text)
        out = tmp_path / f"s{split}.bin"
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", EMUL_CENSUS="1",
                   EMUL_INIT_SEED=str(n), EMUL_PHASE_SPLIT=str(split))
        r = subprocess.run([emul, "plain", str(n), "4", dtype, str(f), str(out)], capture_output=True, text=True,
                           timeout=900, env=env)
        assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
        census = [int(x) for x in next(l for l in r.stdout.splitlines() if l.startswith("census")).split()[1:]]
        assert (census[7] > 0) == bool(split), census
        raw = np.fromfile(out, dtype=np.float64)
        got[split] = raw[0::2] + 1j * raw[1::2]
        scale = np.max(np.abs(want))
        assert np.max(np.abs(got[split] - want)) <= tol * max(1.0, scale), (split, np.max(np.abs(got[split] - want)))
    assert np.max(np.abs(got[1] - got[0])) <= (1e-13 if dtype == "c128" else 2e-5) * max(1.0, np.max(np.abs(want)))


def test_out_of_core_direct_io_matches_pinned(emul, tmp_path, odirect_state_dir):
    # runner_type DirectIO: the state files opened O_DIRECT (create_file,
    # simulator.cpp:71-86), every transfer a 4 KiB-aligned run of the staging
    # buffer (DIO_Runner.cpp:155-164); the bits must equal the pinned tier's
    d, seen = odirect_state_dir
    if d is None:
        pytest.skip("no directory here accepts O_DIRECT: " + "; ".join(seen))
    n, w = 16, 14   # runs of 2^10 amplitudes: 16 / 8 KiB, 4 KiB-aligned for both dtypes
    text = sq.gen_circuit("random", n, 2, seed=4) + random_circuit(n, 60, seed=44)
    for dtype in ("c128", "c64"):
        base = run_case(emul, tmp_path, "ooc", n, w, dtype, text, seed=3)
        os.environ["EMUL_DIRECT_IO"] = "1"
        try:
            f = tmp_path / "c.txt"
            f.write_text(text)
            out = tmp_path / "direct.bin"
            env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", EMUL_INIT_SEED="3",
                       EMUL_HOST_TIER="files", EMUL_FILES_DIR=d, EMUL_EXPECT_TIER="3")
            r = subprocess.run([emul, "ooc", str(n), str(w), dtype, str(f), str(out)], capture_output=True,
                               text=True, timeout=900, env=env)
        finally:
            os.environ.pop("EMUL_DIRECT_IO", None)
        assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
        assert "direct_io=1" in r.stdout
        raw = np.fromfile(out, dtype=np.float64)
        assert np.array_equal(raw[0::2] + 1j * raw[1::2], base)
        files = [os.path.join(d, f"state{i}.bin") for i in range(4)]
        # the files hold the final state in the reference's layout
        dt = np.complex128 if dtype == "c128" else np.complex64
        on_disk = np.concatenate([np.fromfile(p, dtype=dt) for p in files]).astype(np.complex128)
        assert np.array_equal(on_disk, base)


def test_out_of_core_direct_io_rejects_unaligned_runs(emul, tmp_path):
    # runs of 2^(W - h) amplitudes below 4 KiB cannot be O_DIRECT transfers
    n, w = 12, 9
    f = tmp_path / "c.txt"
    f.write_text("H 0\n")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1", EMUL_HOST_TIER="files",
               EMUL_FILES_DIR=str(tmp_path), EMUL_DIRECT_IO="1")
    r = subprocess.run([emul, "ooc", str(n), str(w), "c128", str(f)], capture_output=True, text=True, timeout=300,
                       env=env)
    assert r.returncode == 4 and "4 KiB" in r.stderr, (r.stdout + r.stderr)[-2000:]


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,w", [(16, 13), (13, 12)])
def test_out_of_core_subcircuit_matches_oracle(emul, tmp_path, oracle, dtype, n, w):
    # DIO_Runner::run(subcircuits) (DIO_Runner.cpp:399), the DIO runner's
    # default mode (is_subcircuit = 1, simulator.cpp:149): gate blocks and
    # VSWAP passes streamed through HBM, every host tier the same bits, all of
    # them equal to the oracle's subcircuit sweep.  n = w + 1 puts the VSWAP_2_2
    # qubits above the pass's fixed low run (four free slots are needed).  Both
    # shapes keep qubits 0-7 in every pass's fixed run, so the lane bits hold
    # no gate and the emulation's values are exact.
    chunk = n - 2
    from helpers import high_qubit_subcircuit
    text = high_qubit_subcircuit(n, chunk, seed=n + w)
    want = oracle.run(sq.gen_state(n, 5), n, text, chunk=chunk, sub=True)
    base = run_case(emul, tmp_path, "oocsub", n, w, dtype, text, seed=5)
    tol = 1e-12 if dtype == "c128" else 2e-6
    assert np.max(np.abs(base - want)) <= tol
    for tier in ("pageable", "files"):
        got = run_case(emul, tmp_path, "oocsub", n, w, dtype, text, seed=5, tier=tier)
        assert np.array_equal(got, base), tier


@pytest.mark.parametrize("dtype", ["c128", "c64"])
def test_out_of_core_finder_plan_launches(emul, tmp_path, dtype):
    # a finder-format plan (chunk 10, VSWAP_2_2 / VSWAP_1_1 on any qubits)
    # through the out-of-core runner: every launch in bounds, tiers agree
    n, w = 15, 12
    text = sq.plan(random_circuit(n, 150, seed=9), n, 0, 10)
    assert "VSWAP_2_2" in text
    base = run_case(emul, tmp_path, "oocsub", n, w, dtype, text, seed=2)
    got = run_case(emul, tmp_path, "oocsub", n, w, dtype, text, seed=2, tier="files")
    assert np.array_equal(got, base)


@pytest.mark.parametrize("bitperm", [0, 1])
@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n", [14, 16])
def test_blocked_sweep_final_permutation(emul, tmp_path, oracle, dtype, n, bitperm):
    # sweep_mode 0: VSWAP blocks become a relabelling, and the data moves to
    # the reference's final layout at the end: one bit-permutation pass
    # through LDS and the second state buffer (bitperm 1), or qubit swaps
    # (bitperm 0).  Every access stays in bounds for both.  Gates and VSWAPs
    # act on qubits >= 7 only, so the emulation computes the swap path's
    # values exactly and they must equal the oracle's subcircuit sweep (it
    # does not model LDS shared between threads, so the permutation pass's
    # values are checked on the GPU).
    from helpers import high_qubit_subcircuit
    chunk = n - 3
    text = high_qubit_subcircuit(n, chunk, seed=3 * n + bitperm, n_blocks=9)
    os.environ["EMUL_BITPERM"] = str(bitperm)
    try:
        got = run_case(emul, tmp_path, "sub", n, chunk, dtype, text, seed=n)
    finally:
        os.environ.pop("EMUL_BITPERM", None)
    if not bitperm:
        want = oracle.run(sq.gen_state(n, n), n, text, chunk=chunk, sub=True)
        assert np.max(np.abs(got - want)) <= (1e-12 if dtype == "c128" else 2e-6)


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("n,ckpt", [(4, 0), (11, 3), (13, 7)])
def test_general_adjoint_launches(emul, tmp_path, dtype, n, ckpt):
    # sq_adjoint_grad on a circuit of every gate kind (rotations and phases
    # take angles): the forward's register passes, the checkpoints, the
    # reverse sweep's kernels and register passes on both vectors stay in
    # bounds (values are checked on the GPU)
    text = random_circuit(n, 160, seed=n + ckpt)
    f = tmp_path / "c.txt"
    f.write_text(text)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([emul, "adjgen", str(n), str(ckpt), dtype, str(f)], capture_output=True, text=True,
                       timeout=900, env=env)
    assert r.returncode == 0 and r.stdout.startswith("ok adjgen"), (r.stdout + r.stderr)[-3000:]
