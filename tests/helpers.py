"""Shared test helpers: the CPU oracle loader, random circuit text and
textbook dense gate matrices for brute-force checks.

The oracle (oracle/libsq_oracle.so) and the reference harness
(oracle/_ref/libsqref.so) are test infrastructure: only tests, the smoke
check and bench.py's cpu_baseline leg load them.
"""
import ctypes
import cmath
import math
import os
import random
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "libsq_oracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libsqref.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")

H_LITERAL = 0.70710678118   # circuit.cpp:247


def _build_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), os.path.join(ROOT, "oracle", "libsq_oracle.so")],
                       check=True)


class Oracle:
    """FP64 CPU restatement of the MEM path (oracle/sq_oracle.c)."""

    def __init__(self):
        _build_oracle()
        L = ctypes.CDLL(ORACLE_SO)
        L.sqo_run_plain_text.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p,
                                         ctypes.POINTER(ctypes.c_int)]
        L.sqo_run_sub_text.argtypes = L.sqo_run_plain_text.argtypes
        L.sqo_expect_zz_ring.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.sqo_expect_zz_ring.restype = ctypes.c_double
        L.sqo_bit_string.argtypes = [ctypes.c_longlong, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.sqo_bit_string.restype = ctypes.c_longlong
        L.sqo_set_threads.argtypes = [ctypes.c_int]
        self.L = L

    def run(self, state, n, text, chunk=None, sub=False):
        """Returns a new complex128 array: text applied to state."""
        s = np.array(state, dtype=np.complex128, copy=True)
        chunk = n if chunk is None else chunk
        line = ctypes.c_int(0)
        fn = self.L.sqo_run_sub_text if sub else self.L.sqo_run_plain_text
        rc = fn(s.ctypes.data, n, chunk, text.encode(), ctypes.byref(line))
        if rc != 0:
            raise RuntimeError(f"oracle rc={rc} at line {line.value}")
        return s

    def rc(self, state, n, text, chunk=None, sub=False):
        s = np.array(state, dtype=np.complex128, copy=True)
        fn = self.L.sqo_run_sub_text if sub else self.L.sqo_run_plain_text
        return fn(s.ctypes.data, n, n if chunk is None else chunk, text.encode(), None)

    def expect_zz_ring(self, state, n):
        s = np.ascontiguousarray(state, dtype=np.complex128)
        return self.L.sqo_expect_zz_ring(s.ctypes.data, n)

    def bit_string(self, task, targ):
        arr = (ctypes.c_int * len(targ))(*targ)
        return self.L.sqo_bit_string(task, arr, len(targ))


class RefHarness:
    """The reference's own circuit.cpp gate classes (oracle/_ref), if built."""

    def __init__(self):
        L = ctypes.CDLL(REF_SO)
        L.sqref_run_plain_text.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
        L.sqref_run_sub_text.argtypes = L.sqref_run_plain_text.argtypes
        L.sqref_set_threads.argtypes = [ctypes.c_int]
        self.L = L

    def run(self, state, n, text, chunk=None, sub=False):
        s = np.array(state, dtype=np.complex128, copy=True)
        fn = self.L.sqref_run_sub_text if sub else self.L.sqref_run_plain_text
        rc = fn(s.ctypes.data, n, n if chunk is None else chunk, text.encode())
        if rc != 0:
            raise RuntimeError(f"reference harness rc={rc}")
        return s


def ref_available():
    return os.path.exists(REF_SO)


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    s = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return s / np.linalg.norm(s)


ALL_KINDS = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U1", "U2", "SWAP", "CP", "RZZ"]


def random_unitary(d, rng):
    z = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def gate_line(kind, qs, rng_py, nrng):
    """One circuit line of the MEM grammar with random parameters (targets qs)."""
    if kind in ("H", "X", "Y", "Z"):
        return f"{kind} {qs[0]}"
    if kind in ("P", "RX", "RY", "RZ"):
        return f"{kind} {qs[0]} {rng_py.uniform(0, 2 * math.pi)!r}"
    if kind == "U3":   # U3(theta, phi, lambda) written as U1: m0 = cos(theta/2) is real
        th, ph, la = (rng_py.uniform(0, 2 * math.pi) for _ in range(3))
        c, s = math.cos(th / 2), math.sin(th / 2)
        u = [complex(c, 0.0), -cmath.exp(1j * la) * s, cmath.exp(1j * ph) * s, cmath.exp(1j * (ph + la)) * c]
        return f"U1 {qs[0]} " + " ".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in u)
    if kind == "U1":
        u = random_unitary(2, nrng)
        return f"U1 {qs[0]} " + " ".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in u.flatten())
    a, b = qs
    if kind == "U2":
        a, b = sorted(qs)
        u = random_unitary(4, nrng)
        return f"U2 {a} {b} " + " ".join(f"{float(v.real)!r} {float(v.imag)!r}" for v in u.flatten())
    if kind == "SWAP":
        return f"SWAP {a} {b}"
    return f"{kind} {a} {b} {rng_py.uniform(0, 2 * math.pi)!r}"


def fold_pairs_circuit(n, n_pairs, seed):
    """CP on a random pair followed by a dense gate with real coefficients (H,
    RY, X) on one of its qubits: the complex128 register passes fold most of
    these phases into the dense gate's op, in every target/control placement
    at n = 14.  (A dense gate with complex coefficients is split into phases
    and a rotation, which takes no fold.)"""
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    lines = []
    for _ in range(n_pairs):
        a, b = rng.sample(range(n), 2)
        lines.append(f"CP {min(a, b)} {max(a, b)} {rng.uniform(0, 6.28)!r}")
        lines.append(gate_line(rng.choice(["RY", "H", "X"]), [rng.choice([a, b])], rng, nrng))
    return "\n".join(lines) + "\n"


def random_circuit(n, n_gates, seed, kinds=ALL_KINDS, sorted_2q=True):
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    lines = []
    for _ in range(n_gates):
        k = rng.choice(kinds)
        if k in ("U2", "SWAP", "CP", "RZZ"):
            qs = rng.sample(range(n), 2)
            if sorted_2q or k == "U2":
                qs = sorted(qs)
        else:
            qs = [rng.randrange(n)]
        lines.append(gate_line(k, qs, rng, nrng))
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------------------
# textbook matrices (Qiskit conventions, little-endian qubit order)
# ---------------------------------------------------------------------------

def textbook_1q(kind, phi=0.0, mat=None):
    c, s = math.cos(phi / 2), math.sin(phi / 2)
    if kind == "H":
        return H_LITERAL * np.array([[1, 1], [1, -1]], dtype=complex)
    if kind == "X":
        return np.array([[0, 1], [1, 0]], dtype=complex)
    if kind == "Y":
        return np.array([[0, -1j], [1j, 0]])
    if kind == "Z":
        return np.diag([1, -1]).astype(complex)
    if kind == "P":
        return np.diag([1, np.exp(1j * phi)])
    if kind == "RX":
        return np.array([[c, -1j * s], [-1j * s, c]])
    if kind == "RY":
        return np.array([[c, -s], [s, c]], dtype=complex)
    if kind == "RZ":
        return np.diag([np.exp(-1j * phi / 2), np.exp(1j * phi / 2)])
    if kind == "U1":
        return np.asarray(mat, dtype=complex).reshape(2, 2)
    raise KeyError(kind)


def textbook_2q(kind, phi=0.0, mat=None):
    """4x4 in basis |first second> with the first-listed qubit most significant."""
    if kind == "SWAP":
        m = np.eye(4, dtype=complex)[[0, 2, 1, 3]]
        return m
    if kind == "CP":
        return np.diag([1, 1, 1, np.exp(1j * phi)])
    if kind == "RZZ":
        a, b = np.exp(-1j * phi / 2), np.exp(1j * phi / 2)
        return np.diag([a, b, b, a])
    if kind == "U2":
        return np.asarray(mat, dtype=complex).reshape(4, 4)
    raise KeyError(kind)


def full_1q(n, t, u):
    """2^n x 2^n operator of u on qubit t (qubit q = index bit q)."""
    return np.kron(np.kron(np.eye(1 << (n - 1 - t)), u), np.eye(1 << t))


def full_2q(n, q_first, q_second, u4):
    """u4 in basis |first second> (first most significant) on qubits q_first, q_second."""
    d = 1 << n
    m = np.zeros((d, d), dtype=complex)
    for col in range(d):
        bf, bs = (col >> q_first) & 1, (col >> q_second) & 1
        lc = (bf << 1) | bs
        base = col & ~((1 << q_first) | (1 << q_second))
        for lr in range(4):
            row = base | (((lr >> 1) & 1) << q_first) | ((lr & 1) << q_second)
            m[row, col] += u4[lr, lc]
    return m


def textbook_apply(state, n, text):
    """Applies circuit text with dense textbook matrices (n <= 10)."""
    s = np.array(state, dtype=complex)
    for line in text.strip().splitlines():
        tok = line.split()
        k = tok[0]
        if k in ("H", "X", "Y", "Z"):
            s = full_1q(n, int(tok[1]), textbook_1q(k)) @ s
        elif k in ("P", "RX", "RY", "RZ"):
            s = full_1q(n, int(tok[1]), textbook_1q(k, float(tok[2]))) @ s
        elif k == "U1":
            v = [float(x) for x in tok[2:10]]
            s = full_1q(n, int(tok[1]), textbook_1q("U1", mat=[complex(v[2 * i], v[2 * i + 1]) for i in range(4)])) @ s
        elif k in ("SWAP", "VSWAP_1_1"):
            s = full_2q(n, int(tok[1]), int(tok[2]), textbook_2q("SWAP")) @ s
        elif k in ("CP", "RZZ"):
            s = full_2q(n, int(tok[1]), int(tok[2]), textbook_2q(k, float(tok[3]))) @ s
        elif k == "U2":
            v = [float(x) for x in tok[3:35]]
            u = textbook_2q("U2", mat=[complex(v[2 * i], v[2 * i + 1]) for i in range(16)])
            s = full_2q(n, int(tok[1]), int(tok[2]), u) @ s
        elif k == "VSWAP_2_2":
            a, b, c, d = (int(x) for x in tok[1:5])
            sw = textbook_2q("SWAP")
            s = full_2q(n, b, d, sw) @ (full_2q(n, a, c, sw) @ s)
        else:
            raise KeyError(k)
    return s


def qft_closed_form(state, n):
    """U = R F R scaled by (c*sqrt2)^n, F[y,x] = exp(+2 pi i x y / 2^n) / 2^(n/2)."""
    d = 1 << n
    idx = np.arange(d)
    rev = np.zeros(d, dtype=np.int64)
    for q in range(n):
        rev |= ((idx >> q) & 1) << (n - 1 - q)
    x = state[rev]                       # R
    f = np.fft.ifft(x) * math.sqrt(d)    # F: exp(+2 pi i xy/d)/sqrt(d)
    y = f[rev]                           # R
    return y * (H_LITERAL * math.sqrt(2)) ** n


def infidelity(a, b):
    ov = np.vdot(a, b)
    return 1.0 - abs(ov) ** 2 / (np.vdot(a, a).real * np.vdot(b, b).real)


# ---------------------------------------------------------------------------
# measurement sampling: CPU restatement of the rule in kernels/sample.hpp
# ---------------------------------------------------------------------------

MASK64 = (1 << 64) - 1


def splitmix64(seed, count):
    """First `count` outputs of SplitMix64(seed) (Steele, Lea, Flood 2014)."""
    x = seed & MASK64
    out = []
    for _ in range(count):
        x = (x + 0x9E3779B97F4A7C15) & MASK64
        z = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        out.append(z ^ (z >> 31))
    return out


def sample_oracle(state, seed, shots):
    """Shots of the engine's sampler, restated step by step in FP64 with the
    same summation order (sequential sums of 16-amplitude segments, of 64
    segments per block, and over blocks)."""
    st = np.asarray(state)
    n = int(np.log2(st.size))
    re = st.real.astype(np.float64)
    im = st.imag.astype(np.float64)
    p = re * re + im * im
    seg = p.reshape(-1, 16)
    s = np.zeros(seg.shape[0])
    for j in range(16):
        s = s + seg[:, j]
    spb = 64 if n >= 10 else 1 << (n - 4)
    sb = s.reshape(-1, spb)
    B = np.zeros(sb.shape[0])
    for l in range(spb):
        B = B + sb[:, l]
    inc = np.cumsum(B)
    C = np.concatenate([[0.0], inc[:-1]])
    total = inc[-1]
    nz = np.nonzero(B > 0)[0]
    last_nz = int(nz[-1])
    out = []
    for x in splitmix64(seed, shots):
        u = (float(x >> 11) * 2.0 ** -53) * total
        b = int(np.searchsorted(inc, u, side="right"))
        if b >= B.size:
            b = last_nz
        r = u - C[b]
        L, m, Lm, last, Llast = 0.0, -1, 0.0, 0, 0.0
        for l in range(spb):
            sl = sb[b, l]
            if sl > 0.0:
                last, Llast = l, L
            if m < 0 and L + sl > r:
                m, Lm = l, L
            L = L + sl
        if m < 0:
            m, Lm = last, Llast
        t = r - Lm
        a, j, jlast = 0.0, -1, 0
        for i in range(16):
            pi = p[(b * spb + m) * 16 + i]
            if pi > 0.0:
                jlast = i
            a = a + pi
            if j < 0 and a > t:
                j = i
        if j < 0:
            j = jlast
        out.append((b * spb + m) * 16 + j)
    return np.array(out, dtype=np.uint64)


def fs_type(path):
    """Filesystem type of the mount that holds path (from /proc/mounts)."""
    path = os.path.realpath(path)
    best, kind = "", "?"
    try:
        with open("/proc/mounts") as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and (path == parts[1] or path.startswith(parts[1].rstrip("/") + "/")):
                    if len(parts[1]) >= len(best):
                        best, kind = parts[1], parts[2]
    except OSError:
        pass
    return kind


def odirect_ok(directory):
    """True if a 4 KiB O_DIRECT write and read succeed in directory."""
    import mmap
    probe = os.path.join(directory, f".odirect_probe_{os.getpid()}")
    try:
        fd = os.open(probe, os.O_RDWR | os.O_CREAT | os.O_DIRECT, 0o644)
    except OSError:
        return False
    try:
        buf = mmap.mmap(-1, 4096)   # page-aligned
        buf.write(b"\x5a" * 4096)
        if os.pwrite(fd, buf, 0) != 4096:
            return False
        back = mmap.mmap(-1, 4096)
        return os.preadv(fd, [back], 0) == 4096 and back[:] == buf[:]
    except OSError:
        return False
    finally:
        os.close(fd)
        os.unlink(probe)


def odirect_dir(tmp_path):
    """A directory that accepts O_DIRECT: tmp_path itself, else a new temporary
    folder in $SQ_ODIRECT_DIR, the repository root, /var/tmp or /dev/shm (the
    caller removes it).  Returns (path or None, what each candidate's
    filesystem did)."""
    cands = [str(tmp_path)]
    if os.environ.get("SQ_ODIRECT_DIR"):
        cands.append(os.environ["SQ_ODIRECT_DIR"])
    cands += [ROOT, "/var/tmp", "/dev/shm"]
    seen = []
    for c in cands:
        if not os.path.isdir(c):
            seen.append(f"{c} (absent)")
            continue
        if odirect_ok(c):
            if c == str(tmp_path):
                return c, seen
            import tempfile
            return tempfile.mkdtemp(prefix=".sq_odirect_", dir=c), seen
        seen.append(f"{c} ({fs_type(c)}: O_DIRECT refused)")
    return None, seen


def high_qubit_subcircuit(n, chunk, seed, n_blocks=8):
    """A subcircuit file whose gates and VSWAPs act on qubits >= 7 only, so the
    emulation (which does not model cross-lane exchanges) computes every value
    exactly: gate blocks on qubits [7, chunk) between VSWAP_2_2 / VSWAP_1_1
    passes that swap chunk qubits with qubits >= chunk."""
    rng = random.Random(seed)
    nrng = np.random.default_rng(seed)
    blocks = []
    for b in range(n_blocks):
        lines = []
        for _ in range(12):
            k = rng.choice(["H", "U3", "RX", "RZ", "P", "CP", "RZZ", "U2", "SWAP"])
            if k in ("CP", "RZZ", "U2", "SWAP"):
                qs = sorted(rng.sample(range(7, chunk), 2))
            else:
                qs = [rng.randrange(7, chunk)]
            lines.append(gate_line(k, qs, rng, nrng))
        blocks.append(lines)
        if b % 2 == 0:
            lo = sorted(rng.sample(range(7, chunk), 2))
            hi = sorted(rng.sample(range(chunk, n), 2))
            blocks.append([f"VSWAP_2_2 {lo[0]} {lo[1]} {hi[0]} {hi[1]}"])
        else:
            blocks.append([f"VSWAP_1_1 {rng.randrange(7, chunk)} {rng.randrange(chunk, n)}"])
    return "".join(f"{len(bl)}\n" + "\n".join(bl) + "\n" for bl in blocks)
