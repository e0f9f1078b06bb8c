"""ctypes bindings of include/subquokka_hip.h.

Loading fails loudly when the in-tree library has not been built: there is no
Python or CPU fallback for any engine entry point.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsubquokka_hip.so")

SQ_OK = 0
SQ_ERR_ARG = -1
SQ_ERR_UNKNOWN_GATE = -2
SQ_ERR_QUBIT_INDEX = -3
SQ_ERR_CONTRACT = -4
SQ_ERR_PARSE = -5
SQ_ERR_OOM = -6
SQ_ERR_HIP = -7
SQ_ERR_UNSUPPORTED = -8
SQ_ERR_IO = -9

SQ_C128 = 0
SQ_C64 = 1

GATE_KINDS = ["H", "X", "Y", "Z", "P", "RX", "RY", "RZ", "U1", "U2", "SWAP", "CP", "RZZ",
              "VSWAP_1_1", "VSWAP_2_2"]
KIND = {name: i + 1 for i, name in enumerate(GATE_KINDS)}


class SqGate(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nq", ctypes.c_int32), ("q", ctypes.c_int32 * 4),
                ("param", ctypes.c_double), ("mat", ctypes.c_double * 32)]


class SqFused(ctypes.Structure):
    _fields_ = [("k", ctypes.c_int32), ("q", ctypes.c_int32 * 6), ("u", ctypes.POINTER(ctypes.c_double))]


class SqConfig(ctypes.Structure):
    _fields_ = [("total_qbit", ctypes.c_int32), ("chunk_qbit", ctypes.c_int32), ("dtype", ctypes.c_int32),
                ("device", ctypes.c_int32), ("resident_qbit", ctypes.c_int32), ("sweep_mode", ctypes.c_int32),
                ("pass_qbits", ctypes.c_int32), ("host_tier", ctypes.c_int32)]


class SqZTerm(ctypes.Structure):
    _fields_ = [("coef", ctypes.c_double), ("mask", ctypes.c_uint64)]


class SqKstat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double),
                ("alg_bytes", ctypes.c_double), ("alg_flops", ctypes.c_double)]


class SqError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"sq error {code}: {msg}")
        self.code = code


_lib = None

_P = ctypes.c_void_p
_SIGS = {
    "sq_create": (ctypes.c_int, [ctypes.POINTER(SqConfig), ctypes.POINTER(_P)]),
    "sq_destroy": (ctypes.c_int, [_P]),
    "sq_init_zero": (ctypes.c_int, [_P]),
    "sq_set_state": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_int32]),
    "sq_get_state": (ctypes.c_int, [_P, _P, ctypes.c_uint64, ctypes.c_int32]),
    "sq_device_state": (ctypes.c_int, [_P, ctypes.POINTER(_P), ctypes.POINTER(_P)]),
    "sq_synchronize": (ctypes.c_int, [_P]),
    "sq_last_error": (ctypes.c_char_p, [_P]),
    "sq_apply_gates": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.c_uint64]),
    "sq_apply_gates_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.c_uint64]),
    "sq_apply_blocks": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64]),
    "sq_apply_fused": (ctypes.c_int, [_P, ctypes.POINTER(SqFused), ctypes.c_uint64]),
    "sq_run_fused": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64]),
    "sq_apply_gates_fused": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.c_uint64, ctypes.c_int32]),
    "sq_pass_census": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_int32)]),
    "sq_compose_block": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.c_uint64, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double)]),
    "sq_norm2": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "sq_expect_zz_ring": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "sq_inner": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
    "sq_max_abs_diff": (ctypes.c_int, [_P, _P, ctypes.POINTER(ctypes.c_double)]),
    "sq_hea_energy": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_uint64,
                                     ctypes.POINTER(ctypes.c_double)]),
    "sq_adjoint_grad_hea": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.c_uint64,
                                           ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_double)]),
    "sq_sample": (ctypes.c_int, [_P, ctypes.c_uint64, ctypes.c_uint64, _P]),
    "sq_expect_z": (ctypes.c_int, [_P, ctypes.POINTER(SqZTerm), ctypes.c_uint64, ctypes.POINTER(ctypes.c_double)]),
    "sq_adjoint_grad": (ctypes.c_int, [_P, ctypes.POINTER(SqGate), ctypes.POINTER(ctypes.c_int32), ctypes.c_uint64,
                                       ctypes.POINTER(ctypes.c_double), ctypes.c_uint64, ctypes.POINTER(SqZTerm),
                                       ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]),
    "sq_stats_enable": (ctypes.c_int, [_P, ctypes.c_int32]),
    "sq_stats_get": (ctypes.c_int, [_P, ctypes.POINTER(SqKstat), ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]),
    "sq_stats_reset": (ctypes.c_int, [_P]),
    "sq_stream_info": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "sq_create_file_state": (ctypes.c_int, [ctypes.POINTER(SqConfig), ctypes.POINTER(ctypes.c_char_p), ctypes.c_int32,
                                            ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(_P)]),
    "sq_parse_gate": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.POINTER(SqGate)]),
    "sq_parse_circuit": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(SqGate),
                                        ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                        ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint64,
                                        ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int32)]),
    "sq_gen_circuit": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_double), ctypes.c_char_p, ctypes.c_uint64,
                                      ctypes.POINTER(ctypes.c_uint64)]),
    "sq_gen_state": (ctypes.c_int, [ctypes.c_int32, ctypes.c_uint64, _P]),
    "sq_plan": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_char_p,
                               ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
}


def lib():
    """Loads the engine library (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python __graft_entry__.py build` "
                              "or `python subquokka_dev_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Function names declared in include/subquokka_hip.h."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "subquokka_hip.h")
    text = open(hdr).read()
    return sorted(set(re.findall(r"\b(sq_[a-z0-9_]+)\s*\(", text)))


# ---------------------------------------------------------------------------
# host-side formats
# ---------------------------------------------------------------------------

def _check(rc, eng=None):
    if rc != SQ_OK:
        msg = lib().sq_last_error(eng).decode() if eng else ""
        raise SqError(rc, msg)


def gen_circuit(kind, n, layers=0, seed=0, theta=None):
    """Seeded circuit text: kind 'qft', 'random' or 'hea'."""
    k = {"qft": 0, "random": 1, "hea": 2}[kind]
    th = None
    if theta is not None:
        arr = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if kind != "hea" or arr.size != 2 * n * layers:
            raise ValueError("theta is for the 'hea' circuit and holds 2 * n * layers angles")
        th = arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    need = ctypes.c_uint64(0)
    lib().sq_gen_circuit(k, n, layers, seed, th, None, 0, ctypes.byref(need))
    buf = ctypes.create_string_buffer(need.value)
    _check(lib().sq_gen_circuit(k, n, layers, seed, th, buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


def gen_state(n, seed):
    """Seeded Gaussian normalised state, complex128 numpy array."""
    out = np.empty(1 << n, dtype=np.complex128)
    _check(lib().sq_gen_state(n, seed, out.ctypes.data))
    return out


def parse_circuit(text, n, is_subcircuit=False):
    """Returns (gates array, block sizes or None)."""
    L = lib()
    ng, nb, el = ctypes.c_uint64(0), ctypes.c_uint64(0), ctypes.c_int32(0)
    raw = text.encode() if isinstance(text, str) else text
    rc = L.sq_parse_circuit(raw, n, int(is_subcircuit), None, 0, ctypes.byref(ng), None, 0, ctypes.byref(nb),
                            ctypes.byref(el))
    if rc not in (SQ_OK, SQ_ERR_ARG):
        raise SqError(rc, f"parse error at line {el.value}")
    gates = (SqGate * max(1, ng.value))()
    blocks = (ctypes.c_uint32 * max(1, nb.value))()
    rc = L.sq_parse_circuit(raw, n, int(is_subcircuit), gates, ng.value, ctypes.byref(ng),
                            blocks if is_subcircuit else None, nb.value if is_subcircuit else 0,
                            ctypes.byref(nb), ctypes.byref(el))
    if rc != SQ_OK:
        raise SqError(rc, f"parse error at line {el.value}")
    return gates, (blocks if is_subcircuit else None), ng.value, nb.value


def zterms(terms):
    """Z-string observable: [(coef, qubits), ...] -> ctypes array of sq_zterm
    (h(i) = sum coef * (-1)^(parity of i's bits on qubits))."""
    arr = (SqZTerm * max(1, len(terms)))()
    for k, (coef, qubits) in enumerate(terms):
        mask = 0
        for q in qubits:
            mask |= 1 << int(q)
        arr[k].coef = float(coef)
        arr[k].mask = mask
    return arr, len(terms)


def plan(text, n, mode, param):
    """Planner output text (mode 0: subcircuit/VSWAP plan, 1: fused blocks)."""
    L = lib()
    need = ctypes.c_uint64(0)
    raw = text.encode()
    rc = L.sq_plan(raw, n, mode, param, None, 0, ctypes.byref(need))
    if rc not in (SQ_OK, SQ_ERR_ARG):
        raise SqError(rc, "plan failed")
    buf = ctypes.create_string_buffer(need.value)
    _check(L.sq_plan(raw, n, mode, param, buf, need.value, ctypes.byref(need)))
    return buf.value.decode()


# ---------------------------------------------------------------------------
# engine
# ---------------------------------------------------------------------------

class Engine:
    """One state on one GPU (owns device memory until close())."""

    def __init__(self, n, dtype="c128", chunk=0, device=0, resident_qbit=0, sweep_mode=0, pass_qbits=0,
                 host_tier="auto", state_files=None, direct_io=False, keep_files=False):
        """sweep_mode 0: plain sweeps run as register passes (pass_qbits register
        qubits, 0 = default); 1: one launch per gate.  resident_qbit > 0: the
        out-of-core runner, with the state in host memory (host_tier "auto" or
        "pageable") or, with state_files (a list of 2^f paths), in files."""
        cfg = SqConfig(total_qbit=n, chunk_qbit=chunk, dtype=SQ_C128 if dtype == "c128" else SQ_C64,
                       device=device, resident_qbit=resident_qbit, sweep_mode=sweep_mode, pass_qbits=pass_qbits,
                       host_tier={"auto": 0, "pageable": 1}[host_tier])
        h = ctypes.c_void_p()
        if state_files:
            paths = (ctypes.c_char_p * len(state_files))(*[os.fsencode(p) for p in state_files])
            rc = lib().sq_create_file_state(ctypes.byref(cfg), paths, len(state_files), int(direct_io),
                                            int(keep_files), ctypes.byref(h))
        else:
            rc = lib().sq_create(ctypes.byref(cfg), ctypes.byref(h))
        if rc != SQ_OK:
            raise SqError(rc, "sq_create failed")
        self.h = h
        self.n = n
        self.dtype = dtype

    def close(self):
        if self.h:
            lib().sq_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        if rc != SQ_OK:
            raise SqError(rc, lib().sq_last_error(self.h).decode())

    def init_zero(self):
        self._c(lib().sq_init_zero(self.h))

    def set_state(self, arr):
        arr = np.ascontiguousarray(arr)
        dt = SQ_C128 if arr.dtype == np.complex128 else SQ_C64
        if arr.dtype not in (np.complex128, np.complex64):
            raise TypeError("state must be complex64 or complex128")
        self._c(lib().sq_set_state(self.h, arr.ctypes.data, arr.size, dt))

    def get_state(self, dtype=np.complex128):
        out = np.empty(1 << self.n, dtype=dtype)
        self._c(lib().sq_get_state(self.h, out.ctypes.data, out.size, SQ_C128 if dtype == np.complex128 else SQ_C64))
        return out

    def apply_gates(self, gates, n=None):
        n = len(gates) if n is None else n
        self._c(lib().sq_apply_gates(self.h, gates, n))

    def apply_gates_noalloc(self, gates, n=None):
        """sq_apply_gates_noalloc: the same sweep, no device allocation."""
        n = len(gates) if n is None else n
        self._c(lib().sq_apply_gates_noalloc(self.h, gates, n))

    def apply_text(self, text):
        gates, _, ng, _ = parse_circuit(text, self.n)
        self.apply_gates(gates, ng)

    def apply_blocks(self, gates, blocks, n_blocks):
        self._c(lib().sq_apply_blocks(self.h, gates, blocks, n_blocks))

    def apply_fused(self, blocks):
        arr = (SqFused * len(blocks))()
        keep = []
        for i, (qs, u) in enumerate(blocks):
            u = np.ascontiguousarray(u, dtype=np.complex128)
            keep.append(u)
            arr[i].k = len(qs)
            for j, q in enumerate(qs):
                arr[i].q[j] = q
            arr[i].u = u.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._c(lib().sq_apply_fused(self.h, arr, len(blocks)))

    def run_fused(self, gates, blocks, n_blocks):
        """Planner blocks (<= 5 qubits each): FP64 composition + MFMA application."""
        self._c(lib().sq_run_fused(self.h, gates, blocks, n_blocks))

    def apply_gates_fused(self, gates, n, tile_qbit=0):
        """A plain gate list as fused passes (complex64 engines): four-qubit
        blocks composed in FP64, applied on MFMA inside LDS tiles of
        tile_qbit qubits (0: the engine's default, 13)."""
        self._c(lib().sq_apply_gates_fused(self.h, gates, n, tile_qbit))

    def pass_census(self):
        """Op counts of the last register-pass program by form (sq_pass_census):
        a dict with 'dense_reg', 'dense_lane03', 'dense_lane45', 'dense_vec',
        'diag_free', 'diag_reg1', 'diag_reg2', 'phases' (products of one-qubit
        phases left by dense gates split into phase, real rotation, phase) and
        'fold' -- a 3x4 array of folded controlled phases by target (register,
        lane 0-3, lane 4-5) and control (register slot 0, other slot, lane,
        tile)."""
        c = (ctypes.c_int64 * 20)()
        n = ctypes.c_int32()
        self._c(lib().sq_pass_census(self.h, c, 20, ctypes.byref(n)))
        v = list(c)
        names = ["dense_reg", "dense_lane03", "dense_lane45", "dense_vec", "diag_free", "diag_reg1", "diag_reg2",
                 "phases"]
        out = {k: v[i] for i, k in enumerate(names)}
        out["fold"] = np.array(v[8:20], dtype=np.int64).reshape(3, 4)
        return out

    def compose(self, gates, n, qubits):
        k = len(qubits)
        q = (ctypes.c_int32 * 6)(*qubits)
        out = np.empty((1 << k, 1 << k), dtype=np.complex128)
        self._c(lib().sq_compose_block(self.h, gates, n, k, q, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def sync(self):
        self._c(lib().sq_synchronize(self.h))

    def norm2(self):
        v = ctypes.c_double()
        self._c(lib().sq_norm2(self.h, ctypes.byref(v)))
        return v.value

    def expect_zz_ring(self):
        v = ctypes.c_double()
        self._c(lib().sq_expect_zz_ring(self.h, ctypes.byref(v)))
        return v.value

    def inner(self, other):
        re, im = ctypes.c_double(), ctypes.c_double()
        self._c(lib().sq_inner(self.h, other.h, ctypes.byref(re), ctypes.byref(im)))
        return complex(re.value, im.value)

    def max_abs_diff(self, other):
        v = ctypes.c_double()
        self._c(lib().sq_max_abs_diff(self.h, other.h, ctypes.byref(v)))
        return v.value

    def device_state(self):
        p, s = ctypes.c_void_p(), ctypes.c_void_p()
        self._c(lib().sq_device_state(self.h, ctypes.byref(p), ctypes.byref(s)))
        return p.value, s.value

    def stats_enable(self, on=True):
        self._c(lib().sq_stats_enable(self.h, int(on)))

    def stats_reset(self):
        self._c(lib().sq_stats_reset(self.h))

    def sample(self, shots, seed):
        """Seeded computational-basis shots (uint64 indices)."""
        out = np.empty(max(1, shots), dtype=np.uint64)
        self._c(lib().sq_sample(self.h, seed, shots, out.ctypes.data))
        return out[:shots]

    def stream_info(self):
        """(passes, host-link bytes, host tier name) of the out-of-core runner
        since stats_reset."""
        p, b, t = ctypes.c_uint64(), ctypes.c_double(), ctypes.c_int32()
        self._c(lib().sq_stream_info(self.h, ctypes.byref(p), ctypes.byref(b), ctypes.byref(t)))
        return p.value, b.value, {0: "hbm", 1: "pageable", 2: "pinned", 3: "files"}.get(t.value, "?")

    def stats(self):
        arr = (SqKstat * 64)()
        n = ctypes.c_int32()
        self._c(lib().sq_stats_get(self.h, arr, 64, ctypes.byref(n)))
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms,
                                           alg_bytes=arr[i].alg_bytes, alg_flops=arr[i].alg_flops)
                for i in range(n.value)}

    def expect_z(self, terms):
        """<psi|H|psi> for H = sum coef * prod Z over qubits (terms as for zterms)."""
        arr, nt = zterms(terms)
        v = ctypes.c_double()
        self._c(lib().sq_expect_z(self.h, arr, nt, ctypes.byref(v)))
        return v.value

    def adjoint_grad(self, text, param, theta, terms, checkpoint_every=0):
        """Energy and gradient of a parameterised circuit (sq_adjoint_grad) from
        the current state.  text: circuit lines; param[i] = index into theta of
        gate i's angle, or -1 for a fixed gate."""
        gates, _, ng, _ = parse_circuit(text, self.n)
        pa = np.ascontiguousarray(param, dtype=np.int32).reshape(-1)
        if pa.size != ng:
            raise ValueError(f"param has {pa.size} entries for {ng} gates")
        th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if pa.size and pa.max() >= th.size:
            raise ValueError("a parameter index lies beyond theta")
        arr, nt = zterms(terms)
        g = np.zeros(max(1, th.size))
        v = ctypes.c_double()
        self._c(lib().sq_adjoint_grad(self.h, gates, pa.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), ng,
                                      th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), th.size, arr, nt,
                                      checkpoint_every, ctypes.byref(v),
                                      g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return v.value, g[:th.size]

    def _hea_theta(self, layers, theta):
        th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if th.size != 2 * self.n * layers:
            raise ValueError(f"theta holds {th.size} angles; the ansatz has 2 * {self.n} * {layers}")
        return th

    def hea_energy(self, layers, theta):
        th = self._hea_theta(layers, theta)
        v = ctypes.c_double()
        self._c(lib().sq_hea_energy(self.h, layers, th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), th.size,
                                    ctypes.byref(v)))
        return v.value

    def adjoint_grad_hea(self, layers, theta, checkpoint_every=0):
        th = self._hea_theta(layers, theta)
        g = np.empty_like(th)
        v = ctypes.c_double()
        self._c(lib().sq_adjoint_grad_hea(self.h, layers, th.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          th.size, checkpoint_every, ctypes.byref(v),
                                          g.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return v.value, g
