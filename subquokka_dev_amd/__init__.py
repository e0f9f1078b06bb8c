"""MI355X state-vector engine for SubQuokka's gate-application path.

The engine is native (HIP kernels for gfx950 behind the C ABI in
include/subquokka_hip.h, plus a C++ host with the reference's INI/circuit
formats and runner selection).  This package only loads that library.
"""
from .native import (Engine, SqError, gen_circuit, gen_state, parse_circuit, plan, lib,  # noqa: F401
                     declared_symbols, zterms)
