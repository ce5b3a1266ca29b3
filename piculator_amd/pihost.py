"""ctypes binding of libpiculator.so: the GPU BigFloat and Chudnovsky driver
(piculator_amd/host).  No CPU fallback: every call runs the C++ host code
whose big multiplies go through libpicugpu.so."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import picugpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpiculator.so")

MODE_PLAIN = 0
MODE_BATCHED = 1

OPS = {"mul": 0, "add": 1, "sub": 2, "rcp": 3, "div": 4, "invsqrt": 5, "mul_u32": 6, "ucmp": 7}

_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError("%s is missing: run `make -C piculator_amd`" % LIB_PATH)
        picugpu.load()
        L = C.CDLL(LIB_PATH)
        L.pch_last_error.restype = C.c_char_p
        L.pch_set_engine.argtypes = [C.c_uint32]
        L.pch_set_engine.restype = None
        L.pch_pi_hex.argtypes = [C.c_uint64, C.c_int, C.c_char_p, C.c_uint64,
                                 C.POINTER(C.c_double), C.c_int]
        L.pch_pi_hex.restype = C.c_long
        L.pch_pi_sizes.argtypes = [C.c_uint64] + [C.POINTER(C.c_uint64)] * 3
        L.pch_pi_sizes.restype = None
        L.pch_df_op.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_uint64,
                                C.c_int, C.c_int64, C.c_void_p, C.c_uint64, C.c_uint64,
                                C.POINTER(C.c_int), C.POINTER(C.c_int64), C.c_void_p,
                                C.c_uint64, C.POINTER(C.c_uint64)]
        L.pch_bf_op.argtypes = [C.c_int, C.c_int, C.c_int64, C.c_void_p, C.c_uint64,
                                C.c_int, C.c_int64, C.c_void_p, C.c_uint64, C.c_uint64,
                                C.POINTER(C.c_int), C.POINTER(C.c_int64), C.c_void_p,
                                C.c_uint64, C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def set_engine(flags: int):
    load().pch_set_engine(flags)


PHASES = ["series", "division", "invsqrt", "final_mul", "conversion", "total", "big_muls"]


def pi_sizes(digits: int):
    """(hex digits, precision limbs, series terms) of Pi(digits)."""
    h, p, t = C.c_uint64(), C.c_uint64(), C.c_uint64()
    load().pch_pi_sizes(digits, C.byref(h), C.byref(p), C.byref(t))
    return int(h.value), int(p.value), int(t.value)


def pi_buffer(digits: int):
    """Output buffer large enough for pi.txt of `digits` decimal digits, in
    page-locked host memory (the hex text arrives at full PCIe bandwidth)."""
    return picugpu.host_buffer(int(digits * np.log(10) / np.log(16)) + 128)


def pi_hex_into(digits: int, mode: int, buf, verbose: bool = False):
    """Computes pi.txt into `buf` (from pi_buffer); returns (bytes written,
    phase times dict).  Allocates no result memory on the Python side."""
    picugpu.ensure_init()
    t = (C.c_double * 7)()
    n = load().pch_pi_hex(digits, mode, buf, len(buf), t, 1 if verbose else 0)
    if n < 0:
        raise RuntimeError(load().pch_last_error().decode())
    return n, dict(zip(PHASES, list(t)))


def pi_hex(digits: int, mode: int = MODE_BATCHED, verbose: bool = False):
    """(pi.txt content, phase times dict) computed with GPU products."""
    buf = pi_buffer(digits)
    n, t = pi_hex_into(digits, mode, buf, verbose)
    return buf.raw[:n].decode(), t


def bf_op(op: str, a, b=(True, 0, []), p: int = 0, device: bool = False):
    """One BigFloat operation on (sign, exp, limbs) tuples; returns a tuple.
    device=True runs it on DeviceFloat (values resident in GPU memory);
    op "ucmp" (device only) returns -1, 0 or 1."""
    picugpu.ensure_init()
    sa, ea, la = a
    sb, eb, lb = b
    la = np.ascontiguousarray(np.asarray(la, dtype=np.uint32))
    lb = np.ascontiguousarray(np.asarray(lb, dtype=np.uint32))
    cap = la.size + lb.size + 64 + (p if op in ("rcp", "div", "invsqrt") else 0) * 2
    if op in ("add", "sub"):
        cap += abs(int(ea) - int(eb))  # exponent gap filled by aligned zeros
    out = np.zeros(cap, dtype=np.uint32)
    sz, ez, nz = C.c_int(), C.c_int64(), C.c_uint64()
    fn = load().pch_df_op if device else load().pch_bf_op
    rc = fn(OPS[op], int(sa), ea, la.ctypes.data, la.size, int(sb), eb, lb.ctypes.data, lb.size, p,
            C.byref(sz), C.byref(ez), out.ctypes.data, cap, C.byref(nz))
    if rc != 0:
        raise RuntimeError(load().pch_last_error().decode())
    if op == "ucmp":
        return int(sz.value)
    return (bool(sz.value), int(ez.value), out[: nz.value].tolist())
