"""ctypes front end of the oracle (oracle/orc.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``.  The product (``piculator_amd``)
never imports it.

The oracle restates the reference's CPU arithmetic: the 63-bit-prime NTT
(NTT.cpp), the 8-bit-digit FP64 FFT (FFT.cpp), BigFloat with its truncation
rules (BigFloat.cpp) and the Chudnovsky driver (PI.h).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

ENGINE_REFERENCE = 0  # FFT below k = 27, NTT above (BigFloat.cpp:386)
ENGINE_NTT = 1        # every product through the NTT
ENGINE_FFT = 2        # every product through the FFT


class OrcBF(C.Structure):
    _fields_ = [("sign", C.c_int32), ("exp", C.c_int64), ("L", C.c_uint64),
                ("T", C.POINTER(C.c_uint32))]


def build() -> None:
    """Compile liborc.so from orc.c (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        bfp = C.POINTER(OrcBF)
        sz = C.c_size_t
        L.orc_ntt_prime.argtypes = [C.c_int]; L.orc_ntt_prime.restype = C.c_uint64
        L.orc_ntt_generator.argtypes = [C.c_uint64]; L.orc_ntt_generator.restype = C.c_uint64
        L.orc_modmul.argtypes = [C.c_uint64] * 3; L.orc_modmul.restype = C.c_uint64
        L.orc_modpow.argtypes = [C.c_uint64] * 3; L.orc_modpow.restype = C.c_uint64
        L.orc_ntt_roots.argtypes = [C.c_int] + [C.POINTER(C.c_uint64)] * 4
        L.orc_int_to_ntt.argtypes = [u64p, C.c_int, u32p, sz]
        L.orc_ntt_forward.argtypes = [u64p, C.c_int, C.c_int]
        L.orc_ntt_inverse.argtypes = [u64p, C.c_int, C.c_int]
        L.orc_ntt_pointwise.argtypes = [u64p, u64p, C.c_int]
        L.orc_ntt_to_int.argtypes = [u64p, C.c_int, u32p, sz]; L.orc_ntt_to_int.restype = C.c_int
        L.orc_fft_ensure_table.argtypes = [C.c_int]
        L.orc_int_to_fft.argtypes = [f64p, C.c_int, u32p, sz]
        L.orc_fft_forward.argtypes = [f64p, C.c_int, C.c_int]
        L.orc_fft_inverse.argtypes = [f64p, C.c_int, C.c_int]
        L.orc_fft_pointwise.argtypes = [f64p, f64p, C.c_int]
        L.orc_fft_to_int.argtypes = [f64p, C.c_int, u32p, sz]; L.orc_fft_to_int.restype = C.c_int
        L.orc_mul_limbs.argtypes = [u32p, sz, u32p, sz, u32p, C.c_int, C.c_int]
        L.orc_mul_limbs.restype = C.c_int
        L.orc_bf_free.argtypes = [bfp]
        L.orc_bf_from_u32.argtypes = [bfp, C.c_uint32, C.c_int]
        L.orc_bf_ucmp.argtypes = [bfp, bfp]; L.orc_bf_ucmp.restype = C.c_int
        L.orc_bf_mul_u32.argtypes = [bfp, bfp, C.c_uint32]
        L.orc_bf_add.argtypes = [bfp, bfp, bfp, sz]
        L.orc_bf_sub.argtypes = [bfp, bfp, bfp, sz]
        L.orc_bf_mul.argtypes = [bfp, bfp, bfp, sz, C.c_int, C.c_int]
        L.orc_bf_rcp.argtypes = [bfp, bfp, sz, C.c_int, C.c_int]; L.orc_bf_rcp.restype = C.c_int
        L.orc_bf_div.argtypes = [bfp, bfp, bfp, sz, C.c_int, C.c_int]; L.orc_bf_div.restype = C.c_int
        L.orc_bf_invsqrt.argtypes = [bfp, C.c_uint32, sz, C.c_int, C.c_int]
        L.orc_bf_invsqrt.restype = C.c_int
        L.orc_bf_to_hex.argtypes = [bfp, sz, C.c_char_p, sz]; L.orc_bf_to_hex.restype = C.c_long
        L.orc_pi_sizes.argtypes = [sz, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]
        L.orc_pi_hex.argtypes = [sz, C.c_int, C.c_int, C.c_char_p, sz, C.POINTER(C.c_double)]
        L.orc_pi_hex.restype = C.c_long
        L.orc_mul_count.restype = C.c_uint64
        L.orc_pi_bsr.argtypes = [bfp, bfp, bfp, C.c_uint32, C.c_uint32, sz, C.c_int, C.c_int]
        L.orc_set_plant.argtypes = [C.c_int]
        L.orc_set_plant.restype = None
        L.orc_bbp_hex.argtypes = [C.c_uint64, C.c_int]
        L.orc_bbp_hex.restype = C.c_uint32
        _lib = L
    return _lib


def _u32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


# --------------------------------------------------------------- NTT engine
def ntt_prime(k: int) -> int:
    return int(lib().orc_ntt_prime(k))


def ntt_generator(p: int) -> int:
    return int(lib().orc_ntt_generator(p))


def ntt_roots(k: int):
    """(p, W, iW, rN) of NNT::init for N = 2^k (NTT.cpp:169-188)."""
    v = [C.c_uint64() for _ in range(4)]
    lib().orc_ntt_roots(k, *[C.byref(x) for x in v])
    return tuple(int(x.value) for x in v)


def modmul(a: int, b: int, p: int) -> int:
    return int(lib().orc_modmul(a, b, p))


def int_to_ntt(limbs, k: int) -> np.ndarray:
    a = _u32(limbs)
    T = np.zeros(1 << k, dtype=np.uint64)
    lib().orc_int_to_ntt(T, k, a, a.size)
    return T


def ntt_forward(T: np.ndarray, k: int, threads: int = 1) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.uint64).copy()
    lib().orc_ntt_forward(T, k, threads)
    return T


def ntt_inverse(T: np.ndarray, k: int, threads: int = 1) -> np.ndarray:
    T = np.ascontiguousarray(T, dtype=np.uint64).copy()
    lib().orc_ntt_inverse(T, k, threads)
    return T


def ntt_to_int(T: np.ndarray, k: int, AL: int) -> np.ndarray:
    out = np.zeros(AL, dtype=np.uint32)
    rc = lib().orc_ntt_to_int(np.ascontiguousarray(T, dtype=np.uint64), k, out, AL)
    if rc:
        raise ValueError("FFT length is too small.")
    return out


# --------------------------------------------------------------- FFT engine
def fft_forward(z: np.ndarray, k: int, threads: int = 1) -> np.ndarray:
    lib().orc_fft_ensure_table(k)
    T = np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).view(np.float64).copy()
    lib().orc_fft_forward(T, k, threads)
    return T.view(np.complex128)


def fft_inverse(z: np.ndarray, k: int, threads: int = 1) -> np.ndarray:
    lib().orc_fft_ensure_table(k)
    T = np.ascontiguousarray(np.asarray(z, dtype=np.complex128)).view(np.float64).copy()
    lib().orc_fft_inverse(T, k, threads)
    return T.view(np.complex128)


# ---------------------------------------------------------------- products
def mul_limbs(a, b, engine: int = ENGINE_NTT, threads: int = 1) -> np.ndarray:
    """Untrimmed na+nb limb product through the chosen engine."""
    a, b = _u32(a), _u32(b)
    out = np.zeros(a.size + b.size, dtype=np.uint32)
    rc = lib().orc_mul_limbs(a, a.size, b, b.size, out, engine, threads)
    if rc:
        raise RuntimeError("oracle multiply failed (%d)" % rc)
    return out


def mul_count() -> int:
    return int(lib().orc_mul_count())


# ---------------------------------------------------------------- BigFloat
class BF:
    """Owned oracle BigFloat: (sign, exp, limbs) with the reference's meaning."""

    def __init__(self, sign: bool = True, exp: int = 0, limbs=()):
        self._s = OrcBF(1, 0, 0, None)
        limbs = _u32(limbs)
        if limbs.size:
            buf = (C.c_uint32 * limbs.size)(*limbs.tolist())
            # copy into malloc'd memory owned by the C side
            libc = C.CDLL(None)
            libc.malloc.restype = C.c_void_p
            libc.malloc.argtypes = [C.c_size_t]
            ptr = libc.malloc(4 * limbs.size)
            C.memmove(ptr, buf, 4 * limbs.size)
            self._s = OrcBF(1 if sign else 0, exp, limbs.size, C.cast(ptr, C.POINTER(C.c_uint32)))
        else:
            self._s = OrcBF(1, 0, 0, None)

    def __del__(self):
        try:
            lib().orc_bf_free(C.byref(self._s))
        except Exception:
            pass

    @property
    def ref(self):
        return C.byref(self._s)

    @property
    def sign(self) -> bool:
        return bool(self._s.sign)

    @property
    def exp(self) -> int:
        return int(self._s.exp)

    @property
    def limbs(self) -> np.ndarray:
        n = int(self._s.L)
        if n == 0:
            return np.zeros(0, dtype=np.uint32)
        return np.ctypeslib.as_array(self._s.T, shape=(n,)).copy()

    def tuple(self):
        return (self.sign, self.exp, self.limbs.tolist())

    @staticmethod
    def from_u32(x: int, sign: bool = True) -> "BF":
        z = BF()
        lib().orc_bf_from_u32(z.ref, x, 1 if sign else 0)
        return z

    def mul(self, other: "BF", p: int = 0, engine: int = ENGINE_NTT, threads: int = 1) -> "BF":
        z = BF()
        lib().orc_bf_mul(z.ref, self.ref, other.ref, p, engine, threads)
        return z

    def mul_u32(self, x: int) -> "BF":
        z = BF()
        lib().orc_bf_mul_u32(z.ref, self.ref, x)
        return z

    def ucmp(self, other: "BF") -> int:
        """BigFloat::ucmp (BigFloat.cpp:123-147)."""
        return int(lib().orc_bf_ucmp(self.ref, other.ref))

    def add(self, other: "BF", p: int = 0) -> "BF":
        z = BF()
        lib().orc_bf_add(z.ref, self.ref, other.ref, p)
        return z

    def sub(self, other: "BF", p: int = 0) -> "BF":
        z = BF()
        lib().orc_bf_sub(z.ref, self.ref, other.ref, p)
        return z

    def rcp(self, p: int, engine: int = ENGINE_NTT, threads: int = 1) -> "BF":
        z = BF()
        if lib().orc_bf_rcp(z.ref, self.ref, p, engine, threads):
            raise ZeroDivisionError("Divide by Zero")
        return z

    def div(self, other: "BF", p: int, engine: int = ENGINE_NTT, threads: int = 1) -> "BF":
        z = BF()
        if lib().orc_bf_div(z.ref, self.ref, other.ref, p, engine, threads):
            raise ZeroDivisionError("Divide by Zero")
        return z

    def to_hex(self, digits: int) -> str:
        buf = C.create_string_buffer(digits + 64)
        n = lib().orc_bf_to_hex(self.ref, digits, buf, len(buf))
        if n < 0:
            raise ValueError("hex conversion out of range")
        return buf.value.decode()


def set_plant(k: int) -> None:
    """Plant a mistake in the Newton steps (tests' self-checks only):
    1 = the last sub keeps p + 1 limbs, 2 = the schedule s = p / 2; 0 = none."""
    lib().orc_set_plant(k)


def invsqrt(x: int, p: int, engine: int = ENGINE_NTT, threads: int = 1) -> BF:
    z = BF()
    if lib().orc_bf_invsqrt(z.ref, x, p, engine, threads):
        raise ZeroDivisionError("Divide by Zero")
    return z


# -------------------------------------------------------------------- pi
def pi_sizes(decimal_digits: int):
    h, p, t = C.c_size_t(), C.c_size_t(), C.c_size_t()
    lib().orc_pi_sizes(decimal_digits, C.byref(h), C.byref(p), C.byref(t))
    return int(h.value), int(p.value), int(t.value)


def pi_bsr(a: int, b: int, p: int, engine: int = ENGINE_NTT, threads: int = 1):
    """(P, Q, R) of the binary-splitting recursion over [a, b) (PI.h:64-112)."""
    P, Q, R = BF(), BF(), BF()
    lib().orc_pi_bsr(P.ref, Q.ref, R.ref, a, b, p, engine, threads)
    return P, Q, R


def pi_hex(decimal_digits: int, engine: int = ENGINE_REFERENCE, threads: int = 1):
    """Content of pi.txt for Pi(decimal_digits) and the five phase times."""
    hexd, _, _ = pi_sizes(decimal_digits)
    buf = C.create_string_buffer(hexd + 64)
    times = (C.c_double * 5)()
    n = lib().orc_pi_hex(decimal_digits, engine, threads, buf, len(buf), times)
    if n < 0:
        raise RuntimeError("oracle Pi failed")
    return buf.raw[:n].decode(), list(times)


def bbp_hex(n: int, threads: int = 1, digits: int = 6) -> str:
    """Hex digits of pi from digit n + 1 after the point (BBP extraction)."""
    v = int(lib().orc_bbp_hex(n, threads))
    return ("%08x" % v)[:digits]
