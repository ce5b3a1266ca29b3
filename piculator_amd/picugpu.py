"""ctypes binding of the picugpu C ABI (include/picugpu.h).

Tests and bench.py call the GPU engine through this module, i.e. through
exactly the C entry points a C/C++ caller (piculator_amd/host/BigFloat.cpp)
uses.  There is no fallback: if libpicugpu.so is missing or the GPU is not
usable, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpicugpu.so")

ENGINE_AUTO = 0
ENGINE_NTT = 1
ENGINE_FFT = 2
SYNC = 0x100
BATCH_PLAIN = 0x200
PROFILE = 0x400

STATUS = {
    0: "PCG_OK", 1: "PCG_ERR_INVALID", 2: "PCG_ERR_NOT_INIT", 3: "PCG_ERR_NO_DEVICE",
    4: "PCG_ERR_HIP", 5: "PCG_ERR_TOO_LARGE", 6: "PCG_ERR_OOM", 7: "PCG_ERR_INTERNAL",
}

# Every symbol include/picugpu.h declares.
EXPORTS = [
    "pcg_init", "pcg_shutdown", "pcg_last_error_string", "pcg_reserve", "pcg_mul",
    "pcg_mul_dev", "pcg_mul_dev_noalloc", "pcg_mul_batch", "pcg_mul_batch_dev", "pcg_sync", "pcg_dev_alloc",
    "pcg_dev_free", "pcg_host_alloc", "pcg_host_free", "pcg_memcpy_h2d", "pcg_memcpy_d2h",
    "pcg_fft_last_roundoff",
    "pcg_last_profile", "pcg_ntt_plan", "pcg_timer_start", "pcg_timer_stop",
    "pcg_bsr_exact", "pcg_bsr_bound", "pcg_bsr_prepare", "pcg_bsr_graphs", "pcg_df_free", "pcg_df_set",
    "pcg_df_read", "pcg_df_copy", "pcg_df_ucmp", "pcg_df_add", "pcg_df_sub", "pcg_df_mul",
    "pcg_df_mul_pair", "pcg_df_mul_u32", "pcg_df_to_hex", "pcg_df_bsr_exact",
    "pcg_df_release_pool", "pcg_autotune", "pcg_batch_table", "pcg_bsr_plan_check",
]


class PcgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


class BsrValue(C.Structure):
    _fields_ = [("limbs", C.c_void_p), ("cap", C.c_uint64), ("L", C.c_uint64), ("sign", C.c_int32)]


class DFloat(C.Structure):
    """pcg_dfloat: a BigFloat whose limbs live in the library's device pool."""
    _fields_ = [("limbs", C.c_void_p), ("cap", C.c_uint64), ("L", C.c_uint64), ("exp", C.c_int64),
                ("sign", C.c_int32)]


class MulDesc(C.Structure):
    _fields_ = [("a", C.c_void_p), ("na", C.c_uint64), ("b", C.c_void_p), ("nb", C.c_uint64),
                ("out", C.c_void_p)]


_lib = None


def load():
    """Load libpicugpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            "%s is missing: build it with `make -C piculator_amd/csrc` "
            "(or __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    u64 = C.c_uint64
    vp = C.c_void_p
    L.pcg_init.argtypes = [C.c_int]
    L.pcg_shutdown.argtypes = []
    L.pcg_shutdown.restype = None
    L.pcg_last_error_string.restype = C.c_char_p
    L.pcg_reserve.argtypes = [u64]
    L.pcg_mul.argtypes = [vp, u64, vp, u64, vp, C.c_uint32]
    L.pcg_mul_dev.argtypes = [vp, u64, vp, u64, vp, C.c_uint32]
    L.pcg_mul_dev_noalloc.argtypes = [vp, u64, vp, u64, vp, C.c_uint32]
    L.pcg_mul_batch.argtypes = [C.POINTER(MulDesc), u64, C.c_uint32]
    L.pcg_mul_batch_dev.argtypes = [C.POINTER(MulDesc), u64, C.c_uint32]
    L.pcg_dev_alloc.argtypes = [C.POINTER(vp), u64]
    L.pcg_dev_free.argtypes = [vp]
    L.pcg_host_alloc.argtypes = [C.POINTER(vp), u64]
    L.pcg_host_free.argtypes = [vp]
    L.pcg_memcpy_h2d.argtypes = [vp, vp, u64]
    L.pcg_memcpy_d2h.argtypes = [vp, vp, u64]
    L.pcg_fft_last_roundoff.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.pcg_last_profile.argtypes = [C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    L.pcg_ntt_plan.argtypes = [u64, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pcg_timer_stop.argtypes = [C.POINTER(C.c_double)]
    L.pcg_bsr_exact.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64] + [C.POINTER(BsrValue)] * 3
    L.pcg_bsr_bound.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]
    L.pcg_bsr_prepare.argtypes = [C.POINTER(C.c_uint32), C.c_uint64]
    L.pcg_bsr_graphs.argtypes = [C.c_int]
    L.pcg_bsr_plan_check.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_int), C.c_int]
    L.pcg_autotune.argtypes = [C.POINTER(C.c_int), C.c_int]
    L.pcg_batch_table.argtypes = [C.POINTER(C.c_int), C.c_int, C.c_int]
    dfp = C.POINTER(DFloat)
    L.pcg_df_free.argtypes = [dfp]
    L.pcg_df_set.argtypes = [dfp, C.c_void_p, C.c_uint64, C.c_int64, C.c_int32]
    L.pcg_df_read.argtypes = [dfp, C.c_uint64, C.c_uint64, C.c_void_p]
    L.pcg_df_copy.argtypes = [dfp, dfp]
    L.pcg_df_ucmp.argtypes = [dfp, dfp, C.POINTER(C.c_int)]
    L.pcg_df_add.argtypes = [dfp, dfp, dfp, C.c_uint64]
    L.pcg_df_sub.argtypes = [dfp, dfp, dfp, C.c_uint64]
    L.pcg_df_mul.argtypes = [dfp, dfp, dfp, C.c_uint64, C.c_uint32]
    L.pcg_df_mul_pair.argtypes = [dfp, dfp, dfp, dfp, dfp, C.c_uint64, C.c_uint32]
    L.pcg_df_mul_u32.argtypes = [dfp, dfp, C.c_uint32]
    L.pcg_df_to_hex.argtypes = [dfp, C.c_uint64, C.c_char_p, C.c_uint64, C.POINTER(C.c_uint64)]
    L.pcg_df_bsr_exact.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, dfp, dfp, dfp]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise PcgError(rc, load().pcg_last_error_string().decode())


def _u32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


_inited = False


def init(device: int = 0):
    global _inited
    _check(load().pcg_init(device))
    _inited = True


def shutdown():
    global _inited
    load().pcg_shutdown()
    _inited = False


def ensure_init():
    if not _inited:
        init(0)


def mul(a, b, engine: int = ENGINE_AUTO, profile: bool = False) -> np.ndarray:
    """Exact product of two limb vectors (host memory), na+nb limbs untrimmed."""
    ensure_init()
    a, b = _u32(a), _u32(b)
    out = np.empty(a.size + b.size, dtype=np.uint32)
    flags = engine | (PROFILE if profile else 0)
    _check(load().pcg_mul(_ptr(a), a.size, _ptr(b), b.size, _ptr(out), flags))
    return out


def mul_batch(pairs, engine: int = ENGINE_AUTO, plain: bool = False):
    """Products of a list of (a, b) limb vectors through pcg_mul_batch."""
    ensure_init()
    keep = []
    descs = (MulDesc * max(1, len(pairs)))()
    outs = []
    for i, (a, b) in enumerate(pairs):
        a, b = _u32(a), _u32(b)
        o = np.empty(a.size + b.size, dtype=np.uint32)
        keep += [a, b]
        outs.append(o)
        descs[i] = MulDesc(a.ctypes.data, a.size, b.ctypes.data, b.size, o.ctypes.data)
    flags = engine | (BATCH_PLAIN if plain else 0)
    _check(load().pcg_mul_batch(descs, len(pairs), flags))
    return outs


class DeviceBuffer:
    """Device memory owned through pcg_dev_alloc / pcg_dev_free."""

    def __init__(self, nbytes: int):
        ensure_init()
        p = C.c_void_p()
        _check(load().pcg_dev_alloc(C.byref(p), nbytes))
        self.ptr = p.value
        self.nbytes = nbytes

    def upload(self, arr: np.ndarray):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        _check(load().pcg_memcpy_h2d(self.ptr, _ptr(arr), arr.nbytes))

    def download(self, arr: np.ndarray):
        assert arr.nbytes <= self.nbytes and arr.flags["C_CONTIGUOUS"]
        _check(load().pcg_memcpy_d2h(_ptr(arr), self.ptr, arr.nbytes))
        return arr

    def free(self):
        if self.ptr:
            load().pcg_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def host_buffer(nbytes: int):
    """A ctypes char array of nbytes in page-locked host memory (pcg_host_alloc),
    freed when the array is garbage-collected."""
    import weakref

    ensure_init()
    p = C.c_void_p()
    _check(load().pcg_host_alloc(C.byref(p), nbytes))
    arr = (C.c_char * nbytes).from_address(p.value)
    weakref.finalize(arr, load().pcg_host_free, p.value)
    return arr


def mul_dev(d_a: int, na: int, d_b: int, nb: int, d_out: int, engine: int = ENGINE_AUTO,
            sync: bool = False, profile: bool = False):
    ensure_init()
    flags = engine | (SYNC if sync else 0) | (PROFILE if profile else 0)
    _check(load().pcg_mul_dev(d_a, na, d_b, nb, d_out, flags))


NOALLOC_MAX_LIMBS = 1 << 20


def mul_dev_noalloc(d_a: int, na: int, d_b: int, nb: int, d_out: int, engine: int = ENGINE_AUTO,
                    sync: bool = False):
    """pcg_mul_dev_noalloc: the product with the library's static scratch
    array (na + nb <= NOALLOC_MAX_LIMBS); allocates nothing."""
    ensure_init()
    flags = engine | (SYNC if sync else 0)
    _check(load().pcg_mul_dev_noalloc(d_a, na, d_b, nb, d_out, flags))


def sync():
    _check(load().pcg_sync())


def last_profile():
    st = (C.c_double * 3)()
    tot = C.c_double()
    _check(load().pcg_last_profile(st, 3, C.byref(tot)))
    return list(st), tot.value


def fft_last_roundoff():
    d = C.c_double()
    r = C.c_int()
    _check(load().pcg_fft_last_roundoff(C.byref(d), C.byref(r)))
    return d.value, bool(r.value)


def ntt_plan(total_limbs: int):
    lg = C.c_int()
    ps = C.c_int()
    _check(load().pcg_ntt_plan(total_limbs, C.byref(lg), C.byref(ps)))
    return lg.value, ps.value


def reserve(total_limbs: int):
    ensure_init()
    _check(load().pcg_reserve(total_limbs))


def timer_start():
    _check(load().pcg_timer_start())


def timer_stop() -> float:
    ms = C.c_double()
    _check(load().pcg_timer_stop(C.byref(ms)))
    return ms.value


def bsr_bound(a: int, b: int) -> int:
    v = C.c_uint64()
    _check(load().pcg_bsr_bound(a, b, C.byref(v)))
    return v.value


def bsr_prepare(sizes):
    """Builds the plans of subtrees of these sizes (the first two distinct)."""
    ensure_init()
    arr = (C.c_uint32 * max(1, len(sizes)))(*sizes)
    _check(load().pcg_bsr_prepare(arr, len(sizes)))


def bsr_graphs(enable: bool) -> None:
    """Subtree schedules as replayed HIP graphs (True) or launched kernel by
    kernel (False, the default)."""
    ensure_init()
    _check(load().pcg_bsr_graphs(1 if enable else 0))


def bsr_plan_check(leaves: int, table, bmax: int = 0) -> None:
    """Host-only self-check of the subtree plan layout (no GPU needed) for
    leaves up to index bmax (0: any): raises PcgError naming the first entry
    that differs from the direct node-by-node construction."""
    arr = (C.c_int * len(table))(*table)
    _check(load().pcg_bsr_plan_check(leaves, bmax, arr, len(table)))


def bsr_exact(a: int, b: int, max_limbs: int):
    """(sign, exp, limbs) tuples of P, Q, R of the subtree [a, b) computed on
    the GPU (exact while no value exceeds max_limbs)."""
    ensure_init()
    bound = bsr_bound(a, b)
    bufs = [np.zeros(bound + 1, dtype=np.uint32) for _ in range(3)]
    vals = [BsrValue(buf.ctypes.data, buf.size, 0, 1) for buf in bufs]
    _check(load().pcg_bsr_exact(a, b, max_limbs, C.byref(vals[0]), C.byref(vals[1]),
                                C.byref(vals[2])))
    return [(bool(v.sign), 0, buf[: v.L].tolist()) for v, buf in zip(vals, bufs)]


def df_to_tuple(x: DFloat):
    """(sign, exp, limbs) of a device value (copied to the host)."""
    n = int(x.L)
    out = np.zeros(max(n, 1), dtype=np.uint32)
    if n:
        _check(load().pcg_df_read(C.byref(x), 0, n, out.ctypes.data))
    return (bool(x.sign), int(x.exp), out[:n].tolist())


def df_from_tuple(v) -> DFloat:
    """A device value holding the (sign, exp, limbs) tuple v."""
    ensure_init()
    sign, exp, limbs = v
    arr = _u32(limbs)
    x = DFloat()
    _check(load().pcg_df_set(C.byref(x), arr.ctypes.data, arr.size, int(exp), int(bool(sign))))
    return x


def df_mul_pair(x1, x2, s, p: int = 0, flags: int = 0, alias: str = ""):
    """(x1.mul(s, p), x2.mul(s, p)) through pcg_df_mul_pair on device values
    made from (sign, exp, limbs) tuples.  alias "x1" / "s" writes the first
    result over that operand (the entry point allows it)."""
    vals = [df_from_tuple(v) for v in (x1, x2, s)]
    z1, z2 = DFloat(), DFloat()
    t1 = {"x1": vals[0], "s": vals[2]}.get(alias, z1)
    try:
        _check(load().pcg_df_mul_pair(C.byref(t1), C.byref(vals[0]), C.byref(z2), C.byref(vals[1]),
                                      C.byref(vals[2]), p, flags))
        return df_to_tuple(t1), df_to_tuple(z2)
    finally:
        for v in vals + [z1, z2]:
            load().pcg_df_free(C.byref(v))


KEEP_B = 0x1000
REUSE_B = 0x2000


def df_mul_seq(s, xs, p: int = 0, flags=None, between=None):
    """[x.mul(s, p) for x in xs] through pcg_df_mul on device values made from
    (sign, exp, limbs) tuples, with flags[i] for product i (PCG_KEEP_B /
    PCG_REUSE_B: s's transform kept and reused); between(i), if given, runs
    after product i (e.g. another product that overwrites the kept one)."""
    flags = flags or [0] * len(xs)
    sv = df_from_tuple(s)
    vals = [df_from_tuple(x) for x in xs]
    zs = [DFloat() for _ in xs]
    try:
        for i, (x, z) in enumerate(zip(vals, zs)):
            _check(load().pcg_df_mul(C.byref(z), C.byref(x), C.byref(sv), p, flags[i]))
            if between:
                between(i)
        return [df_to_tuple(z) for z in zs]
    finally:
        for v in [sv] + vals + zs:
            load().pcg_df_free(C.byref(v))


def df_bsr_exact(a: int, b: int, max_limbs: int):
    """(sign, exp, limbs) of P, Q, R of the subtree [a, b) through
    pcg_df_bsr_exact (results left in device memory, then read back)."""
    ensure_init()
    vals = [DFloat() for _ in range(3)]
    try:
        _check(load().pcg_df_bsr_exact(a, b, max_limbs, *[C.byref(v) for v in vals]))
        return [df_to_tuple(v) for v in vals]
    finally:
        for v in vals:
            load().pcg_df_free(C.byref(v))


SMALL_BUCKETS = 13  # b = 0..12: small-product classes
TABLE_ENTRIES = 21  # and b = 13..20: pass geometry of batched products at 2^b points


def autotune():
    """Re-runs the autotuner; returns the class per bucket (entries 0..12)
    and the pass geometry per batched size (entries 13..20)."""
    ensure_init()
    t = (C.c_int * TABLE_ENTRIES)()
    _check(load().pcg_autotune(t, TABLE_ENTRIES))
    return list(t)


def batch_table(table=None):
    """Reads the class / geometry table, or installs `table`."""
    ensure_init()
    t = (C.c_int * TABLE_ENTRIES)(*(table or [0] * TABLE_ENTRIES))
    _check(load().pcg_batch_table(t, TABLE_ENTRIES, 1 if table is not None else 0))
    return list(t)
