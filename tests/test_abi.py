"""The C-ABI library loads and exports every entry point include/picugpu.h
declares (no GPU needed); error paths work without a device."""
import ctypes
import os
import re

import pytest

from piculator_amd import picugpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "picugpu.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(pcg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(picugpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(picugpu.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_probe_library_exports_the_probes():
    """The measurement probes bench.py loads (FP64 mulmod peak, HBM access
    patterns) are built beside the engine; an unknown pattern is refused
    without touching the GPU."""
    lib = ctypes.CDLL(os.path.join(os.path.dirname(picugpu.LIB_PATH), "libpcg_probe.so"))
    assert hasattr(lib, "pcg_probe_fp64_rates")
    assert hasattr(lib, "pcg_probe_stream")
    us, bps = ctypes.c_double(), ctypes.c_double()
    assert lib.pcg_probe_stream(3, 5, ctypes.byref(us), ctypes.byref(bps)) == 2


def test_calls_before_init_fail_cleanly():
    lib = picugpu.load()
    assert lib.pcg_mul(None, 0, None, 0, None, 0) == 2  # PCG_ERR_NOT_INIT
    assert b"pcg_init" in lib.pcg_last_error_string()


def test_plan_limits():
    # 2^21 .. 2^25 points: a column pass and a 2^14-point row pass
    lg, passes = picugpu.ntt_plan(2 ** 25)
    assert lg == 25 and passes == 2
    assert picugpu.ntt_plan(2 ** 21) == (21, 2)
    assert picugpu.ntt_plan(2 ** 26)[1] == 3 and picugpu.ntt_plan(2 ** 30)[1] == 3
    # every transform length has a kernel instance for each of its passes
    for lg in range(8, 31):
        assert picugpu.ntt_plan(2 ** lg)[0] == lg
    assert picugpu.ntt_plan(3)[0] == 8
    with pytest.raises(picugpu.PcgError):
        picugpu.ntt_plan(2 ** 31)


def _leaf_bits(b: int):
    """Bit lengths of |P|, Q, R of the leaf b (PI.h:67-88, uint32 factors)."""
    m32 = (1 << 32) - 1
    f1, f2, f3 = (2 * b - 1) & m32, (6 * b - 5) & m32, (6 * b - 1) & m32
    P = (13591409 + 545140134 * b) * f1 * f2 * f3
    Q = b ** 3 * 26726400 * 409297880
    return P.bit_length(), Q.bit_length(), (f1 * f2 * f3).bit_length()


def _tree_bound(a: int, b: int, leaf) -> int:
    """Largest value length (limbs) of the subtree [a, b), node by node over
    the explicit tree (midpoints m = (a + b) / 2 of PI.h:90), from bit bounds:
    every leaf's P, Q, R below 2^leaf[0], 2^leaf[1], 2^leaf[2], x = P_L Q_R
    and y = P_R R_L add bits, P = max(x, y) + 1, Q and R add; a value's
    length is ceil(bits / 32)."""
    mx = 0

    def limbs(bits):
        return (bits + 31) // 32

    def rec(a, b):
        nonlocal mx
        if b - a == 1:
            mx = max(mx, limbs(max(leaf)))
            return leaf
        m = (a + b) // 2
        lp, lq, lr = rec(a, m)
        rp, rq, rr = rec(m, b)
        x, y = lp + rq, rp + lr
        v = (max(x, y) + 1, lq + rq, lr + rr)
        mx = max(mx, limbs(x), limbs(y), *(limbs(t) for t in v))
        return v

    rec(a, b)
    return mx


def test_leaf_bit_bounds():
    """The leaf bounds hold for the largest leaf index (P, Q, R of PI.h:67-88
    with the reference's uint32 factors 2b-1, 6b-5, 6b-1)."""
    m32 = (1 << 32) - 1
    for b in (1, 2, 3, 70513670, 715827882, 715827883, m32 - 1, m32):
        f1, f2, f3 = (2 * b - 1) & m32, (6 * b - 5) & m32, (6 * b - 1) & m32
        P = (13591409 + 545140134 * b) * f1 * f2 * f3
        Q = b ** 3 * 26726400 * 409297880
        R = f1 * f2 * f3
        assert P.bit_length() <= 158 and Q.bit_length() <= 150 and R.bit_length() <= 96


def test_leaf_bits_grow_with_b():
    """Below the wrap of 6b - 1 (b <= 715827882) every leaf term grows with
    b, so the bits of the leaf b bound every leaf of [a, b); checked leaf by
    leaf on ranges up to the wrap."""
    for a, b in ((0, 3000), (70510000, 70513670), (715825000, 715827882)):
        top = _leaf_bits(b)
        for k in range(a + 1, b + 1):
            assert all(x <= y for x, y in zip(_leaf_bits(k), top)), k


def test_bsr_bound_matches_tree_walk():
    """pcg_bsr_bound: the tree walk from the bits of the leaf b, and from the
    bounds of any leaf (158, 150, 96) past the wrap."""
    from piculator_amd import picugpu

    for d in (1, 2, 3, 5, 64, 100, 1000, 4097, 70515):
        assert picugpu.bsr_bound(1000, 1000 + d) == _tree_bound(1000, 1000 + d, _leaf_bits(1000 + d)), d
    for a, d in ((70000000, 4097), (715827000, 882), (715827000, 883), (2000000000, 1000)):
        b = a + d
        leaf = _leaf_bits(b) if b <= 715827882 else (158, 150, 96)
        assert picugpu.bsr_bound(a, b) == _tree_bound(a, b, leaf), (a, b)


# Small-product class tables (pcg_batch_table layout, buckets 0..12): the
# defaults before tuning, and tables that move products into the
# single-workgroup NTT and batched-pass classes, so the plan's launch groups
# (paired, single, wrapped) change with them.
_PLAN_TABLES = [
    [1, 1, 1, 1, 1, 1, 1, 2, 2, 0, 0, 11, 12],
    [1, 1, 1, 1, 1, 1, 1, 2, 2, 0, 10, 11, 12],
    [1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 110, 111, 112],
    [1, 1, 1, 1, 1, 1, 2, 2, 3, 110, 10, 11, 112],
    [1, 1, 1, 1, 1, 1, 1, 2, 4, 4, 4, 11, 12],
]


def test_bsr_plan_layout_matches_direct_construction():
    """The subtree plan the GPU lays out from the size table (the same
    per-node code run on the CPU) equals the direct node-by-node
    construction: arena offsets, leaf lists, product and add lists in launch
    order, per-height launch groups.  Every size up to 300 leaves (all the
    small shapes, odd and even splits) and larger ones around powers of two."""
    from piculator_amd import picugpu

    sizes = list(range(1, 301)) + [1023, 1024, 1025, 4097, 12345, 65537, 100003]
    for t in _PLAN_TABLES:
        for d in sizes:
            picugpu.bsr_plan_check(d, t)
    # plans sized from the leaf bounds at the largest leaf index
    for bmax in (300, 2203553, 70513670, 715827882):
        for d in (1, 2, 3, 7, 100, 300, 4097):
            picugpu.bsr_plan_check(d, _PLAN_TABLES[0], bmax)


def test_bsr_plan_layout_at_pi_subtree_sizes():
    """The two subtree sizes of pi to 10^9 digits at depth 5 (70,513,670
    terms / 32) with the default table, for any leaf and for the last one."""
    from piculator_amd import picugpu

    for d in (2203552, 2203553):
        picugpu.bsr_plan_check(d, _PLAN_TABLES[0])
        picugpu.bsr_plan_check(d, _PLAN_TABLES[0], 70513670)


def test_bsr_plan_check_rejects_bad_tables():
    from piculator_amd import picugpu

    with pytest.raises(picugpu.PcgError):
        picugpu.bsr_plan_check(100, [1] * 12)
    with pytest.raises(picugpu.PcgError):
        picugpu.bsr_plan_check(100, [0] * 12 + [5])  # class 5 is not a class


def test_host_library_maps_allocation_failure_to_status():
    """A std::bad_alloc inside the host library (here: an operand of 2^60
    limbs, whose limb array cannot be allocated) comes back as status -1
    with a message instead of crossing extern "C" and terminating the
    process.  No GPU call is made: the operand is built before any product."""
    import ctypes as C

    import numpy as np

    from piculator_amd import pihost

    L = pihost.load()
    la = np.ones(4, dtype=np.uint32)
    sz, ez, nz = C.c_int(), C.c_int64(), C.c_uint64()
    out = np.zeros(8, dtype=np.uint32)
    rc = L.pch_bf_op(1, 1, 0, la.ctypes.data, 1 << 60, 1, 0, la.ctypes.data, 1, 0,
                     C.byref(sz), C.byref(ez), out.ctypes.data, out.size, C.byref(nz))
    assert rc == -1
    assert b"memory" in L.pch_last_error()
