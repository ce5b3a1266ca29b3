"""GPU parity of the NTT product (through the C ABI) against exact arithmetic
and the oracle's restatement of the reference multiply (BigFloat.cpp:377-441)."""
import random

import numpy as np
import pytest

from oracle import oracle as O
from piculator_amd import workload
from tests.helpers import (exact_product, limb_digest, oracle_product, residues_match,
                           single_product_digest)

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF


def rand_limbs(rng, n):
    return np.array([rng.getrandbits(32) for _ in range(n)], dtype=np.uint32)


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 7), (3, 5), (17, 33), (100, 77), (255, 1),
                                   (129, 127), (1000, 999), (2047, 2049), (2048, 2048),
                                   (4095, 1), (3000, 5000), (12345, 6789)])
def test_random_small(gpu, na, nb):
    rng = random.Random(na * 7919 + nb)
    a, b = rand_limbs(rng, na), rand_limbs(rng, nb)
    got = gpu.mul(a, b, gpu.ENGINE_NTT)
    assert np.array_equal(got, exact_product(a, b))


@pytest.mark.parametrize("n", [1, 2, 3, 64, 511, 4096, 70000])
def test_all_ones_carry_chains(gpu, n):
    # (2^(32n) - 1)^2 has the longest carry chains
    a = np.full(n, M32, dtype=np.uint32)
    got = gpu.mul(a, a, gpu.ENGINE_NTT)
    assert np.array_equal(got, exact_product(a, a))


def test_zero_operand(gpu):
    a = np.zeros(0, dtype=np.uint32)
    b = np.array([5, 6], dtype=np.uint32)
    assert np.array_equal(gpu.mul(a, b), np.zeros(2, dtype=np.uint32))
    z = np.zeros(50, dtype=np.uint32)
    assert np.array_equal(gpu.mul(z, b), np.zeros(52, dtype=np.uint32))


def test_sparse_and_power_of_two(gpu):
    a = np.zeros(3000, dtype=np.uint32)
    a[-1] = 1 << 31
    b = np.zeros(2000, dtype=np.uint32)
    b[0] = 1
    b[-1] = M32
    assert np.array_equal(gpu.mul(a, b), exact_product(a, b))


@pytest.mark.parametrize("na,nb", [(4000, 3000), (20000, 9000)])
def test_against_oracle_ntt(gpu, na, nb):
    rng = random.Random(na)
    a, b = rand_limbs(rng, na), rand_limbs(rng, nb)
    ref = O.mul_limbs(a, b, O.ENGINE_NTT)
    assert np.array_equal(gpu.mul(a, b, gpu.ENGINE_NTT), ref)


@pytest.mark.parametrize("logn", [16, 20, 24, 26])
@pytest.mark.parametrize("unbalanced", [False, True])
def test_workload_exact(gpu, logn, unbalanced):
    """The single-multiply workload (n = 2^16 .. 2^26 limbs, n x n and
    n x n/2): every limb equals the oracle's product (the digest of its NTT
    product, which its FFT and, up to 2^20, Python integers reproduce)."""
    n = 1 << logn
    nb = n // 2 if unbalanced else n
    a, b = workload.operand_pair(n, nb)
    got = gpu.mul(a, b, gpu.ENGINE_NTT)
    assert got.size == n + nb
    if logn <= 16:
        assert np.array_equal(got, exact_product(a, b))
    assert limb_digest(got) == single_product_digest(n, nb)


@pytest.mark.parametrize("logN", [21, 22, 23, 24, 25])
def test_half_filled_operands(gpu, logN):
    """Operands of at most N/2 limbs at every large-tile transform length
    (2^21 .. 2^25 points: their forward column pass runs as two half-length
    DFTs per tile): the NTT product equals the FFT engine's and matches the
    check-prime residues; an operand one limb longer than N/2 (the full
    column pass) too; all-ones operands of N/2 limbs square to the closed
    form (2^(32n) - 1)^2."""
    N = 1 << logN
    rng = np.random.default_rng(logN)
    for na, nb in ((N // 2, N // 2 - 12345), (N // 2 + 1, N // 4 + 3)):
        a = rng.integers(0, 2**32, na, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, 2**32, nb, dtype=np.uint64).astype(np.uint32)
        a[-1] |= 1
        b[-1] |= 1
        got = gpu.mul(a, b, gpu.ENGINE_NTT)
        assert got.size == na + nb
        assert residues_match(a, b, got), (na, nb)
        assert np.array_equal(got, gpu.mul(a, b, gpu.ENGINE_FFT)), (na, nb)
    n = N // 2
    ones = np.full(n, M32, dtype=np.uint32)
    got = gpu.mul(ones, ones.copy(), gpu.ENGINE_NTT)
    want = np.zeros(2 * n, dtype=np.uint32)
    want[0] = 1
    want[n] = 0xFFFFFFFE
    want[n + 1:] = M32
    assert np.array_equal(got, want)


def _low_limbs(a, b, k):
    """The lowest k limbs of a*b (they depend only on the lowest k limbs)."""
    from tests.helpers import from_int, to_int

    m = (1 << (32 * k)) - 1
    return from_int((to_int(a[:k]) * to_int(b[:k])) & m, k)


# Products whose zL - 1 coefficients exceed 2^20 by at most 2^18 run as a
# cyclic convolution of 2^20 points plus the short product of the operands'
# top limbs (the wrapped coefficients); the boundary cases are one wrapped
# coefficient, none (zL - 1 = 2^20 fits), and the largest wrap.
@pytest.mark.parametrize("na,nb", [((1 << 19) + 4096, (1 << 19) + 4096),
                                   ((1 << 20) - 5, 70000),
                                   ((1 << 19) + 1, (1 << 19) + 1),
                                   (1 << 19, (1 << 19) + 1),
                                   ((1 << 19) + 65536, (1 << 19) + 65536),
                                   ((1 << 19) + 131072, (1 << 19) + 131072)])
def test_wrapped_products(gpu, na, nb):
    rng = np.random.default_rng(na ^ (nb << 3))
    a = rng.integers(0, 2**32, na, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, nb, dtype=np.uint64).astype(np.uint32)
    a[-1] |= 1
    b[-1] |= 1
    got = gpu.mul(a, b, gpu.ENGINE_NTT)
    assert got.size == na + nb
    k = min(na, nb, 270000)  # covers every wrapped (corrected) coefficient
    assert np.array_equal(got[:k], _low_limbs(a, b, k))
    assert np.array_equal(got, oracle_product(a, b))


@pytest.mark.parametrize("n", [(1 << 19) + 3000])
def test_wrapped_carry_chains(gpu, n):
    a = np.full(n, M32, dtype=np.uint32)
    got = gpu.mul(a, a, gpu.ENGINE_NTT)
    # (2^(32n) - 1)^2 = 2^(64n) - 2^(32n+1) + 1: limbs 1, 0..0, 0xFFFFFFFE, F..F
    want = np.zeros(2 * n, dtype=np.uint32)
    want[0] = 1
    want[n] = 0xFFFFFFFE
    want[n + 1:] = M32
    assert np.array_equal(got, want)


@pytest.mark.parametrize("n", [3000, 5000, (1 << 16) + 7, (1 << 19) + 65536, (1 << 20) - 3,
                               (1 << 21) + 100000])
def test_squares_on_one_device_buffer(gpu, n):
    """a * a with both operands at one device address takes the square path
    (a's forward transform reused as the fused pass's input); plain, wrapped
    (2n - 1 coefficients a little above a power of two) and all-ones."""
    import ctypes as C

    L = gpu.load()
    rng = random.Random(n)
    for a in (rand_limbs(rng, n), np.full(n, M32, dtype=np.uint32)):
        d_a, d_o = C.c_void_p(), C.c_void_p()
        gpu._check(L.pcg_dev_alloc(C.byref(d_a), 4 * n))
        gpu._check(L.pcg_dev_alloc(C.byref(d_o), 8 * n))
        try:
            gpu._check(L.pcg_memcpy_h2d(d_a, a.ctypes.data, 4 * n))
            gpu.mul_dev(d_a.value, n, d_a.value, n, d_o.value, gpu.ENGINE_NTT, sync=True)
            got = np.empty(2 * n, dtype=np.uint32)
            gpu._check(L.pcg_memcpy_d2h(got.ctypes.data, d_o, 8 * n))
        finally:
            L.pcg_dev_free(d_a)
            L.pcg_dev_free(d_o)
        if n <= 1 << 17:
            assert np.array_equal(got, exact_product(a, a))
        else:  # the general path (operands in two buffers) and the oracle
            assert np.array_equal(got, gpu.mul(a, a, gpu.ENGINE_NTT))
            assert np.array_equal(got, oracle_product(a, a))
