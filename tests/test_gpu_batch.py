"""Batched products (pcg_mul_batch / pcg_mul_batch_dev) against exact
arithmetic: every size class (schoolbook, single-workgroup NTT, multi-pass
NTT), uneven lengths, and identity with the plain one-launch-per-product path."""
import random

import numpy as np
import pytest

from piculator_amd import workload
from tests.helpers import exact_product, limb_digest, mul_digests

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF


def rand_limbs(rng, n, ones=False):
    if ones:
        return np.full(n, M32, dtype=np.uint32)
    return np.array([rng.getrandbits(32) for _ in range(n)], dtype=np.uint32)


def test_every_class_exact(gpu):
    rng = random.Random(11)
    sizes = [(1, 1), (1, 1023), (512, 512), (3, 700), (1000, 24), (1025, 1), (600, 1000),
             (2048, 2048), (1, 4095), (3000, 1096), (4097, 1), (5000, 5000), (20000, 3)]
    pairs = []
    for k, (na, nb) in enumerate(sizes):
        ones = k % 4 == 3
        pairs.append((rand_limbs(rng, na, ones), rand_limbs(rng, nb, ones)))
    outs = gpu.mul_batch(pairs)
    for (a, b), o in zip(pairs, outs):
        assert np.array_equal(o, exact_product(a, b)), (a.size, b.size)


def test_zero_length_operands_in_batch(gpu):
    a = np.array([1, 2, 3], dtype=np.uint32)
    e = np.zeros(0, dtype=np.uint32)
    outs = gpu.mul_batch([(a, e), (e, e), (a, a)])
    assert np.array_equal(outs[0], np.zeros(3, dtype=np.uint32))
    assert outs[1].size == 0
    assert np.array_equal(outs[2], exact_product(a, a))


def test_batch_equals_plain_on_uneven_workload(gpu):
    pairs = workload.batch_operands(512)
    got = gpu.mul_batch(pairs)
    plain = gpu.mul_batch(pairs, plain=True)
    for (a, b), g, p in zip(pairs, got, plain):
        assert np.array_equal(g, p)
        assert np.array_equal(g, exact_product(a, b))


def test_mid_size_products_batched(gpu):
    """Products of 2^18 .. 2^21 transform points run as batched classes (one
    launch sequence per size class) and 2^22 as a single product: every
    output equals the plain one-product-per-launch path, the smallest one the
    exact product, all of them the product modulo the check primes."""
    from tests.helpers import residues_match

    rng = np.random.default_rng(2024)
    sizes = [(70000, 65000), (70001, 64999), (150000, 110000), (300000, 200000), (600000, 400000),
             (1000000, 900000), (1500000, 1500000)]
    pairs = []
    for k, (na, nb) in enumerate(sizes):
        if k == 1:
            pairs.append((np.full(na, M32, dtype=np.uint32), np.full(nb, M32, dtype=np.uint32)))
        else:
            pairs.append((rng.integers(0, 2**32, na, dtype=np.uint64).astype(np.uint32),
                          rng.integers(0, 2**32, nb, dtype=np.uint64).astype(np.uint32)))
    got = gpu.mul_batch(pairs)
    plain = gpu.mul_batch(pairs, plain=True)
    for (a, b), g, p in zip(pairs, got, plain):
        assert np.array_equal(g, p), (a.size, b.size)
        assert residues_match(a, b, g), (a.size, b.size)
    for k in (0, 1):
        assert np.array_equal(got[k], exact_product(*pairs[k]))


def test_full_uneven_batch_4096(gpu):
    pairs = workload.batch_operands(4096)
    lens = workload.batch_lengths(4096)
    assert all(16 <= na <= 16384 and 16 <= nb <= 16384 for na, nb in lens)
    got = gpu.mul_batch(pairs)
    want = mul_digests()["batch"]["sha256"]
    assert len(want) == 4096
    for i, ((a, b), g) in enumerate(zip(pairs, got)):
        assert g.size == a.size + b.size
        assert limb_digest(g) == want[i], (i, a.size, b.size)


def test_batch_dev_pointers(gpu):
    rng = random.Random(5)
    pairs = [(rand_limbs(rng, rng.randint(1, 3000)), rand_limbs(rng, rng.randint(1, 3000)))
             for _ in range(40)]
    bufs = []
    descs = (gpu.MulDesc * len(pairs))()
    for i, (a, b) in enumerate(pairs):
        da, db = gpu.DeviceBuffer(a.nbytes), gpu.DeviceBuffer(b.nbytes)
        do = gpu.DeviceBuffer(a.nbytes + b.nbytes)
        da.upload(a)
        db.upload(b)
        bufs.append((da, db, do))
        descs[i] = gpu.MulDesc(da.ptr, a.size, db.ptr, b.size, do.ptr)
    gpu._check(gpu.load().pcg_mul_batch_dev(descs, len(pairs), 0))
    for (a, b), (_, _, do) in zip(pairs, bufs):
        out = np.empty(a.size + b.size, dtype=np.uint32)
        do.download(out)
        assert np.array_equal(out, exact_product(a, b))


SCHOOL_MAX = {0: 1024, 1: 64, 2: 256, 3: 512}


def class_valid(b, c):
    if c in SCHOOL_MAX:
        return (1 << b) <= SCHOOL_MAX[c]
    if c == 4:  # one Karatsuba level over the schoolbook
        return (1 << b) <= 1024
    if 110 <= c <= 112:  # batched pass kernels at 2^(c - 100) points
        return c - 100 >= b
    return 10 <= c <= 12 and c >= b


def test_autotune_table_is_valid(gpu):
    t = gpu.autotune()
    assert len(t) == 21
    for b, c in enumerate(t[:13]):
        assert class_valid(b, c), (b, c)
        if (1 << b) > 1024:
            assert c in (b, 100 + b)
    # pass geometry per batched size: 0 the 4096-point passes, 1 large tiles
    assert all(g in (0, 1) for g in t[13:])
    assert gpu.batch_table() == t


@pytest.mark.parametrize("geom", [0, 1])
def test_forced_pass_geometry_exact(gpu, geom):
    """Batched multi-pass products of 2^13 .. 2^17 points on each pass
    geometry (4096-point passes; large tiles: one fused pass from the limbs
    at 2^14 points -- and at 2^13, where it runs at 2^14 points --, register
    column passes at 2^15 .. 2^17) are exact, unbalanced and all-ones
    operands included."""
    saved = gpu.batch_table()
    gpu.batch_table(saved[:13] + [geom] * 8)
    try:
        rng = random.Random(77 + geom)
        sizes = [(4000, 4100), (8191, 8190), (5000, 11000), (16000, 15000), (30000, 33000),
                 (65000, 65000), (100, 60000), (12000, 4500)]
        pairs = [(rand_limbs(rng, na, k % 3 == 1), rand_limbs(rng, nb, k % 3 == 1))
                 for k, (na, nb) in enumerate(sizes)]
        pairs = pairs * 2
        outs = gpu.mul_batch(pairs)
        for (a, b), o in zip(pairs, outs):
            assert np.array_equal(o, exact_product(a, b)), (geom, a.size, b.size)
    finally:
        gpu.batch_table(saved)


def test_subtree_on_each_pass_geometry(gpu):
    """A subtree of 2^17 leaves near 7e7 (batched products up to 2^20 points)
    gives the same P, Q, R on both pass geometries for every size."""
    saved = gpu.batch_table()
    a, b = 70_000_000, 70_000_000 + (1 << 17)
    got = []
    try:
        for geom in (0, 1):
            gpu.batch_table(saved[:13] + [geom] * 8)
            got.append(gpu.bsr_exact(a, b, 1 << 28))
    finally:
        gpu.batch_table(saved)
    assert got[0] == got[1]


@pytest.mark.parametrize("forced", [0, 1, 2, 3, 4, 10, 110])
def test_forced_small_classes_exact(gpu, forced):
    """Every kernel of the tuned crossover (schoolbook with 256 threads, with
    16 or 64 threads per product and 4 or 8 columns per thread, one
    Karatsuba level over the schoolbook, 2^10-point single-workgroup NTT,
    batched pass kernels at 2^10 points) gives the exact product in every
    bucket it may serve, including all-ones carry chains and unbalanced
    shapes."""
    saved = gpu.batch_table()
    table = [forced if class_valid(b, forced) else c for b, c in enumerate(saved)]
    gpu.batch_table(table)
    try:
        rng = random.Random(5 + forced)
        sizes = [(1, 1), (2, 1), (3, 4), (7, 9), (16, 16), (33, 31), (64, 64), (100, 27),
                 (128, 128), (255, 257), (300, 212), (384, 384), (511, 513), (1, 1023),
                 (1023, 1), (700, 324), (1, 63), (31, 33), (40, 24), (1, 255), (128, 127),
                 (512, 512), (1000, 24), (513, 511), (24, 1000), (2, 1022)]
        pairs = []
        for k, (na, nb) in enumerate(sizes):
            ones = k % 3 == 2
            pairs.append((rand_limbs(rng, na, ones), rand_limbs(rng, nb, ones)))
        pairs = pairs * 3  # several products per launch of each class
        outs = gpu.mul_batch(pairs)
        for (a, b), o in zip(pairs, outs):
            assert np.array_equal(o, exact_product(a, b)), (forced, a.size, b.size)
    finally:
        gpu.batch_table(saved)
