"""DeviceFloat (values resident in GPU memory, pcg_df_* entry points)
against the oracle's BigFloat: mul with its precision chop, add/sub with
exponent alignment, bottom cut and leading-zero strip, mul by a word,
ucmp, and the Newton reciprocal, division and 1/sqrt (reference
BigFloat.cpp:123-567).  Bit-exact on every (sign, exp, limbs)."""
import random

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def host(gpu):
    from piculator_amd import pihost

    pihost.load()
    return pihost


def rand_val(rng, n, exp=None, sign=None, top_nonzero=True):
    limbs = [rng.getrandbits(32) for _ in range(n)]
    if n and top_nonzero:
        limbs[-1] |= 1
    return (rng.random() < 0.5 if sign is None else sign,
            rng.randint(-40, 40) if exp is None else exp, limbs)


def same(got, want):
    """Equality of (sign, exp, limbs) with a short description of the first
    difference (lengths, index, values) for the assertion message."""
    if got == want:
        return True, ""
    gl, wl = got[2], want[2]
    i = next((k for k in range(min(len(gl), len(wl))) if gl[k] != wl[k]), None)
    return False, "sign %s/%s exp %s/%s len %d/%d first diff %s%s" % (
        got[0], want[0], got[1], want[1], len(gl), len(wl), i,
        "" if i is None else " got %#x want %#x" % (gl[i], wl[i]))


def check(got, want):
    ok, msg = same(got, want)
    assert ok, msg


def oracle_op(op, a, b, p):
    oa, ob = O.BF(*a), O.BF(*b)
    if op == "mul":
        return oa.mul(ob, p).tuple()
    if op == "add":
        return oa.add(ob, p).tuple()
    if op == "sub":
        return oa.sub(ob, p).tuple()
    if op == "mul_u32":
        return oa.mul_u32(b[2][0] if b[2] else 0).tuple()
    raise ValueError(op)


SIZES = [1, 2, 3, 7, 100, 1023, 1024, 1025, 4097, 70001]


@pytest.mark.parametrize("seed", range(24))
def test_ops_match_oracle(host, seed):
    rng = random.Random(1000 + seed)
    a = rand_val(rng, rng.choice(SIZES))
    b = rand_val(rng, rng.choice(SIZES))
    for p in (0, 1, 10, 1000, 50000):
        for op in ("mul", "add", "sub"):
            got = host.bf_op(op, a, b, p, device=True)
            assert got == oracle_op(op, a, b, p), (op, p, len(a[2]), len(b[2]))
    x = (True, 0, [rng.getrandbits(32) | 1])
    assert host.bf_op("mul_u32", a, x, device=True) == oracle_op("mul_u32", a, x, 0)


def test_carry_and_borrow_chains(host):
    n = 5000
    ones = (True, 0, [0xFFFFFFFF] * n)
    one = (True, 0, [1])
    # carry through every limb and out of the top (within one block of
    # 1024 limbs, then across blocks)
    for m in (3, 1024, 1500, n):
        x = (True, 0, [0xFFFFFFFF] * m)
        check(host.bf_op("add", x, one, 0, device=True), oracle_op("add", x, one, 0))
    # borrow through every limb, then the strip of all leading zeros
    big = (True, 0, [0] * n + [1])
    check(host.bf_op("sub", big, one, 0, device=True), oracle_op("sub", big, one, 0))
    # equal magnitudes: zero with exponent 0, positive
    a = (False, 3, [5, 6, 7])
    b = (False, 3, [5, 6, 7])
    check(host.bf_op("sub", a, b, 0, device=True), oracle_op("sub", a, b, 0))
    check(host.bf_op("add", a, (True, 3, [5, 6, 7]), 0, device=True),
          oracle_op("add", a, (True, 3, [5, 6, 7]), 0))
    # multiply by a word with maximal carries
    w = (True, 0, [0xFFFFFFFF])
    check(host.bf_op("mul_u32", ones, w, device=True), oracle_op("mul_u32", ones, w, 0))
    # one-limb trim of a product
    s = (True, 0, [0, 0, 1])
    check(host.bf_op("mul", s, s, 0, device=True), oracle_op("mul", s, s, 0))


def test_alignment_and_cut(host):
    rng = random.Random(7)
    for _ in range(30):
        a = rand_val(rng, rng.randint(1, 3000), exp=rng.randint(-5000, 5000))
        b = rand_val(rng, rng.randint(1, 3000), exp=rng.randint(-5000, 5000))
        p = rng.choice([0, 1, 5, 100, 2500])
        for op in ("add", "sub"):
            assert host.bf_op(op, a, b, p, device=True) == oracle_op(op, a, b, p), (op, p)


def test_ucmp(host):
    rng = random.Random(11)
    cases = []
    for _ in range(20):
        a = rand_val(rng, rng.randint(1, 2000), exp=0)
        b = (a[0], 0, list(a[2]))
        i = rng.randrange(len(b[2]))
        b[2][i] ^= 1 << rng.randrange(32)
        if b[2][-1] == 0:
            b[2][-1] = 1
        cases.append((a, b))
    cases.append(((True, 2, [1, 2]), (True, 0, [0, 0, 1, 2])))   # same value, other exponent
    cases.append(((True, 2, [1, 2]), (True, 0, [5, 0, 1, 2])))
    cases.append(((True, 0, []), (True, 0, [])))
    for a, b in cases:
        assert host.bf_op("ucmp", a, b, device=True) == O.BF(*a).ucmp(O.BF(*b))


@pytest.mark.parametrize("p", [0, 1, 2, 3, 17, 500, 5000])
def test_newton_match_oracle(host, p):
    rng = random.Random(p)
    a = rand_val(rng, 600, exp=-7, sign=False)
    b = rand_val(rng, 400, exp=2, sign=True)
    oa, ob = O.BF(*a), O.BF(*b)
    assert host.bf_op("rcp", a, p=p, device=True) == oa.rcp(p).tuple()
    assert host.bf_op("div", b, a, p=p, device=True) == ob.div(oa, p).tuple()
    assert host.bf_op("invsqrt", (True, 0, [10005]), p=p, device=True) == O.invsqrt(10005, p).tuple()


def test_errors_are_reference_literals(host):
    with pytest.raises(RuntimeError, match="Divide by Zero"):
        host.bf_op("rcp", (True, 0, []), p=3, device=True)


def test_large_values_residue(host):
    """Multi-million-limb add and product: size-independent residue checks."""
    from tests.helpers import residue, CHECK_PRIMES

    rng = np.random.default_rng(5)
    n = 3_000_000
    a = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    a[-1] |= 1
    b[-1] |= 1
    s = host.bf_op("add", (True, 0, a), (True, 0, b), 0, device=True)
    z = np.asarray(s[2], dtype=np.uint32)
    for q in CHECK_PRIMES[:2]:
        assert residue(z, q) == (residue(a, q) + residue(b, q)) % q
    m = host.bf_op("mul", (True, 0, a), (True, 0, b), 0, device=True)
    z = np.asarray(m[2], dtype=np.uint32)
    assert m[1] == 0
    for q in CHECK_PRIMES[:2]:
        assert residue(z, q) == residue(a, q) * residue(b, q) % q


# (len x1, len x2, len s, p): one transform length for both products (the
# shared forward transform), two lengths (two separate products), the
# precision cut of every operand, one-limb operands and zeros
PAIR_CASES = [
    (100, 120, 110, 0),
    (3000, 2900, 3100, 0),
    (1000, 5000, 3000, 0),
    (70001, 65000, 70001, 0),
    (70001, 65000, 70001, 50000),
    (70001, 3, 70001, 1000),
    (1, 1, 1, 0),
    (0, 50, 40, 0),
    (50, 40, 0, 7),
]


@pytest.mark.parametrize("n1,n2,ns,p", PAIR_CASES)
def test_mul_pair_matches_oracle(gpu, n1, n2, ns, p):
    """pcg_df_mul_pair gives exactly x1.mul(s, p) and x2.mul(s, p)."""
    rng = random.Random(n1 * 7 + n2 * 13 + ns * 3 + p)
    x1, x2, s = rand_val(rng, n1), rand_val(rng, n2), rand_val(rng, ns)
    want1 = O.BF(*x1).mul(O.BF(*s), p).tuple()
    want2 = O.BF(*x2).mul(O.BF(*s), p).tuple()
    got1, got2 = gpu.df_mul_pair(x1, x2, s, p)
    check(got1, want1)
    check(got2, want2)


@pytest.mark.parametrize("alias", ["x1", "s"])
def test_mul_pair_result_over_operand(gpu, alias):
    """The first result may replace an operand: both products are complete
    before either result is assigned."""
    rng = random.Random(31)
    x1, x2, s = rand_val(rng, 9000), rand_val(rng, 8500), rand_val(rng, 9100)
    got1, got2 = gpu.df_mul_pair(x1, x2, s, 0, alias=alias)
    check(got1, O.BF(*x1).mul(O.BF(*s), 0).tuple())
    check(got2, O.BF(*x2).mul(O.BF(*s), 0).tuple())


def test_mul_pair_fft_engine(gpu):
    """With the FFT engine requested the pair runs as two products."""
    rng = random.Random(32)
    x1, x2, s = rand_val(rng, 4000), rand_val(rng, 4100), rand_val(rng, 3900)
    got1, got2 = gpu.df_mul_pair(x1, x2, s, 0, flags=2)
    check(got1, O.BF(*x1).mul(O.BF(*s), 0).tuple())
    check(got2, O.BF(*x2).mul(O.BF(*s), 0).tuple())


def test_mul_pair_rejects_one_result_twice(gpu):
    x = gpu.df_from_tuple((True, 0, [1, 2, 3]))
    r = gpu.C.byref
    try:
        assert gpu.load().pcg_df_mul_pair(r(x), r(x), r(x), r(x), r(x), 0, 0) == 1  # PCG_ERR_INVALID
        assert gpu.df_to_tuple(x) == (True, 0, [1, 2, 3])
    finally:
        gpu.load().pcg_df_free(r(x))


def test_mul_pair_wrapped_products(gpu):
    """A pair whose products wrap (zL - 1 just above 2^20): one forward
    transform of s of 2^20 points, each product's short top product; exact
    (p = 0) results checked by residues and by their lowest limbs."""
    from tests.helpers import residue, CHECK_PRIMES, from_int, to_int

    rng = np.random.default_rng(77)
    n = (1 << 19) + 5000
    s = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    x1 = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    x2 = rng.integers(0, 2**32, n - 3000, dtype=np.uint64).astype(np.uint32)
    for v in (s, x1, x2):
        v[-1] |= 1
    got1, got2 = gpu.df_mul_pair((True, 0, x1), (True, 0, x2), (True, 0, s), 0)
    for x, got in ((x1, got1), (x2, got2)):
        z = np.asarray(got[2], dtype=np.uint32)
        assert z.size in (x.size + s.size, x.size + s.size - 1)
        for q in CHECK_PRIMES:
            assert residue(z, q) == residue(x, q) * residue(s, q) % q
        k = 12000
        assert np.array_equal(z[:k], from_int((to_int(x[:k]) * to_int(s[:k])) % (1 << (32 * k)), k))


@pytest.mark.parametrize("ns,nx,p", [(3000, 6000, 0), (30000, 60000, 0), (40000, 90000, 50000),
                                     ((1 << 19) + 9000, (1 << 19) + 20000, 0),
                                     ((1 << 19) + 9000, (1 << 19) - 8993, 0)])
def test_kept_transform_reused(gpu, ns, nx, p):
    """x.mul(s) with PCG_KEEP_B, then more products by s with PCG_REUSE_B
    (the Newton pattern): every result exactly as without the flags.  The
    second operand is 7 limbs shorter: the same transform length, for the
    wrapped sizes another wrap (a new short product) or, at the last size,
    no wrap (2^20 - 1 coefficients after 2^20 + 6); the last operand has a
    different transform length (recomputed)."""
    rng = random.Random(ns + nx + p)
    s = rand_val(rng, ns)
    xs = [rand_val(rng, nx), rand_val(rng, nx - 7), rand_val(rng, nx // 3)]
    got = gpu.df_mul_seq(s, xs, p, [gpu.KEEP_B, gpu.REUSE_B, gpu.REUSE_B])
    for g, x in zip(got, xs):
        check(g, O.BF(*x).mul(O.BF(*s), p).tuple())


@pytest.mark.parametrize("ns", [(1 << 21) - 3, (1 << 21) + 5])
def test_kept_square_transform_reused(gpu, ns):
    """s.mul(s) with PCG_KEEP_B (the square path), then x.mul(s) with
    PCG_REUSE_B -- the 1/sqrt Newton step's r0 r0, then t r0 -- at a
    large-tile transform length (2^22 points; the second size wraps onto a
    2^22-point main transform): both results exact."""
    from tests.helpers import CHECK_PRIMES, residue

    rng = np.random.default_rng(ns)
    s = rng.integers(0, 2**32, ns, dtype=np.uint64).astype(np.uint32)
    x = rng.integers(0, 2**32, ns - 7, dtype=np.uint64).astype(np.uint32)
    s[-1] |= 1
    x[-1] |= 1
    L, r = gpu.load(), gpu.C.byref

    def products(flags):
        sv, xv = gpu.df_from_tuple((True, 0, s)), gpu.df_from_tuple((False, 3, x))
        z1, z2 = gpu.DFloat(), gpu.DFloat()
        try:
            gpu._check(L.pcg_df_mul(r(z1), r(sv), r(sv), 0, flags[0]))
            gpu._check(L.pcg_df_mul(r(z2), r(xv), r(sv), 0, flags[1]))
            return gpu.df_to_tuple(z1), gpu.df_to_tuple(z2)
        finally:
            for v in (sv, xv, z1, z2):
                L.pcg_df_free(r(v))

    got = products((gpu.KEEP_B, gpu.REUSE_B))
    want = products((0, 0))
    for g, w in zip(got, want):
        check(g, w)
    for (sign, exp, z), y, sgn in ((got[0], s, True), (got[1], x, False)):
        z = np.asarray(z, dtype=np.uint32)
        assert sign == sgn and exp == (0 if sgn else 3)
        assert z.size == y.size + s.size or (z.size == y.size + s.size - 1)
        for q in CHECK_PRIMES:
            assert residue(z, q) == residue(y, q) * residue(s, q) % q


def test_kept_transform_overwritten(gpu):
    """Another product between the keep and the reuse overwrites the kept
    transform: the reuse recomputes it, results exact."""
    rng = random.Random(5)
    s, x0, x1 = rand_val(rng, 20000), rand_val(rng, 20000), rand_val(rng, 20001)
    other = (rand_val(rng, 20000), rand_val(rng, 19000))

    def between(i):
        if i == 0:
            gpu.df_mul_pair(other[0], other[1], other[1], 0)
            gpu.mul(np.array(other[0][2], dtype=np.uint32), np.array(other[1][2], dtype=np.uint32))

    got = gpu.df_mul_seq(s, [x0, x1], 0, [gpu.KEEP_B, gpu.REUSE_B], between)
    check(got[0], O.BF(*x0).mul(O.BF(*s), 0).tuple())
    check(got[1], O.BF(*x1).mul(O.BF(*s), 0).tuple())
