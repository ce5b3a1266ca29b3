"""Pins the oracle's BigFloat semantics and Chudnovsky driver (no GPU).

Closed forms used (values are sum T[i] 2^(32 (exp + i)), sign-magnitude):
  * mul (BigFloat.cpp:326-444): the exact product of the operands each
    chopped to their top p+2 limbs (p = 0: no chop), one zero top limb trimmed;
  * add/sub (BigFloat.cpp:184-325): exact sum/difference of the operands
    with every word below top - (p+2) dropped;
  * rcp/div/invsqrt (BigFloat.cpp:445-567): accuracy bounds;
  * Pi (PI.h:138-210): hex digits of pi (mpmath) on the correct prefix,
    Bailey-Borwein-Plouffe digit extraction at spot positions.
"""
from fractions import Fraction
import hashlib
import random

import mpmath
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import to_int


def val(bf: O.BF) -> Fraction:
    s, e, limbs = bf.tuple()
    v = Fraction(to_int(limbs)) * Fraction(2) ** (32 * e)
    return v if s else -v


def chop(limbs, exp, p):
    limbs = list(limbs)
    if len(limbs) > p:
        c = len(limbs) - p
        return limbs[c:], exp + c
    return limbs, exp


def rand_bf(rng, n, exp=None, sign=None):
    limbs = [rng.getrandbits(32) for _ in range(n)]
    limbs[-1] |= 1
    return O.BF(rng.random() < 0.5 if sign is None else sign,
                rng.randint(-20, 20) if exp is None else exp, limbs)


@pytest.mark.parametrize("seed", range(12))
def test_mul_truncation_closed_form(seed):
    rng = random.Random(seed)
    a, b = rand_bf(rng, rng.randint(1, 60)), rand_bf(rng, rng.randint(1, 60))
    p = rng.choice([0, 1, 3, 10, 40])
    z = a.mul(b, p)
    pp = len(a.limbs) + len(b.limbs) if p == 0 else p + 2
    al, ae = chop(a.limbs, a.exp, pp)
    bl, be = chop(b.limbs, b.exp, pp)
    ref = Fraction(to_int(al) * to_int(bl)) * Fraction(2) ** (32 * (ae + be))
    if a.sign != b.sign:
        ref = -ref
    assert val(z) == ref
    assert z.exp == ae + be
    assert len(z.limbs) in (len(al) + len(bl), len(al) + len(bl) - 1)
    assert z.limbs[-1] != 0


def test_mul_zero_and_u32():
    z = O.BF()
    x = O.BF(True, 3, [5, 6])
    assert len(z.mul(x).limbs) == 0 and len(x.mul(z).limbs) == 0
    y = x.mul_u32(0xFFFFFFFF)
    assert val(y) == val(x) * 0xFFFFFFFF and y.exp == 3


@pytest.mark.parametrize("seed", range(16))
def test_add_sub_truncation_closed_form(seed):
    rng = random.Random(1000 + seed)
    a = rand_bf(rng, rng.randint(1, 30))
    b = rand_bf(rng, rng.randint(1, 30))
    p = rng.choice([0, 1, 4, 12])
    for op in ("add", "sub"):
        z = getattr(a, op)(b, p)
        top = max(a.exp + len(a.limbs), b.exp + len(b.limbs))
        bot = min(a.exp, b.exp)
        if p and top - bot > p + 2:
            bot = top - (p + 2)

        def tv(x):
            s, e, limbs = x.tuple()
            kept = sum(int(w) << (32 * (e + i - bot)) for i, w in enumerate(limbs) if e + i >= bot)
            v = Fraction(kept) * Fraction(2) ** (32 * bot)
            return v if s else -v

        ref = tv(a) + tv(b) if op == "add" else tv(a) - tv(b)
        assert val(z) == ref, (op, p)
        assert len(z.limbs) == 0 or z.limbs[-1] != 0 or op == "add"


def test_sign_rules():
    a = O.BF(True, 0, [5])
    b = O.BF(True, 0, [5])
    z = a.sub(b)
    assert len(z.limbs) == 0 and z.sign and z.exp == 0
    n = O.BF(False, 0, [7])
    assert val(a.add(n)) == -2 and not a.add(n).sign
    assert val(n.sub(a)) == -12


@pytest.mark.parametrize("p", [1, 2, 5, 33, 200])
def test_rcp_div_accuracy(p):
    rng = random.Random(p)
    x = rand_bf(rng, 50, exp=-10, sign=True)
    r = x.rcp(p)
    err = abs(val(r) * val(x) - 1)
    assert err < Fraction(2) ** (-32 * (p - 1)) * 4
    y = rand_bf(rng, 40, exp=3, sign=False)
    q = y.div(x, p)
    assert abs(val(q) / (val(y) / val(x)) - 1) < Fraction(2) ** (-32 * (p - 1)) * 8


@pytest.mark.parametrize("p", [1, 2, 7, 64])
def test_invsqrt_accuracy(p):
    r = O.invsqrt(10005, p)
    err = abs(val(r) ** 2 * 10005 - 1)
    assert err < Fraction(2) ** (-32 * (p - 1)) * 8


def test_divide_by_zero():
    with pytest.raises(ZeroDivisionError):
        O.BF().rcp(3)
    with pytest.raises(ZeroDivisionError):
        O.invsqrt(0, 3)


@pytest.mark.parametrize("D,hexd,p,terms", [(10000, 8304, 923, 706), (100000, 83048, 9228, 7052),
                                            (1000000, 830482, 92276, 70515)])
def test_pi_sizes(D, hexd, p, terms):
    assert O.pi_sizes(D) == (hexd, p, terms)


def pi_hex_mpmath(n):
    with mpmath.workdps(int(n * 1.21) + 30):
        x = mpmath.mp.pi - 3
        v = int(mpmath.floor(x * mpmath.mpf(16) ** n))
    return "3." + format(v, "0%dx" % n)


def bbp_digit(n):
    """Hex digit n+1 after the point of pi (Bailey-Borwein-Plouffe 1995)."""
    def S(j):
        s = 0.0
        for k in range(n + 1):
            r = 8 * k + j
            s = (s + pow(16, n - k, r) / r) % 1.0
        t, k = 0.0, n + 1
        while True:
            term = 16.0 ** (n - k) / (8 * k + j)
            if term < 1e-17:
                break
            t += term
            k += 1
        return (s + t) % 1.0
    x = (4 * S(1) - 2 * S(4) - S(5) - S(6)) % 1.0
    return "%x" % int(x * 16)


@pytest.mark.parametrize("D", [10000, 100000])
def test_pi_prefix_and_engines_agree(D):
    s_ref, _ = O.pi_hex(D, O.ENGINE_REFERENCE, threads=2)
    s_ntt, _ = O.pi_hex(D, O.ENGINE_NTT, threads=1)
    hexd, p, _ = O.pi_sizes(D)
    assert len(s_ref) == 2 + hexd
    assert s_ref == s_ntt  # engines give exact products
    good = 8 * (p - 4)  # digits inside the working precision
    assert s_ref[: 2 + good] == pi_hex_mpmath(good)
    for pos in (0, 1000, good - 20):
        assert s_ref[2 + pos] == bbp_digit(pos)


def test_pi_digest_fixtures():
    """SHA-256 of pi.txt (reference semantics) as committed by
    tests/golden/make_pi_digests.py; the GPU pipeline must match them too."""
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "pi_digests.json")
    with open(path) as f:
        fx = json.load(f)
    for D in ("10000", "100000"):
        s, _ = O.pi_hex(int(D), O.ENGINE_REFERENCE, threads=2)
        assert hashlib.sha256(s.encode()).hexdigest() == fx[D]["sha256"]
        assert s[:66] == fx[D]["head"]
    # every entry (10^4 ... 10^9): sizes of Pi(D) and the first 64 digits of pi
    head = pi_hex_mpmath(64)
    for D, e in fx.items():
        assert (e["hex_digits"], e["p_limbs"], e["terms"]) == O.pi_sizes(int(D))
        assert e["bytes"] == 2 + e["hex_digits"]
        assert e["head"] == head


def test_planted_truncation_mistake_is_caught():
    """Chopping p+1 instead of p+2 limbs (a plausible slip) changes pi."""
    rng = random.Random(9)
    a, b = rand_bf(rng, 30, sign=True), rand_bf(rng, 30, sign=True)
    z = a.mul(b, 5)
    al, ae = chop(a.limbs, a.exp, 6)
    bl, be = chop(b.limbs, b.exp, 6)
    wrong = Fraction(to_int(al) * to_int(bl)) * Fraction(2) ** (32 * (ae + be))
    assert val(z) != wrong


def test_bbp_extraction_matches_mpmath_and_python():
    """The C BBP extraction (oracle.bbp_hex) against mpmath's hex expansion of
    pi and the pure-Python digit extraction above."""
    with mpmath.workdps(3000):
        x = mpmath.mp.pi - 3
        s = format(int(mpmath.floor(x * mpmath.mpf(16) ** 2400)), "02400x")
    for n in (0, 1, 7, 100, 999, 2000, 2390):
        assert O.bbp_hex(n, threads=2) == s[n:n + 6]
    for n in (5000, 123457):
        assert O.bbp_hex(n, threads=3)[0] == bbp_digit(n)


def test_bbp_fixture_and_spot_check_helper():
    """tests/golden/bbp_spots.json agrees with the extraction (small positions)
    and with the oracle's pi.txt at 10^6 digits; pi_spot_checks accepts that
    pi.txt and rejects planted single-digit mistakes."""
    import json
    import os

    from tests.helpers import pi_correct_digits, pi_spot_checks

    path = os.path.join(os.path.dirname(__file__), "golden", "bbp_spots.json")
    with open(path) as f:
        spots = json.load(f)
    for k in ("0", "1000", "99990"):
        assert O.bbp_hex(int(k), threads=2) == spots[k]
    s, _ = O.pi_hex(1000000, O.ENGINE_REFERENCE, threads=os.cpu_count() or 1)
    good = pi_correct_digits(len(s) - 2)
    assert good == 738176
    for k, v in spots.items():
        n = int(k)
        if n + 6 <= good:
            assert s[2 + n:2 + n + 6] == v, n
    res = pi_spot_checks(s)
    assert res["ok"] and res["bbp_positions"] == [0, 1000, 99990, 738100]
    assert res["prefix_digits"] == [10000, 100000]

    def flip(text, i):
        c = "0" if text[i] != "0" else "1"
        return text[:i] + c + text[i + 1:]

    bad = pi_spot_checks(flip(s, 2 + 738103))
    assert not bad["bbp_ok"] and not bad["ok"]
    bad = pi_spot_checks(flip(s, 2 + 5000))
    assert bad["bbp_ok"] and not bad["prefix_ok"] and not bad["ok"]
