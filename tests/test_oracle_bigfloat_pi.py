"""Pins the oracle's BigFloat semantics and Chudnovsky driver (no GPU).

Closed forms used (values are sum T[i] 2^(32 (exp + i)), sign-magnitude):
  * mul (BigFloat.cpp:326-444): the exact product of the operands each
    chopped to their top p+2 limbs (p = 0: no chop), one zero top limb trimmed;
  * add/sub (BigFloat.cpp:184-325): exact sum/difference of the operands
    with every word below top - (p+2) dropped;
  * rcp/div/invsqrt (BigFloat.cpp:445-567): the IEEE-double seeds at p = 0,
    the exponent and length the last sub of a Newton step fixes, an
    accuracy bound at the p + 2-limb working precision, and one Newton step
    rebuilt from the mul/sub closed forms above on the oracle's own
    half-precision result (the schedule s = p/2 + 1 read from
    BigFloat.cpp:502-504,553-555 -- pinned by this restatement only);
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


# ---- Newton steps rebuilt from the closed forms (exact rational arithmetic)
class X:
    """(sign, exp, magnitude m): value (-1)^(!sign) m 2^(32 exp); its limb
    count is that of m (every BigFloat result has a nonzero top limb)."""

    def __init__(self, sign, exp, m):
        self.sign, self.exp, self.m = (True, 0, 0) if m == 0 else (bool(sign), exp, m)

    @staticmethod
    def of(bf):
        s, e, limbs = bf.tuple()
        return X(s, e, to_int(limbs))

    @property
    def L(self):
        return (self.m.bit_length() + 31) // 32

    def tuple(self):
        return (self.sign, self.exp, [(self.m >> (32 * i)) & 0xFFFFFFFF for i in range(self.L)])


def x_chop(x, pp):
    """x cut to its top pp limbs (BigFloat.cpp:355-366)."""
    c = x.L - pp
    return X(x.sign, x.exp + c, x.m >> (32 * c)) if c > 0 else x


def x_mul(a, b, p):
    """BigFloat::mul closed form: exact product of the chopped operands."""
    if a.m == 0 or b.m == 0:
        return X(True, 0, 0)
    pp = a.L + b.L if p == 0 else p + 2
    a, b = x_chop(a, pp), x_chop(b, pp)
    return X(a.sign == b.sign, a.exp + b.exp, a.m * b.m)


def x_add(a, b, p):
    """BigFloat::add closed form: words below top - (p + 2) dropped."""
    if a.m == 0:
        return b
    if b.m == 0:
        return a
    top = max(a.exp + a.L, b.exp + b.L)
    bot = min(a.exp, b.exp)
    if p and top - bot > p + 2:
        bot = top - (p + 2)

    def aligned(x):
        v = (x.m >> (32 * (bot - x.exp))) if bot > x.exp else x.m << (32 * (x.exp - bot))
        return v if x.sign else -v

    v = aligned(a) + aligned(b)
    return X(v >= 0, bot, abs(v))


def x_sub(a, b, p):
    return x_add(a, X(not b.sign, b.exp, b.m), p)


def newton_half(p):
    """BigFloat.cpp:502-504 (rcp) and 553-555 (invsqrt)."""
    return 0 if p == 1 else (1 if p == 2 else p // 2 + 1)


def rcp_step(x, T, p):
    """r1 = r0 - (r0 x - 1) r0 at precision p (BigFloat.cpp:510)."""
    one = X(True, 0, 1)
    return x_sub(T, x_mul(x_sub(x_mul(x, T, p), one, p), T, p), p)


def invsqrt_step(xv, T, p):
    """r1 = r0 - (r0^2 x - 1) / 2 r0 at precision p (BigFloat.cpp:561-567)."""
    one = X(True, 0, 1)
    t = x_mul(T, T, p)
    t = x_mul(t, X(True, 0, xv), 0)
    t = x_sub(t, one, p)
    t = X(t.sign, t.exp - 1, t.m * 2147483648)
    t = x_mul(t, T, p)
    return x_sub(T, t, p)


NEWTON_P = list(range(1, 24)) + [33, 64, 100, 257]


def newton_operand(p):
    rng = random.Random(p)
    return rand_bf(rng, 50, exp=-10, sign=bool(p % 2))


def top_position(v: Fraction) -> int:
    """t with 2^(32 (t - 1)) <= |v| < 2^(32 t)."""
    v, t = abs(v), 0
    while v >= Fraction(2) ** (32 * t):
        t += 1
    while v < Fraction(2) ** (32 * (t - 1)):
        t -= 1
    return t


def test_newton_seeds_are_ieee_doubles():
    """p = 0 seeds (BigFloat.cpp:459-499, 527-551): 1/x of the top three
    limbs and 1/sqrt(x) in IEEE double, scaled by 2^32 until >= 2^32, kept
    as two limbs -- recomputed here with Python floats (the same IEEE
    operations, correctly rounded)."""
    import math

    for p in (1, 5, 33):
        x = newton_operand(p)
        r = x.rcp(0)
        limbs, e = x.limbs[-3:].tolist(), x.exp + max(0, len(x.limbs) - 3)
        v = float(limbs[0]) + limbs[1] * 4294967296. + limbs[2] * 4294967296. * 4294967296.
        v, ex = 1. / v, -e
        while v < 4294967296.:
            v *= 4294967296.
            ex -= 1
        w = int(v)
        assert r.tuple() == (x.sign, ex, [w & 0xFFFFFFFF, w >> 32])
    for xv in (1, 2, 10005, 0xFFFFFFFF):
        v, ex = 1. / math.sqrt(float(xv)), 0
        while v < 4294967296.:
            v *= 4294967296.
            ex -= 1
        w = int(v)
        assert O.invsqrt(xv, 0).tuple() == (True, ex, [w & 0xFFFFFFFF, w >> 32])


def _newton_invariants_hold(p) -> bool:
    """What the last sub of a Newton step fixes: the result's bottom is cut
    p + 2 limbs below the top of 1/x (so exp = top - (p + 2)), its length is
    p + 2 (+1 carry limb, -1 stripped zero), and |r x - 1| stays within a
    few units of 2^(-32 (p + 1)), the p + 2-limb working precision."""
    x = newton_operand(p)
    r = x.rcp(p)
    t = top_position(1 / val(x))
    ok = r.exp == t - (p + 2) and p + 1 <= len(r.limbs) <= p + 3
    ok &= abs(val(r) * val(x) - 1) < 4 * Fraction(2) ** (-32 * (p + 1))
    q = O.invsqrt(10005, p)
    ok &= q.exp == -(p + 2) and p + 1 <= len(q.limbs) <= p + 3  # 1/sqrt(10005) < 2^0
    ok &= abs(val(q) ** 2 * 10005 - 1) < 4 * Fraction(2) ** (-32 * (p + 1))
    return bool(ok)


@pytest.mark.parametrize("p", NEWTON_P)
def test_newton_result_invariants(p):
    assert _newton_invariants_hold(p)


@pytest.mark.parametrize("p", NEWTON_P)
def test_newton_step_from_closed_forms(p):
    """rcp(p) and invsqrt(p) equal one Newton step, computed with the
    mul/sub closed forms in exact arithmetic, on the oracle's own result at
    the half precision s = p/2 + 1: limb for limb, exponent and sign."""
    s = newton_half(p)
    x = newton_operand(p)
    assert x.rcp(p).tuple() == rcp_step(X.of(x), X.of(x.rcp(s)), p).tuple()
    assert O.invsqrt(10005, p).tuple() == invsqrt_step(10005, X.of(O.invsqrt(10005, s)), p).tuple()
    y = rand_bf(random.Random(p + 7), 40, exp=3, sign=False)
    assert y.div(x, p).tuple() == x_mul(X.of(y), X.of(x.rcp(p)), p).tuple()


def test_planted_newton_mistakes_are_caught():
    """A p + 1 chop in the last sub of each Newton step fails the
    invariants; the schedule s = p/2 passes every accuracy bound (its error
    is as small as the right schedule's) but fails the rebuilt step."""
    try:
        O.set_plant(1)
        assert not all(_newton_invariants_hold(p) for p in (3, 5, 33))
        O.set_plant(2)
        assert all(_newton_invariants_hold(p) for p in (5, 33))
        caught = False
        for p in (5, 33, 64):
            x = newton_operand(p)
            caught |= x.rcp(p).tuple() != rcp_step(X.of(x), X.of(x.rcp(newton_half(p))), p).tuple()
        assert caught
    finally:
        O.set_plant(0)


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
