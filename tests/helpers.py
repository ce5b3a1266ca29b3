"""Shared checking helpers for the tests (no GPU needed)."""
from __future__ import annotations

import numpy as np

# Primes below 2^31 used for residue checks of big products.
CHECK_PRIMES = [2147483647, 2147483629, 2147483587, 2147483579]


def to_int(limbs) -> int:
    a = np.asarray(limbs, dtype=np.uint32)
    return int.from_bytes(a.astype("<u4").tobytes(), "little")


def from_int(x: int, n: int) -> np.ndarray:
    return np.frombuffer(x.to_bytes(4 * n, "little"), dtype="<u4").astype(np.uint32)


def exact_product(a, b) -> np.ndarray:
    """na+nb limb product by Python big-int arithmetic (small inputs)."""
    a = np.asarray(a, dtype=np.uint32)
    b = np.asarray(b, dtype=np.uint32)
    return from_int(to_int(a) * to_int(b), a.size + b.size)


def residue(limbs, q: int) -> int:
    """(sum limb_i 2^(32 i)) mod q, vectorised; q < 2^31."""
    x = np.asarray(limbs, dtype=np.uint32).astype(np.uint64)
    n = x.size
    if n == 0:
        return 0
    qq = np.uint64(q)
    pw = np.ones(1, dtype=np.uint64)
    base = (1 << 32) % q
    while pw.size < n:
        step = pow(base, pw.size, q)
        pw = np.concatenate([pw, (pw * np.uint64(step)) % qq])
    pw = pw[:n]
    terms = ((x % qq) * pw) % qq
    return int(terms.sum(dtype=np.uint64) % qq)


def residues_match(a, b, c) -> bool:
    """c == a * b modulo every check prime (a size-independent property)."""
    for q in CHECK_PRIMES:
        if residue(c, q) != residue(a, q) * residue(b, q) % q:
            return False
    return True


def pi_correct_digits(hex_digits: int) -> int:
    """Hex digits of Pi(D)'s output that are digits of pi: the computation
    carries p = (hex + 9) / 9 limbs (PI.h:141-144), 8 hex digits each, of
    which the last few absorb the truncation error."""
    return 8 * ((hex_digits + 9) // 9 - 4)


def pi_spot_checks(s: bytes | str) -> dict:
    """Checks of a pi.txt ("3." + hex digits) that need no run at the same
    size: BBP digits at the fixture positions inside the correct region,
    and the digest of the correct region of every oracle pi.txt fixture
    shorter than s."""
    import hashlib
    import json
    import os

    if isinstance(s, str):
        s = s.encode()
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    with open(os.path.join(gold, "bbp_spots.json")) as f:
        spots = json.load(f)
    with open(os.path.join(gold, "pi_digests.json")) as f:
        digests = json.load(f)
    good = pi_correct_digits(len(s) - 2)
    res = {"bbp_positions": [], "bbp_ok": True, "prefix_digits": [], "prefix_ok": True}
    for k, v in sorted(spots.items(), key=lambda kv: int(kv[0])):
        n = int(k)
        if n + 6 <= good:
            res["bbp_positions"].append(n)
            res["bbp_ok"] &= s[2 + n:2 + n + 6].decode() == v
    for k, fx in sorted(digests.items(), key=lambda kv: int(kv[0])):
        if "prefix_sha256" in fx and fx["bytes"] < len(s):
            res["prefix_digits"].append(int(k))
            pre = s[:2 + pi_correct_digits(fx["hex_digits"])]
            res["prefix_ok"] &= hashlib.sha256(pre).hexdigest() == fx["prefix_sha256"]
    res["ok"] = bool(res["bbp_ok"] and res["prefix_ok"] and res["bbp_positions"])
    return res


def limb_digest(limbs) -> str:
    """SHA-256 of the little-endian uint32 limbs (the mul_digests.json form)."""
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


_MUL_DIGESTS = None


def mul_digests() -> dict:
    """tests/golden/mul_digests.json: oracle product digests of the workloads."""
    global _MUL_DIGESTS
    if _MUL_DIGESTS is None:
        import json
        import os

        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mul_digests.json")
        with open(path) as f:
            _MUL_DIGESTS = json.load(f)
    return _MUL_DIGESTS


def single_product_digest(na: int, nb: int) -> str:
    return mul_digests()["single"]["%dx%d" % (na, nb)]["sha256"]


def oracle_threads() -> int:
    """Threads for oracle products inside tests: the CPU share of a GPU box
    is 16 cores whatever os.cpu_count() reports there."""
    import os

    return max(1, min(16, os.cpu_count() or 1))


def oracle_product(a, b) -> np.ndarray:
    """The oracle's exact product (its restatement of the reference NTT
    multiply), multi-threaded; for operands too large for Python integers."""
    from oracle import oracle as O

    return O.mul_limbs(a, b, O.ENGINE_NTT, oracle_threads())
