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
