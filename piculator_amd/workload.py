"""Synthetic operands for the multiply benchmarks and parity tests.

Operands are SplitMix64 streams (Steele, Lea & Flood 2014): one 64-bit output
per limb, low 32 bits kept, top limb forced nonzero.  Seeds follow the
measurement plan: operand A of n limbs uses 0xA000 + log2(n), operand B of nb
limbs uses 0xB000 + log2(nb).
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    """First n outputs of SplitMix64 started at `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        i = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def limbs(seed: int, n: int) -> np.ndarray:
    """n random uint32 limbs, little-endian, top limb nonzero."""
    x = (splitmix64(seed, n) & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    if n and x[-1] == 0:
        x[-1] = 1
    return x


def operand_pair(n: int, nb: int | None = None):
    """The (A, B) operand pair of the single-multiply workload: n x nb limbs."""
    nb = n if nb is None else nb
    a = limbs(0xA000 + int(math.log2(n)), n)
    b = limbs(0xB000 + int(math.log2(nb)), nb)
    return a, b


def batch_lengths(count: int = 4096, seed: int = 0xC0DE):
    """(na, nb) pairs of the uneven batch workload: floor(2^(4 + 10u)) limbs,
    u = next()/2^64, so lengths are log-uniform in [2^4, 2^14]."""
    z = splitmix64(seed, 2 * count)
    u = z.astype(np.float64) / 18446744073709551616.0
    L = np.floor(np.exp2(4.0 + 10.0 * u)).astype(np.int64)
    return list(zip(L[0::2].tolist(), L[1::2].tolist()))


def batch_operands(count: int = 4096, seed: int = 0xC0DE):
    """Operands of the batch workload: product i uses seeds (seed, i) mixed."""
    pairs = []
    for i, (na, nb) in enumerate(batch_lengths(count, seed)):
        a = limbs((seed << 20) + 2 * i, na)
        b = limbs((seed << 20) + 2 * i + 1, nb)
        pairs.append((a, b))
    return pairs


def bytes_alg(na: int, nb: int) -> int:
    """Algorithmic HBM bytes of one multiply: 64 N + 8 (na + nb) with
    N = 2^ceil(log2(2 (na + nb))) 16-bit-digit points stored as 8-byte residues:
    two read+write passes per transform, three transforms, pointwise fused."""
    N = 1 << math.ceil(math.log2(2 * (na + nb)))
    return 64 * N + 8 * (na + nb)


def mulmods_alg(na: int, nb: int) -> int:
    """Algorithmic modular multiplies: 3 (N/2) log2 N + 2 N."""
    N = 1 << math.ceil(math.log2(2 * (na + nb)))
    return 3 * (N // 2) * int(math.log2(N)) + 2 * N
