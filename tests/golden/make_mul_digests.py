"""Writes tests/golden/mul_digests.json: SHA-256 digests of exact products
of the multiply workloads, as the oracle's restatement of the reference
multiply (BigFloat.cpp:377-441 with NTT.cpp / FFT.cpp) computes them on the
CPU.  Run from the repository root (about three minutes on 8 cores):
    python tests/golden/make_mul_digests.py

A digest is the SHA-256 of the untrimmed na + nb little-endian uint32 limbs
of the product (the array BigFloat::mul fills before its one-limb trim,
BigFloat.cpp:411-412,436-437).  Entries:
  * "single": the SplitMix64 operand pairs of the single-multiply workload
    (piculator_amd/workload.operand_pair), n = 2^16, 2^20, 2^24, 2^26 limbs,
    n x n and n x n/2.  Every product is computed by the oracle's NTT
    (16-bit digits, one 63-bit prime, NTT.cpp) and, up to 2^24 limbs, also
    by its FFT (8-bit digits, FP64 complex, FFT.cpp): two independent
    engines that must agree limb for limb before a digest is written.  Up
    to 2^20 limbs Python's own big-integer product is a third check.
  * "batch": the 4096 products of the uneven batch workload (seed 0xC0DE,
    workload.batch_operands), one digest per product, each checked against
    Python's big-integer product.
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from piculator_amd import workload  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mul_digests.json")
THREADS = os.cpu_count() or 1


def digest(limbs: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(limbs, dtype="<u4").tobytes()).hexdigest()


def py_product(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    x = int.from_bytes(a.astype("<u4").tobytes(), "little")
    y = int.from_bytes(b.astype("<u4").tobytes(), "little")
    return np.frombuffer((x * y).to_bytes(4 * (a.size + b.size), "little"), dtype="<u4")


def single_entries():
    out = {}
    for logn in (16, 20, 24, 26):
        n = 1 << logn
        for nb in (n, n // 2):
            a, b = workload.operand_pair(n, nb)
            t = time.time()
            ntt = O.mul_limbs(a, b, O.ENGINE_NTT, THREADS)
            checks = ["oracle NTT (NTT.cpp)"]
            if logn <= 24:
                fft = O.mul_limbs(a, b, O.ENGINE_FFT, THREADS)
                if not np.array_equal(ntt, fft):
                    raise SystemExit("oracle NTT and FFT disagree at %d x %d" % (n, nb))
                checks.append("oracle FFT (FFT.cpp)")
                del fft
            if logn <= 20:
                if not np.array_equal(ntt, py_product(a, b)):
                    raise SystemExit("oracle and Python products disagree at %d x %d" % (n, nb))
                checks.append("Python int")
            key = "%dx%d" % (n, nb)
            out[key] = {"na": n, "nb": nb, "sha256": digest(ntt), "agree": checks,
                        "low_limb": int(ntt[0]), "top_limb": int(ntt[-1])}
            print(key, out[key]["sha256"][:16], "%.1fs" % (time.time() - t), flush=True)
            del ntt
    return out


def batch_entry():
    pairs = workload.batch_operands(4096, 0xC0DE)
    ds = []
    for a, b in pairs:
        z = O.mul_limbs(a, b, O.ENGINE_REFERENCE, 1)
        if not np.array_equal(z, py_product(a, b)):
            raise SystemExit("oracle and Python products disagree in the batch")
        ds.append(digest(z))
    return {"count": 4096, "seed": 0xC0DE, "sha256": ds,
            "agree": ["oracle reference engine (BigFloat.cpp:386)", "Python int"]}


def main():
    data = {"format": "sha256 of the untrimmed na+nb little-endian uint32 product limbs",
            "single": single_entries(), "batch": batch_entry()}
    with open(OUT, "w") as f:
        json.dump(data, f, indent=0, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
