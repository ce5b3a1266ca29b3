"""Writes tests/golden/pi_digests.json: SHA-256 and head of pi.txt as the
oracle (CPU restatement of the reference, reference engine crossover)
produces it for Pi(D) (PI.h:138-210).  Run from the repository root:
    python tests/golden/make_pi_digests.py [D ...]
The GPU pipeline must reproduce these bytes exactly.  prefix_sha256 covers
"3." and the first 8 (p - 4) hex digits, the part that is digits of pi
(the computation carries p = (hex + 9) / 9 limbs, PI.h:141-144), so a run
at more digits must reproduce it too."""
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pi_digests.json")


def main(digits):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for D in digits:
        t = time.time()
        s, times = O.pi_hex(D, O.ENGINE_REFERENCE, threads=os.cpu_count() or 1)
        hexd, p, terms = O.pi_sizes(D)
        data[str(D)] = {
            "sha256": hashlib.sha256(s.encode()).hexdigest(),
            "head": s[:66],
            "tail": s[-32:],
            "prefix_sha256": hashlib.sha256(s[:2 + 8 * (p - 4)].encode()).hexdigest(),
            "bytes": len(s),
            "hex_digits": hexd,
            "p_limbs": p,
            "terms": terms,
        }
        print(D, data[str(D)]["sha256"], "%.2fs" % (time.time() - t), flush=True)
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(x) for x in sys.argv[1:]] or [10000, 100000, 1000000])
