"""Writes tests/golden/bbp_spots.json: hexadecimal digits of pi at chosen
positions by Bailey-Borwein-Plouffe digit extraction (oracle.bbp_hex), which
is independent of the binary-splitting pipeline.

Pi(D) prints hex = floor(D log 10 / log 16) hex digits but works with
p = (hex + 9) / 9 limbs (PI.h:141-144), i.e. about 8 p hex digits: only the
first 8 (p - 4) digits are digits of pi, the rest is the deterministic tail
of the truncated arithmetic.  Positions therefore sit inside that correct
region, up to its end for D = 10^6 .. 10^9 (738206216 digits at 10^9), so a
10^9 run can be spot-checked where no full CPU run is affordable.  Run from
the repository root:
    python tests/golden/make_bbp_spots.py
Key n holds the 6 hex digits from digit n + 1 after the point."""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bbp_spots.json")
POSITIONS = [0, 1000, 99990, 738100, 7382000, 73820500, 250000000, 500000000,
             700000000, 738206000]


def main():
    data = {}
    for n in POSITIONS:
        t = time.time()
        data[str(n)] = O.bbp_hex(n, threads=os.cpu_count() or 1)
        print(n, data[str(n)], "%.1fs" % (time.time() - t), flush=True)
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)


if __name__ == "__main__":
    main()
