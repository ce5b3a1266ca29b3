"""Benchmark of the big-number multiply hot path on one MI355X.

Step = one exact 2^24 x 2^24-limb multiply (operands SplitMix64, radix 2^32,
resident in HBM) through the C ABI (pcg_mul_dev), the size at which the
measurement plan gates the NTT multiply against the HBM roofline.  The timed
region is K back-to-back multiplies on the library stream, bracketed by HIP
events recorded on that stream and by device synchronisation.

Also reported on the same JSON line:
  roofline      algorithmic bytes of the multiply / time vs 8 TB/s
  cpu_baseline  the oracle (CPU restatement of the reference NTT multiply,
                NTT.cpp with its 128/64 divide) on a bounded 2^22 x 2^22
                sample, scaled by N log N to the 2^24 workload
  pi            pi to --pi-digits (default 10^6) on the GPU BigFloat, batched
                and plain schedules, checked against the oracle digest
Prints exactly one JSON line on stdout; progress goes to stderr.
"""
import argparse
import hashlib
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pi 1e9-digit wall s, batched vs plain GPU (bit-exact); NTT mul %HBM roofline"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(threads: int, logn_sample: int = 22, logn_target: int = 24):
    """Oracle NTT multiply of a 2^s x 2^s sample, scaled to 2^t x 2^t by N log N
    of the reference transform length N = 2^ceil(log2(4 (na + nb)))."""
    from oracle import oracle as O
    from piculator_amd import workload

    n = 1 << logn_sample
    a, b = workload.operand_pair(n)
    t0 = time.perf_counter()
    O.mul_limbs(a, b, O.ENGINE_NTT, threads=threads)
    dt = time.perf_counter() - t0
    ks = math.ceil(math.log2(4 * 2 * n))
    kt = math.ceil(math.log2(4 * 2 * (1 << logn_target)))
    scale = (2 ** kt * kt) / (2 ** ks * ks)
    t_target = dt * scale
    nt = 1 << logn_target
    return {
        "value": workload.bytes_alg(nt, nt) / t_target / 1e9,
        "unit": "GB/s (algorithmic bytes of one 2^24 x 2^24 multiply / time)",
        "cores": threads,
        "kind": "port",
        "sample": "oracle NTT-forced multiply (63-bit prime, 128/64 divide, std::async-style "
                  "thread split) of 2^%d x 2^%d SplitMix64 limbs: %.2f s, scaled x%.2f by "
                  "N log N (N = 2^%d -> 2^%d) to %.1f s per 2^24 multiply"
                  % (logn_sample, logn_sample, dt, scale, ks, kt, t_target),
        "seconds_per_step": t_target,
    }


def batch_compare(count: int = 4096, reps: int = 3):
    """Uneven batch (lengths log-uniform in 2^4..2^14 limbs, seed 0xC0DE),
    operands resident in HBM: one pcg_mul_batch_dev call vs the plain form
    (one launch sequence per product, PCG_BATCH_PLAIN)."""
    import numpy as np

    from piculator_amd import picugpu, workload
    from tests.helpers import residues_match

    pairs = workload.batch_operands(count)
    total = sum(a.size + b.size for a, b in pairs)
    pool = picugpu.DeviceBuffer(8 * total + 64)
    descs = (picugpu.MulDesc * count)()
    host = np.zeros(2 * total, dtype=np.uint32)
    off = 0
    outs = []
    for i, (a, b) in enumerate(pairs):
        pa, pb = off, off + a.size
        host[pa:pa + a.size] = a
        host[pb:pb + b.size] = b
        po = total + off
        descs[i] = picugpu.MulDesc(pool.ptr + 4 * pa, a.size, pool.ptr + 4 * pb, b.size,
                                   pool.ptr + 4 * po)
        outs.append((po, a.size + b.size))
        off += a.size + b.size
    pool.upload(host)
    lib = picugpu.load()
    res = {"products": count, "operand_limbs": int(total)}
    for name, flags in (("batched", 0), ("plain", picugpu.BATCH_PLAIN)):
        picugpu._check(lib.pcg_mul_batch_dev(descs, count, flags))  # warm-up
        best = None
        for _ in range(reps):
            picugpu.timer_start()
            picugpu._check(lib.pcg_mul_batch_dev(descs, count, flags))
            ms = picugpu.timer_stop()
            best = ms if best is None else min(best, ms)
        res[name + "_ms"] = best
    res["speedup"] = res["plain_ms"] / res["batched_ms"]
    pool.download(host)
    ok = True
    for i in range(0, count, 97):
        a, b = pairs[i]
        po, n = outs[i]
        ok &= bool(residues_match(a, b, host[po:po + n]))
    res["checked_products_ok"] = ok
    return res


def pi_compare(digits: int, plain: bool):
    from piculator_amd import pihost

    res = {"digits": digits}
    s, t = pihost.pi_hex(digits, pihost.MODE_BATCHED)
    res["batched_s"] = t["total"]
    res["batched_phases_s"] = {k: t[k] for k in ("series", "division", "invsqrt", "final_mul",
                                                  "conversion")}
    res["big_multiplies"] = int(t["big_muls"])
    digest = hashlib.sha256(s.encode()).hexdigest()
    res["sha256"] = digest
    gold = os.path.join(ROOT, "tests", "golden", "pi_digests.json")
    with open(gold) as f:
        fx = json.load(f)
    if str(digits) in fx:
        res["matches_oracle"] = digest == fx[str(digits)]["sha256"]
    if plain:
        s2, t2 = pihost.pi_hex(digits, pihost.MODE_PLAIN)
        res["plain_s"] = t2["total"]
        res["plain_equals_batched"] = s2 == s
        res["batched_speedup"] = t2["total"] / t["total"]
    return res


def sweep(steps: int):
    """One JSON line per (size, shape, engine) of the single-multiply table:
    n = 2^16, 2^20, 2^24, 2^26 limbs, n x n and n x n/2, NTT and FFT engines."""
    import numpy as np

    from piculator_amd import picugpu, workload
    from tests.helpers import residues_match

    picugpu.init(0)
    for logn in (16, 20, 24, 26):
        n = 1 << logn
        for nb in (n, n // 2):
            a, b = workload.operand_pair(n, nb)
            da, db = picugpu.DeviceBuffer(a.nbytes), picugpu.DeviceBuffer(b.nbytes)
            dout = picugpu.DeviceBuffer(a.nbytes + b.nbytes)
            da.upload(a)
            db.upload(b)
            picugpu.reserve(n + nb)
            for name, eng in (("ntt", picugpu.ENGINE_NTT), ("fft", picugpu.ENGINE_FFT)):
                picugpu.mul_dev(da.ptr, n, db.ptr, nb, dout.ptr, eng, sync=True)
                k = steps if logn < 26 else max(2, steps // 4)
                picugpu.timer_start()
                for _ in range(k):
                    picugpu.mul_dev(da.ptr, n, db.ptr, nb, dout.ptr, eng, sync=(eng == picugpu.ENGINE_FFT))
                ms = picugpu.timer_stop() / k
                rec = {"na": n, "nb": nb, "engine": name, "ms": ms}
                if eng == picugpu.ENGINE_FFT:
                    dist, recomputed = picugpu.fft_last_roundoff()
                    rec["max_roundoff"] = dist
                    rec["ntt_recompute"] = recomputed
                balg = workload.bytes_alg(n, nb)
                rec["gbs_alg"] = balg / (ms / 1e3) / 1e9
                rec["hbm_frac"] = rec["gbs_alg"] / HBM_PEAK_GBS
                rec["mulmods_per_s"] = workload.mulmods_alg(n, nb) / (ms / 1e3)
                out = np.empty(n + nb, dtype=np.uint32)
                dout.download(out)
                rec["exact"] = bool(residues_match(a, b, out))
                print(json.dumps(rec), flush=True)
            for buf in (da, db, dout):
                buf.free()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--logn", type=int, default=24, help="operand size 2^logn limbs")
    ap.add_argument("--pi-digits", type=int, default=1000000)
    ap.add_argument("--no-pi", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain pi schedule")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run one device copy of 512 MiB (counter calibration)")
    ap.add_argument("--sweep", action="store_true",
                    help="print the single-multiply table (sizes x shapes x engines) instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(the sharding design is described in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        # a plain copy of known size (512 MiB read + 512 MiB written), so that
        # FETCH_SIZE / WRITE_SIZE of the multiply kernels can be scaled by what
        # the same counters read for it
        import torch

        src = torch.ones(1 << 27, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.sweep:
        sweep(args.steps)
        return

    import numpy as np

    from piculator_amd import picugpu, workload
    from tests.helpers import residues_match

    picugpu.init(0)
    n = 1 << args.logn
    a, b = workload.operand_pair(n)
    da, db = picugpu.DeviceBuffer(a.nbytes), picugpu.DeviceBuffer(b.nbytes)
    dout = picugpu.DeviceBuffer(a.nbytes + b.nbytes)
    da.upload(a)
    db.upload(b)
    picugpu.reserve(2 * n)
    logN, passes = picugpu.ntt_plan(2 * n)
    log("multiply %d x %d limbs: 2^%d points, %d passes" % (n, n, logN, passes))
    for _ in range(args.warmup):
        picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_NTT, sync=True)
    picugpu.sync()
    t0 = time.perf_counter()
    picugpu.timer_start()
    for _ in range(args.steps):
        picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_NTT)
    ev_ms = picugpu.timer_stop()
    picugpu.sync()
    wall_ms = (time.perf_counter() - t0) * 1e3
    ms = ev_ms / args.steps
    log("timed %d steps: %.3f ms/step (events), wall %.3f ms" % (args.steps, ms, wall_ms / args.steps))
    picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_NTT, sync=True, profile=True)
    stages, total = picugpu.last_profile()
    out = np.empty(2 * n, dtype=np.uint32)
    dout.download(out)
    ok = bool(residues_match(a, b, out)) and int(out[0]) == (int(a[0]) * int(b[0])) & 0xFFFFFFFF
    log("product check (residues mod 4 primes):", ok)

    balg = workload.bytes_alg(n, n)
    achieved = balg / (ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_mul_2p%d_traffic.json" % args.logn)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_multiply")
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}

    res = {
        "metric": METRIC,
        "value": achieved,
        "unit": "GB/s (algorithmic bytes 64 N + 8 (na + nb) of one 2^%d x 2^%d multiply / time)"
                % (args.logn, args.logn),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "single exact big-integer multiply, 2^%d x 2^%d limbs of radix 2^32 "
                               "(SplitMix64 operands resident in HBM)" % (args.logn, args.logn),
                   "na": n, "nb": n, "engine": "two-prime FP64 NTT + CRT + carry scan",
                   "transform_points": 1 << logN, "passes": passes},
        "roofline": roofline,
        "product_check": ok,
        "stage_ms": {"forward_a": stages[0], "forward_b_pointwise_inverse": stages[1],
                     "crt_carry": stages[2], "total": total},
    }
    if not args.no_batch:
        log("uneven batch of 4096 products...")
        res["batch"] = batch_compare()
        log("batch:", json.dumps(res["batch"]))
    if not args.no_cpu:
        th = args.cpu_threads or min(16, os.cpu_count() or 1)
        log("cpu baseline (%d threads)..." % th)
        res["cpu_baseline"] = cpu_baseline(th)
        log("cpu baseline:", res["cpu_baseline"]["sample"])
    if not args.no_pi:
        log("pi %d digits..." % args.pi_digits)
        res["pi"] = pi_compare(args.pi_digits, not args.no_plain)
        log("pi:", json.dumps(res["pi"]))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
