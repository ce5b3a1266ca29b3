"""Benchmark of pi to 10^9 decimal digits on one MI355X, with its big-number
multiply hot path measured against the HBM and FP64-mulmod rooflines.

Step = one computation of pi.txt for --digits decimal digits (default 10^9)
through the batched schedule: Chudnovsky binary splitting with the lower tree
evaluated on the GPU (pcg_bsr_exact), the top combines, division, 1/sqrt(10005)
and the final product on the GPU BigFloat, and the hex conversion.  W untimed
steps, then K timed steps, each bracketed by device synchronisation; `value`
is the mean wall time of a step.  The output of every timed step is checked
outside its timed window (SHA-256 against the oracle's pi.txt digest where a
fixture exists, BBP digits, the prefix digest of the oracle's 10^8-digit
pi.txt); if any check fails, `value` is null and the exit status is 1.

Also reported on the same JSON line:
  roofline      the 2^24 x 2^24-limb NTT multiply (operands resident in HBM):
                algorithmic bytes / time vs 8 TB/s, time from HIP events on
                the library stream; traffic from the committed counter profile
  mulmod        the same multiply's algorithmic mulmods per second against
                the FP64 mulmod rate measured by a register-resident probe
  ntt_mul       that multiply's stage times and full-product digest check
  fft_mul       the same operands on the FP64 FFT engine: time, HBM
                fraction, maximum round-off, NTT recompute, digest check
  batch         4096 uneven products: one batched call vs one launch sequence
                per product
  pi_small      pi to 10^6 digits, batched vs plain schedule, both checked
                against the oracle digest
  cpu_baseline  the oracle (CPU restatement of the reference, reference engine
                choice) timed in this run on the box's CPU share: pi to 10^8
                digits beside the GPU's own 10^8 run, scaled to --digits by the
                ratio of full oracle runs, and the oracle's 2^24 x 2^24
                multiply (NTT at k = 27, BigFloat.cpp:417-441) beside the GPU's
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


def pi_cost(digits: int) -> float:
    """Cost model of the binary-splitting pi computation, N log^2 N in the
    hex digit count N (multiply M(N) ~ N log N, log N tree levels)."""
    n = digits * math.log(10) / math.log(16)
    return n * math.log2(n) ** 2


def cpu_share():
    """(threads, how): the reference's default thread count is
    std::thread::hardware_concurrency() (main.cpp:6), the CPUs this process
    may run on: its affinity mask, capped by a cgroup CPU quota and by
    OMP_NUM_THREADS (the GPU box pins both to its CPU share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    how = ["affinity %d" % n]
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            quota, period = f.read().split()[:2]
        if quota != "max":
            q = max(1, int(int(quota) / int(period)))
            how.append("cgroup quota %d" % q)
            n = min(n, q)
    except (OSError, ValueError):
        pass
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        how.append("OMP_NUM_THREADS %s" % omp)
        n = min(n, int(omp))
    how.append("os.cpu_count %s" % os.cpu_count())
    return max(1, n), ", ".join(how)


def cpu_baseline(threads: int, how: str, digits: int, sample_digits: int, gpu_sample_s, logn: int,
                 gpu_mul_ms: float):
    """The oracle's Pi() (reference engine crossover: FFT below k = 27, NTT
    above, BigFloat.cpp:386) on `sample_digits`, scaled to `digits`, and the
    oracle's 2^logn x 2^logn multiply on the reference engine."""
    from oracle import oracle as O
    from piculator_amd import workload
    from tests.helpers import limb_digest, single_product_digest

    t0 = time.perf_counter()
    s, _ = O.pi_hex(sample_digits, O.ENGINE_REFERENCE, threads=threads)
    dt = time.perf_counter() - t0
    ok = None
    with open(os.path.join(ROOT, "tests", "golden", "pi_digests.json")) as f:
        fx = json.load(f)
    if str(sample_digits) in fx:
        ok = hashlib.sha256(s.encode()).hexdigest() == fx[str(sample_digits)]["sha256"]
    del s
    with open(os.path.join(ROOT, "tests", "golden", "oracle_pi_times.json")) as f:
        secs = json.load(f)["seconds"]
    if str(digits) in secs and str(sample_digits) in secs:
        scale = secs[str(digits)] / secs[str(sample_digits)]
        scale_how = "the ratio of full oracle runs at both sizes (%.0f s / %.1f s, 8 threads)" % (
            secs[str(digits)], secs[str(sample_digits)])
    else:
        scale = pi_cost(digits) / pi_cost(sample_digits)
        scale_how = "N log^2 N in hex digits"
    n = 1 << logn
    a, b = workload.operand_pair(n)
    t0 = time.perf_counter()
    z = O.mul_limbs(a, b, O.ENGINE_REFERENCE, threads)
    mul_s = time.perf_counter() - t0
    mul_ok = limb_digest(z) == single_product_digest(n, n)
    del z, a, b
    res = {
        "value": dt * scale,
        "unit": "s (pi to %d digits; measured at %d, scaled)" % (digits, sample_digits),
        "cores": threads,
        "kind": "port",
        "sample": "oracle Pi(%d), reference engines, %d threads (%s): %.2f s measured in this run, "
                  "digest %s; x%.1f to %d digits by %s"
                  % (sample_digits, threads, how, dt, "ok" if ok else ("MISMATCH" if ok is False else "n/a"),
                     scale, digits, scale_how),
        "sample_s": dt,
        "sample_digits": sample_digits,
        "sample_digest_ok": ok,
        "mul": {"na": n, "nb": n, "engine": "reference (NTT at k = 27, BigFloat.cpp:417-441)",
                "threads": threads, "s": mul_s, "digest_ok": bool(mul_ok),
                "gpu_ms": gpu_mul_ms, "gpu_speedup": mul_s * 1e3 / gpu_mul_ms},
    }
    if gpu_sample_s:
        res["gpu_sample_s"] = gpu_sample_s
        res["same_size_speedup"] = dt / gpu_sample_s
    return res


def probe_rates():
    """FP64 mulmod and FMA rates of the whole GPU (libpcg_probe.so)."""
    import ctypes as C

    lib = C.CDLL(os.path.join(ROOT, "piculator_amd", "libpcg_probe.so"))
    mm, fma, cu = C.c_double(), C.c_double(), C.c_int()
    if lib.pcg_probe_fp64_rates(C.byref(mm), C.byref(fma), C.byref(cu)) != 0:
        raise RuntimeError("mulmod probe failed")
    return {"mulmods_per_s": mm.value, "fma_per_s": fma.value, "compute_units": cu.value,
            "probe": "register-resident chains of the engine's FP64 Shoup mulmod (6 FP64 ops) and of "
                     "FP64 FMAs, 16 workgroups of 256 threads per CU (piculator_amd/csrc/probe_mulmod.hip)"}


STREAM_PATTERNS = ((0, 0), (0, 8), (8, 0), (8, 8), (0, 16), (16, 0), (16, 16), (32, 32), (4, 4))


def stream_probe():
    """Markdown table of the HBM access-pattern probe (libpcg_probe.so,
    piculator_amd/csrc/probe_stream.hip): a 2^25-point, two-prime tile copy
    with the large-tile passes' thread shape, each side contiguous (row
    tiles) or strided with C adjacent columns (runs of 8 C bytes)."""
    import ctypes as C

    lib = C.CDLL(os.path.join(ROOT, "piculator_amd", "libpcg_probe.so"))
    side = lambda c: "contiguous" if c == 0 else "strided, %d columns (%d-byte runs)" % (c, 8 * c)
    print("| read side | write side | us | TB/s (read + write) |")
    print("|---|---|---|---|")
    for rc, wc in STREAM_PATTERNS:
        us, bps = C.c_double(), C.c_double()
        if lib.pcg_probe_stream(rc, wc, C.byref(us), C.byref(bps)) != 0:
            raise RuntimeError("stream probe failed for pattern %d -> %d" % (rc, wc))
        print("| %s | %s | %.1f | %.2f |" % (side(rc), side(wc), us.value, bps.value / 1e12), flush=True)


def ntt_mul(logn: int, steps: int, warmup: int):
    """The 2^logn x 2^logn NTT multiply through pcg_mul_dev: mean time from HIP
    events over `steps` back-to-back multiplies, stage profile, product check."""
    import numpy as np

    from piculator_amd import picugpu, workload
    from tests.helpers import limb_digest, single_product_digest

    n = 1 << logn
    a, b = workload.operand_pair(n)
    da, db = picugpu.DeviceBuffer(a.nbytes), picugpu.DeviceBuffer(b.nbytes)
    dout = picugpu.DeviceBuffer(a.nbytes + b.nbytes)
    da.upload(a)
    db.upload(b)
    picugpu.reserve(2 * n)
    logN, passes = picugpu.ntt_plan(2 * n)
    log("multiply %d x %d limbs: 2^%d points, %d passes" % (n, n, logN, passes))
    for _ in range(warmup):
        picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_NTT, sync=True)
    picugpu.sync()
    picugpu.timer_start()
    for _ in range(steps):
        picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_NTT)
    ms = picugpu.timer_stop() / steps
    picugpu.sync()
    picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_NTT, sync=True, profile=True)
    stages, total = picugpu.last_profile()
    out = np.empty(2 * n, dtype=np.uint32)
    dout.download(out)
    ok = limb_digest(out) == single_product_digest(n, n)
    del out
    for buf in (da, db, dout):
        buf.free()
    balg = workload.bytes_alg(n, n)
    achieved = balg / (ms / 1e3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "ntt_mul_2p%d_traffic.json" % logn)
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_multiply")
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    info = {"na": n, "nb": n, "transform_points": 1 << logN, "passes": passes, "steps": steps,
            "ms": ms, "algorithmic_bytes": balg, "product_check": ok,
            "stage_ms": {"forward_a": stages[0], "forward_b_pointwise_inverse": stages[1],
                         "crt_carry": stages[2], "total": total}}
    log("multiply: %.3f ms, %.1f GB/s algorithmic, check %s" % (ms, achieved, ok))
    return roofline, info


def fft_mul(logn: int, steps: int):
    """The same operands on the FP64 FFT engine (PCG_ENGINE_FFT): mean time
    from HIP events, round-off report, full-product digest."""
    import numpy as np

    from piculator_amd import picugpu, workload
    from tests.helpers import limb_digest, single_product_digest

    n = 1 << logn
    a, b = workload.operand_pair(n)
    da, db = picugpu.DeviceBuffer(a.nbytes), picugpu.DeviceBuffer(b.nbytes)
    dout = picugpu.DeviceBuffer(a.nbytes + b.nbytes)
    da.upload(a)
    db.upload(b)
    picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_FFT, sync=True)
    picugpu.timer_start()
    for _ in range(steps):
        # each FFT product reads its round-off back (the recompute decision)
        picugpu.mul_dev(da.ptr, n, db.ptr, n, dout.ptr, picugpu.ENGINE_FFT, sync=True)
    ms = picugpu.timer_stop() / steps
    dist, recomputed = picugpu.fft_last_roundoff()
    out = np.empty(2 * n, dtype=np.uint32)
    dout.download(out)
    ok = limb_digest(out) == single_product_digest(n, n)
    del out
    for buf in (da, db, dout):
        buf.free()
    achieved = workload.bytes_alg(n, n) / (ms / 1e3) / 1e9
    res = {"na": n, "nb": n, "ms": ms, "gbs_alg": achieved, "hbm_frac": achieved / HBM_PEAK_GBS,
           "max_roundoff": dist, "ntt_recompute": bool(recomputed), "digest_ok": bool(ok),
           "digits": "16-bit balanced, right-angle convolution in complex FP64"}
    log("fft multiply: %.3f ms, round-off %.4f, check %s" % (ms, dist, ok))
    return res


def batch_size_sweep(total_limbs: int = 1 << 25, reps: int = 3):
    """One JSON line per transform size 2^12 .. 2^24: one pcg_mul_batch_dev
    call of equal device-resident products with na = nb = 2^(s-1) - 1 limbs
    (the product fills 2^s points), as many as make `total_limbs` operand
    limbs, best of `reps` after one untimed call; the algorithmic bytes are
    those of the multiply roofline, 64 N_alg + 8 (na + nb) per product."""
    import numpy as np

    from piculator_amd import picugpu, workload

    picugpu.init(0)
    lib = picugpu.load()
    for s in range(12, 25):
        n = (1 << (s - 1)) - 1
        count = max(1, total_limbs // (2 * n))
        total = count * 2 * n
        pool = picugpu.DeviceBuffer(8 * total + 64)
        host = np.zeros(2 * total, dtype=np.uint32)
        host[:total] = np.random.default_rng(s).integers(0, 2**32, total, dtype=np.uint64).astype(np.uint32)
        pool.upload(host)
        descs = (picugpu.MulDesc * count)()
        for i in range(count):
            descs[i] = picugpu.MulDesc(pool.ptr + 4 * (2 * n * i), n, pool.ptr + 4 * (2 * n * i + n), n,
                                       pool.ptr + 4 * (total + 2 * n * i))
        picugpu._check(lib.pcg_mul_batch_dev(descs, count, 0))
        best = None
        for _ in range(reps):
            picugpu.timer_start()
            picugpu._check(lib.pcg_mul_batch_dev(descs, count, 0))
            ms = picugpu.timer_stop()
            best = ms if best is None else min(best, ms)
        alg = count * workload.bytes_alg(n, n)
        print(json.dumps({"points": 1 << s, "limbs": n, "products": count, "ms": round(best, 4),
                          "alg_GBs": round(alg / best / 1e6, 1), "hbm_frac": round(alg / best / 1e6 / 8000, 3)}),
              flush=True)
        pool.free()


def class_sweep(count: int = 2048, reps: int = 3):
    """One JSON line per small-product bucket b = 6..12: the time of one
    pcg_mul_batch_dev call of `count` device-resident products of
    na = nb = 3 * 2^(b-3) limbs in every class valid for the bucket (forced
    through the class table: schoolbook 0..3, Karatsuba 4, single-workgroup
    NTT 10..12, batched pass kernels 110..112), best of `reps` after one."""
    import numpy as np

    from piculator_amd import picugpu

    picugpu.init(0)
    saved = picugpu.batch_table()
    lib = picugpu.load()
    school_max = {0: 1024, 1: 64, 2: 256, 3: 512, 4: 1024}
    try:
        for b in range(6, 13):
            n = 3 << (b - 3)
            total = count * 2 * n
            pool = picugpu.DeviceBuffer(8 * total + 64)
            host = np.zeros(2 * total, dtype=np.uint32)
            host[:total] = np.random.default_rng(b).integers(0, 2**32, total, dtype=np.uint64).astype(np.uint32)
            pool.upload(host)
            descs = (picugpu.MulDesc * count)()
            for i in range(count):
                descs[i] = picugpu.MulDesc(pool.ptr + 4 * (2 * n * i), n, pool.ptr + 4 * (2 * n * i + n), n,
                                           pool.ptr + 4 * (total + 2 * n * i))
            res = {}
            lg = max(b, 10)
            for cls in (0, 1, 2, 3, 4, lg, 100 + lg):
                if cls in school_max and (1 << b) > school_max[cls]:
                    continue
                t = list(saved)
                t[b] = cls
                picugpu.batch_table(t)
                picugpu._check(lib.pcg_mul_batch_dev(descs, count, 0))
                best = None
                for _ in range(reps):
                    picugpu.timer_start()
                    picugpu._check(lib.pcg_mul_batch_dev(descs, count, 0))
                    ms = picugpu.timer_stop()
                    best = ms if best is None else min(best, ms)
                res[str(cls)] = best
            print(json.dumps({"bucket": b, "limbs": n, "products": count, "ms": res,
                              "tuned_class": saved[b]}), flush=True)
            pool.free()
    finally:
        picugpu.batch_table(saved)


def batch_compare(count: int = 4096, reps: int = 3):
    """Uneven batch (lengths log-uniform in 2^4..2^14 limbs, seed 0xC0DE),
    operands resident in HBM: one pcg_mul_batch_dev call vs the plain form
    (one launch sequence per product, PCG_BATCH_PLAIN)."""
    import numpy as np

    from piculator_amd import picugpu, workload
    from tests.helpers import limb_digest, mul_digests

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
    want = mul_digests()["batch"]["sha256"] if count == 4096 else None
    ok = want is not None
    for i in range(count if want else 0):
        po, n = outs[i]
        ok &= limb_digest(host[po:po + n]) == want[i]
    res["all_products_digest_ok"] = bool(ok)
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
    from tests.helpers import limb_digest, single_product_digest

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
                rec["exact"] = limb_digest(out) == single_product_digest(n, nb)
                print(json.dumps(rec), flush=True)
            for buf in (da, db, dout):
                buf.free()


def _rocpd(dirname: str):
    """sqlite connection to the rocpd database rocprofv3 wrote under dirname."""
    import glob
    import sqlite3

    dbs = glob.glob(os.path.join(dirname, "**", "*.db"), recursive=True)
    if not dbs:
        raise SystemExit("no rocpd database under %s" % dirname)
    return sqlite3.connect(dbs[0])


def summarize_trace(dirname: str):
    """Markdown table of kernel durations from a rocprofv3 --kernel-trace run."""
    con = _rocpd(dirname)
    rows = con.execute("select name, count(*), avg(duration), min(duration), max(duration), "
                       "sum(duration) from kernels group by name order by sum(duration) desc").fetchall()
    tot = sum(r[5] for r in rows)
    print("| kernel | calls | avg us | min us | max us | total % |")
    print("|---|---|---|---|---|---|")
    for n, c, a, mn, mx, sm in rows:
        n = n.replace("(pcg::PassArgs)", "").replace("void ", "")
        print("| `%s` | %d | %.1f | %.1f | %.1f | %.1f |" % (n, c, a / 1e3, mn / 1e3, mx / 1e3,
                                                          100.0 * sm / tot))
    print("total kernel time %.3f ms" % (tot / 1e6))


def summarize_traffic(fetch_dir: str, write_dir: str):
    """HBM bytes per multiply from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch) of `bench.py --mul-only --calib-copy`, each
    scaled by what the same counter reads for the 512 MiB calibration copy."""
    out = {}
    for key, d, counter in (("read", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        con = _rocpd(d)
        rows = con.execute("select kernel_name, sum(value) from counters_collection where counter_name = ? "
                           "group by dispatch_id, kernel_name", (counter,)).fetchall()
        copies = [v for n, v in rows if "copyBuffer" in n]
        calib = max(copies) if copies else None
        scale = (512 * 1024) / calib if calib else 1.0
        mults = sum(1 for n, _ in rows if "k_carry_fix" in n)
        per_kernel = {}
        for n, v in rows:
            if "pcg::" not in n:  # the calibration copy and its setup
                continue
            name = n.split("(")[0].replace("void ", "")
            per_kernel[name] = per_kernel.get(name, 0.0) + v * 1024 * scale / max(mults, 1)
        out[key] = {"scale": scale, "multiplies": mults, "bytes_per_multiply": sum(per_kernel.values()),
                    "per_kernel_bytes": per_kernel}
    out["hbm_bytes_per_multiply"] = out["read"]["bytes_per_multiply"] + out["write"]["bytes_per_multiply"]
    print(json.dumps(out, indent=1, sort_keys=True))


def summarize_counters(dirname: str):
    """Per-kernel averages per dispatch of every counter in a rocprofv3
    --pmc run, with the ratios used for the compute-side evidence: VALU
    instructions per wave-cycle (x4 waves per SIMD for the issue share),
    wait share, LDS bank-conflict share."""
    con = _rocpd(dirname)
    rows = con.execute("select kernel_name, counter_name, dispatch_id, sum(value) from counters_collection "
                       "group by dispatch_id, kernel_name, counter_name").fetchall()
    acc, disp = {}, {}
    for kn, cn, d, v in rows:
        k = kn.split("(")[0].replace("void ", "")
        acc.setdefault(k, {}).setdefault(cn, 0.0)
        acc[k][cn] += v
        disp.setdefault(k, set()).add(d)
    out = {}
    for k, m in acc.items():
        n = len(disp[k])
        e = {cn: v / n for cn, v in m.items()}
        wc = e.get("SQ_WAVE_CYCLES")
        if wc:
            if "SQ_ACTIVE_INST_VALU" in e:
                e["valu_per_wave_cycle"] = e["SQ_ACTIVE_INST_VALU"] / wc
            if "SQ_WAIT_INST_ANY" in e:
                e["wait_inst_share"] = e["SQ_WAIT_INST_ANY"] / wc
        if e.get("SQ_LDS_IDX_ACTIVE") and "SQ_LDS_BANK_CONFLICT" in e:
            e["lds_conflict_share"] = e["SQ_LDS_BANK_CONFLICT"] / e["SQ_LDS_IDX_ACTIVE"]
        e["dispatches"] = n
        out[k] = e
    print(json.dumps(out, indent=1, sort_keys=True))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed pi computations")
    ap.add_argument("--warmup", type=int, default=1, help="untimed pi computations")
    ap.add_argument("--digits", type=int, default=1000000000)
    ap.add_argument("--logn", type=int, default=24, help="roofline multiply size 2^logn limbs")
    ap.add_argument("--mul-steps", type=int, default=20)
    ap.add_argument("--small-digits", type=int, default=1000000)
    ap.add_argument("--no-small", action="store_true", help="skip the batched-vs-plain pi run")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain pi schedule")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-sample-digits", type=int, default=100000000)
    ap.add_argument("--no-fft", action="store_true", help="skip the FFT-engine multiply line")
    ap.add_argument("--no-probe", action="store_true", help="skip the FP64 mulmod peak probe")
    ap.add_argument("--graph", action="store_true",
                    help="replay captured HIP graphs of the subtree level schedules (default: kernel by kernel)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="first run one device copy of 512 MiB (counter calibration)")
    ap.add_argument("--fft-only", action="store_true",
                    help="only the FFT-engine multiply line (the operands of --logn), for profiling")
    ap.add_argument("--mul-only", action="store_true",
                    help="only the roofline multiply (for counter runs)")
    ap.add_argument("--sweep", action="store_true",
                    help="print the single-multiply table (sizes x shapes x engines) instead")
    ap.add_argument("--class-sweep", action="store_true",
                    help="print the time of every small-product class per bucket instead")
    ap.add_argument("--batch-sizes", action="store_true",
                    help="print the batched-product throughput per transform size instead")
    ap.add_argument("--stream-probe", action="store_true",
                    help="print the HBM access-pattern probe table instead")
    ap.add_argument("--summarize-trace", metavar="DIR",
                    help="print the kernel table of a rocprofv3 --kernel-trace run (no GPU)")
    ap.add_argument("--summarize-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="print HBM bytes per multiply from two rocprofv3 --pmc runs (no GPU)")
    ap.add_argument("--summarize-counters", metavar="DIR",
                    help="print per-kernel counter averages of a rocprofv3 --pmc run (no GPU)")
    args = ap.parse_args()
    if args.summarize_counters:
        summarize_counters(args.summarize_counters)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace)
        return
    if args.summarize_traffic:
        summarize_traffic(*args.summarize_traffic)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path "
              "(the binary-splitting tree would shard by subtrees, one GPU each)", file=sys.stderr)
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
        sweep(args.mul_steps)
        return
    if args.stream_probe:
        stream_probe()
        return
    if args.class_sweep:
        class_sweep()
        return
    if args.batch_sizes:
        batch_size_sweep()
        return

    from piculator_amd import picugpu, pihost
    from tests.helpers import pi_spot_checks

    picugpu.init(0)
    picugpu.bsr_graphs(args.graph)
    if args.fft_only:
        print(json.dumps({"fft_mul": fft_mul(args.logn, args.mul_steps)}), flush=True)
        return
    roofline, mul_info = ntt_mul(args.logn, args.mul_steps, 3)
    rates = None
    if not args.no_probe:
        rates = probe_rates()
        log("fp64 probe: %.3g mulmods/s, %.3g FMA/s" % (rates["mulmods_per_s"], rates["fma_per_s"]))
    mulmod = None
    if rates:
        from piculator_amd import workload

        mm = workload.mulmods_alg(1 << args.logn, 1 << args.logn) / (mul_info["ms"] / 1e3)
        mulmod = {"bound": "fp64 mulmod", "achieved": mm, "peak": rates["mulmods_per_s"], "unit": "mulmod/s",
                  "frac": mm / rates["mulmods_per_s"],
                  "algorithmic_mulmods": workload.mulmods_alg(1 << args.logn, 1 << args.logn),
                  "fp64_fma_per_s": rates["fma_per_s"], "probe": rates["probe"]}
    if args.mul_only:
        print(json.dumps({"roofline": roofline, "mulmod": mulmod, "ntt_mul": mul_info}), flush=True)
        return

    D = args.digits
    with open(os.path.join(ROOT, "tests", "golden", "pi_digests.json")) as f:
        fx = json.load(f)
    buf = pihost.pi_buffer(D)
    cold = None
    for i in range(args.warmup):
        log("warm-up pi %d digits (%d/%d)..." % (D, i + 1, args.warmup))
        t0 = time.perf_counter()
        n, t = pihost.pi_hex_into(D, pihost.MODE_BATCHED, buf)
        if cold is None:
            cold = time.perf_counter() - t0
        log("  %.3f s" % t["total"])
    log("timed: %d x pi %d digits..." % (args.steps, D))
    phases, walls, digests = [], [], []
    for _ in range(args.steps):
        picugpu.sync()
        t0 = time.perf_counter()
        n, t = pihost.pi_hex_into(D, pihost.MODE_BATCHED, buf)
        picugpu.sync()
        walls.append(time.perf_counter() - t0)
        phases.append(t)
        digests.append(hashlib.sha256(buf.raw[:n]).hexdigest())  # outside the timed window
    wall = sum(walls) / len(walls)
    log("  %.3f s per pi computation" % wall)
    s = buf.raw[:n]
    del buf
    check = pi_spot_checks(s)
    check["sha256"] = digests[-1]
    check["every_step_same_digest"] = len(set(digests)) == 1
    check["ok"] = bool(check["ok"] and check["every_step_same_digest"])
    if str(D) in fx:
        check["full_digest_ok"] = all(d == fx[str(D)]["sha256"] for d in digests)
        check["ok"] = bool(check["ok"] and check["full_digest_ok"])
    check["head"] = s[:34].decode()
    del s
    log("pi check:", json.dumps(check))
    hexd, p, terms = pihost.pi_sizes(D)
    last = phases[-1]
    res = {
        "metric": METRIC,
        "value": wall,
        "unit": "s (wall time of one pi computation to %d decimal digits)" % D,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "pi to %d decimal digits (Chudnovsky binary splitting, batched "
                               "schedule, every big multiply on the GPU)" % D,
                   "digits": D, "hex_digits": hexd, "terms": terms, "precision_limbs": p,
                   "roofline_kernel": "2^%d x 2^%d-limb two-prime FP64 NTT multiply"
                                      % (args.logn, args.logn),
                   "subtree_schedule": "captured HIP graph" if args.graph else "kernel by kernel"},
        "roofline": roofline,
        "mulmod": mulmod,
        "step_s": walls,
        "cold_first_run_s": cold,
        "cold_note": "the first (warm-up) run in this process: includes the small-product autotuner "
                     "and the first plan layouts; the 2^%d multiply ran before it" % args.logn,
        "phases_s": {k: last[k] for k in ("series", "division", "invsqrt", "final_mul",
                                          "conversion", "total")},
        "big_multiplies": int(last["big_muls"]),
        "pi_check": check,
        "ntt_mul": mul_info,
    }
    if not args.no_fft:
        res["fft_mul"] = fft_mul(args.logn, max(2, args.mul_steps // 4))
    if not args.no_batch:
        log("uneven batch of 4096 products...")
        res["batch"] = batch_compare()
        log("batch:", json.dumps(res["batch"]))
    if not args.no_small:
        log("pi %d digits, batched vs plain..." % args.small_digits)
        res["pi_small"] = pi_compare(args.small_digits, not args.no_plain)
        log("pi_small:", json.dumps(res["pi_small"]))
    if not args.no_cpu:
        th, how = cpu_share()
        if args.cpu_threads:
            th, how = args.cpu_threads, "--cpu-threads"
        sd = args.cpu_sample_digits
        gpu_sample = None
        if sd <= D:
            s2, t2 = pihost.pi_hex(sd, pihost.MODE_BATCHED)
            gpu_sample = t2["total"]
            del s2
        log("cpu baseline (%d threads: %s)..." % (th, how))
        res["cpu_baseline"] = cpu_baseline(th, how, D, sd, gpu_sample, args.logn, mul_info["ms"])
        log("cpu baseline:", res["cpu_baseline"]["sample"])
    ok = check["ok"] and mul_info["product_check"]
    if not ok:
        res["value"] = None
        res["ms_per_step"] = None
        res["invalid"] = "a result check failed (pi_check / ntt_mul.product_check): no timing is reported"
    print(json.dumps(res), flush=True)
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
