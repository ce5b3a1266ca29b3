"""pcg_mul_dev_noalloc (the entry point for callers that cannot allocate)
against pcg_mul_dev through the C ABI: the same limbs on two calls in a row,
the second smaller (the static scratch keeps the first call's transforms),
and the size limit.  Exact against Python integers too."""
import numpy as np
import pytest

from tests.helpers import exact_product

pytestmark = pytest.mark.gpu


def _run(gpu, fn, a, b):
    da, db = gpu.DeviceBuffer(a.nbytes), gpu.DeviceBuffer(b.nbytes)
    do = gpu.DeviceBuffer(a.nbytes + b.nbytes)
    try:
        da.upload(a)
        db.upload(b)
        fn(da.ptr, a.size, db.ptr, b.size, do.ptr, gpu.ENGINE_NTT, sync=True)
        out = np.empty(a.size + b.size, dtype=np.uint32)
        do.download(out)
        return out
    finally:
        for x in (da, db, do):
            x.free()


@pytest.mark.parametrize("sizes", [((600000, 448576), (3000, 1234)),
                                   ((70000, 65535), (1, 77)),
                                   ((1 << 19, 1 << 19), (100000, 100000))])
def test_noalloc_matches_mul_dev_twice(gpu, sizes):
    print("noalloc sizes", sizes, flush=True)
    rng = np.random.default_rng(sizes[0][0])
    for k, (na, nb) in enumerate(sizes):
        a = rng.integers(0, 2**32, na, dtype=np.uint64).astype(np.uint32)
        b = rng.integers(0, 2**32, nb, dtype=np.uint64).astype(np.uint32)
        if k == 1:
            a[:] = 0xFFFFFFFF  # longest carry chains on the second, smaller call
        got = _run(gpu, gpu.mul_dev_noalloc, a, b)
        want = _run(gpu, gpu.mul_dev, a, b)
        assert np.array_equal(got, want), (na, nb)
        if na + nb <= 20000:
            assert np.array_equal(got, exact_product(a, b))


def test_noalloc_limit_and_engine(gpu):
    L = gpu.load()
    n = gpu.NOALLOC_MAX_LIMBS
    # beyond the limit: PCG_ERR_TOO_LARGE before anything is launched (no
    # device memory is touched, so null-free dummy addresses are enough)
    rc = L.pcg_mul_dev_noalloc(16, n, 16, 1, 16, gpu.ENGINE_NTT)
    assert rc == 5
    assert b"static scratch limit" in L.pcg_last_error_string()
    rc = L.pcg_mul_dev_noalloc(16, 10, 16, 10, 16, gpu.ENGINE_FFT)
    assert rc == 1
