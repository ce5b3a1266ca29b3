"""FP64 FFT engine (PCG_ENGINE_FFT): 16-bit balanced digits, right-angle
convolution, round-off report and NTT recompute above 0.25.  The product must
be bit-exact against the NTT engine and exact arithmetic in every case
(the reference's FFT.cpp path uses 8-bit digits and no check; FFT.cpp:485-571)."""
import random

import numpy as np
import pytest

from piculator_amd import workload
from tests.helpers import exact_product, limb_digest, oracle_product, single_product_digest

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF


@pytest.mark.parametrize("na,nb", [(1, 1), (5, 3), (300, 200), (1000, 1000), (4096, 4096),
                                   (20000, 777), (50000, 50000)])
def test_fft_exact_small(gpu, na, nb):
    rng = random.Random(na + 3 * nb)
    a = np.array([rng.getrandbits(32) for _ in range(na)], dtype=np.uint32)
    b = np.array([rng.getrandbits(32) for _ in range(nb)], dtype=np.uint32)
    got = gpu.mul(a, b, gpu.ENGINE_FFT)
    dist, recomputed = gpu.fft_last_roundoff()
    assert 0.0 <= dist
    assert np.array_equal(got, exact_product(a, b)), (dist, recomputed)


def test_fft_all_ones(gpu):
    a = np.full(3000, M32, dtype=np.uint32)
    got = gpu.mul(a, a, gpu.ENGINE_FFT)
    assert np.array_equal(got, exact_product(a, a))


@pytest.mark.parametrize("logn", [16, 20, 24, 26])
@pytest.mark.parametrize("unbalanced", [False, True])
def test_fft_roundoff_report_and_exactness(gpu, logn, unbalanced):
    """The single-multiply workload on the FFT engine (n = 2^16 .. 2^26,
    n x n and n x n/2): the round-off report is consistent with the
    recompute decision and every limb equals the oracle's product."""
    n = 1 << logn
    nb = n // 2 if unbalanced else n
    a, b = workload.operand_pair(n, nb)
    got = gpu.mul(a, b, gpu.ENGINE_FFT)
    dist, recomputed = gpu.fft_last_roundoff()
    assert 0.0 <= dist
    assert recomputed == (dist > 0.25)
    assert limb_digest(got) == single_product_digest(n, nb)


def test_fft_recompute_path_is_exercised(gpu):
    # Every balanced digit 32767: coefficients reach 2^21 * 2^30 = 2^51, so the
    # FP64 round-off exceeds 0.25 and the NTT must recompute the product.
    n = 1 << 20
    a = np.full(n, 0x7FFF7FFF, dtype=np.uint32)
    got = gpu.mul(a, a, gpu.ENGINE_FFT)
    dist, recomputed = gpu.fft_last_roundoff()
    assert recomputed and dist > 0.25
    assert np.array_equal(got, gpu.mul(a, a, gpu.ENGINE_NTT))
    assert np.array_equal(got, oracle_product(a, a))
