"""Device-resident binary splitting (pcg_bsr_exact) against the oracle's
restatement of Pi_BSR (PI.h:64-112): bit-identical P, Q, R for subtrees whose
values fit the precision window, and the size error otherwise."""
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

BIG_P = 1 << 28  # no truncation anywhere


@pytest.mark.parametrize("a,b", [(0, 1), (0, 2), (1, 3), (0, 7), (0, 100), (37, 1000),
                                 (5000, 13000), (0, 70515)])
def test_subtree_matches_oracle(gpu, a, b):
    got = gpu.bsr_exact(a, b, BIG_P)
    P, Q, R = O.pi_bsr(a, b, BIG_P)
    assert got[0] == P.tuple()
    assert got[1] == Q.tuple()
    assert got[2] == R.tuple()


def test_large_leaf_indices(gpu):
    # leaves near 7e7 (pi to 1e9 digits) exercise the uint32 products 6b-5, 6b-1
    a, b = 70_000_000, 70_002_000
    got = gpu.bsr_exact(a, b, BIG_P)
    P, Q, R = O.pi_bsr(a, b, BIG_P)
    assert got == [P.tuple(), Q.tuple(), R.tuple()]


@pytest.mark.parametrize("a,b", [(715827870, 715827901), (1431655700, 1431655800)])
def test_wrapping_leaf_factors(gpu, a, b):
    # 6b-5 and 6b-1 wrap mod 2^32 from b = 715827883 (twice from
    # 1431655766): leaves and two-leaf nodes keep the reference's uint32
    # factors.  (Ranges stay below a + b = 2^32, where PI.h:90's uint32
    # midpoint itself wraps.)
    got = gpu.bsr_exact(a, b, BIG_P)
    P, Q, R = O.pi_bsr(a, b, BIG_P)
    assert got == [P.tuple(), Q.tuple(), R.tuple()]


def test_captured_schedule_replays(gpu):
    """Subtrees of one plan (same size, same leaf bounds near 7e7) run as
    replays of the plan's captured level schedule, with other products and
    synchronisations between the replays: every replay equals the
    kernel-by-kernel evaluation, and that equals the oracle."""
    import numpy as np

    d = 20000
    starts = [70_000_000 + k * d for k in range(3)]
    try:
        gpu.bsr_graphs(False)
        want = [gpu.bsr_exact(a, a + d, BIG_P) for a in starts]
        gpu.bsr_graphs(True)
        other = (np.arange(1, 5000, dtype=np.uint32), np.arange(7, 3007, dtype=np.uint32))
        for _ in range(2):
            for a, w in zip(starts, want):
                assert gpu.bsr_exact(a, a + d, BIG_P) == w, a
                gpu.mul(*other)
    finally:
        gpu.bsr_graphs(False)
    P, Q, R = O.pi_bsr(starts[1], starts[1] + d, BIG_P)
    assert want[1] == [P.tuple(), Q.tuple(), R.tuple()]


def test_window_too_small_is_refused(gpu):
    with pytest.raises(gpu.PcgError) as e:
        gpu.bsr_exact(0, 5000, 100)
    assert e.value.code == 5  # PCG_ERR_TOO_LARGE


@pytest.mark.parametrize("a,b", [(0, 1), (0, 100), (5000, 13000)])
def test_device_resident_results(gpu, a, b):
    """pcg_df_bsr_exact leaves the same values in device memory (exponent 0)."""
    got = gpu.df_bsr_exact(a, b, BIG_P)
    P, Q, R = O.pi_bsr(a, b, BIG_P)
    assert got == [P.tuple(), Q.tuple(), R.tuple()]


def test_prepared_plans(gpu):
    """Plans built ahead (pcg_bsr_prepare, laid out on the GPU) give the same
    values; a size named twice, a third size (beyond the two cached plans)
    or one never evaluated is harmless."""
    gpu.bsr_prepare([8000, 8001, 8000, 123])
    for a, b in [(5000, 13000), (40, 8041)]:
        got = gpu.bsr_exact(a, b, BIG_P)
        P, Q, R = O.pi_bsr(a, b, BIG_P)
        assert got == [P.tuple(), Q.tuple(), R.tuple()]
    gpu.bsr_prepare([])
