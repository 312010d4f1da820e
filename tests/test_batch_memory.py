"""Host logic of the batched operator's memory plan: grouping TODs so that
each group's filter work buffer fits a cap."""
import numpy as np
import pytest


def _check(work, cap, groups):
    assert [i for g in groups for i in g] == list(range(len(work)))  # consecutive, all, once
    assert all(g for g in groups)
    assert max(sum(work[i] for i in g) for g in groups) <= cap


def test_one_group_when_everything_fits():
    from minkasi_amd.mapmaking.batch import filter_groups
    assert filter_groups([5, 7, 3], 15) == [[0, 1, 2]]
    assert filter_groups([5, 7, 3], 100) == [[0, 1, 2]]
    assert filter_groups([], 10) == []


def test_cap_below_largest_tod_raises():
    from minkasi_amd.mapmaking.batch import filter_groups
    with pytest.raises(MemoryError):
        filter_groups([5, 9, 3], 8)


def test_groups_respect_cap_and_balance():
    from minkasi_amd.mapmaking.batch import filter_groups
    rng = np.random.default_rng(3)
    for trial in range(300):
        work = [int(v) for v in rng.integers(1, 1000, size=int(rng.integers(1, 300)))]
        cap = int(rng.integers(max(work), 2 * sum(work)))
        groups = filter_groups(work, cap)
        _check(work, cap, groups)
        # never more than twice the lower bound on the group count, plus one
        lower = -(-sum(work) // cap)
        assert len(groups) <= 2 * lower + 1, (len(groups), lower)
    # equal TODs split evenly
    g = filter_groups([10] * 12, 40)
    _check([10] * 12, 40, g)
    assert [len(x) for x in g] == [4, 4, 4]
    # the largest TOD alone fills a group
    g = filter_groups([1, 50, 1], 50)
    _check([1, 50, 1], 50, g)


def test_memory_plan_decisions():
    """The tiled gather's index copy is kept while the filter still gets at
    least MIN_FILTER_WORK (or its largest TOD's work), dropped before that,
    and too little memory for even the largest TOD's filter work raises."""
    from minkasi_amd.mapmaking.batch import MEM_RESERVE, MIN_FILTER_WORK, BatchedDot
    GB = 1 << 30
    work = [3 * GB, 2 * GB, 4 * GB] * 10          # 90 GiB of filter work in all, 4 GiB for the largest TOD
    nsamp = 8 * 10 ** 9                            # 32 GB of tiled-gather indices
    plan = BatchedDot._memory_plan
    tiled, cap = plan(work, nsamp, None, None, free=MEM_RESERVE + 4 * nsamp + 20 * GB)
    assert tiled and cap == 20 * GB
    tiled, cap = plan(work, nsamp, None, None, free=MEM_RESERVE + 4 * nsamp + MIN_FILTER_WORK - 1)
    assert not tiled and cap == 4 * nsamp + MIN_FILTER_WORK - 1
    tiled, cap = plan(work, nsamp, True, 5 * GB, free=MEM_RESERVE + 4 * nsamp + 20 * GB)
    assert tiled and cap == 5 * GB                 # the caller's cap wins when smaller
    tiled, cap = plan(work, nsamp, False, None, free=MEM_RESERVE + 200 * GB)
    assert not tiled and cap == 200 * GB
    with pytest.raises(MemoryError):
        plan(work, nsamp, None, None, free=MEM_RESERVE + 4 * GB - 1)
    with pytest.raises(MemoryError):               # forced tiled gather leaves too little
        plan(work, nsamp, True, None, free=MEM_RESERVE + 4 * nsamp + GB)
