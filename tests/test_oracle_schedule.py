"""The one-by-one pool schedule is recoverable from per-path lengths (CPU, oracle).

A path's evolution depends only on its (pixel, sample) sampler stream, so the
number of rounds it stays in the pool (finish - start) is the same whatever
pool it runs in. The batched renderer records these lengths and rebuilds the
reference's schedule for the job's pool: ids [0, pool) start in round 0 and the
paths finishing in round r hand their slots, in id order, to the next ids,
which start in round r + 1 (wavefront_path.cu:265-285, :641-700). This test
restates that reconstruction in Python and checks it against the oracle's own
rounds at several pool sizes.
"""
import numpy as np
import pytest

import oracle as O
import scenes


def schedule(lengths, pool):
    """Finishing round of every path id for a pool of `pool` slots, from path lengths."""
    n = len(lengths)
    finish = np.zeros(n, np.int64)
    ends = {}
    lo, hi, r = 0, min(pool, n), 0
    while True:
        f = r + lengths[lo:hi].astype(np.int64)
        finish[lo:hi] = f
        for v, c in zip(*np.unique(f, return_counts=True)):
            ends[int(v)] = ends.get(int(v), 0) + int(c)
        t = ends.pop(r, 0)
        if hi == n:
            break
        lo, hi, r = hi, min(hi + t, n), r + 1
    return finish


@pytest.mark.parametrize("name", ["cornell", "soup"])
def test_lengths_do_not_depend_on_the_pool(name):
    if name == "cornell":
        sc = scenes.cornell34(24, 20, spp=4, max_depth=6)
    else:
        sc = scenes.soup_room(3000, 20, 16, spp=4, max_depth=4)
    runs = {}
    for pool in (97, 1000, 1 << 21):
        start, finish, st = O.render_paths(sc, pool_size=pool, threads=8)
        assert (finish >= start).all()
        runs[pool] = (start, finish, st)
    lengths = runs[1 << 21][1] - runs[1 << 21][0]
    assert (lengths <= sc["max_depth"] + 1).all()
    for pool, (start, finish, st) in runs.items():
        assert np.array_equal(finish - start, lengths), pool
        rebuilt = schedule(lengths, pool)
        assert np.array_equal(rebuilt, finish), pool
        assert int(rebuilt.max()) + 1 == st["rounds"], pool


def test_schedule_closed_forms():
    # a pool at least as large as the job: every path starts in round 0
    ln = np.array([3, 0, 5, 1, 1], np.uint32)
    assert np.array_equal(schedule(ln, 8), ln)
    # one slot: paths run back to back, each starting the round after the previous ends
    assert np.array_equal(schedule(ln, 1), np.array([3, 4, 10, 12, 14]))
    # two slots, all of length 0: two paths finish per round
    assert np.array_equal(schedule(np.zeros(6, np.uint32), 2), np.array([0, 0, 1, 1, 2, 2]))
