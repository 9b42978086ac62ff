"""The block-scope queue (the reference's CTAQueue, include/ops/DeviceQueue.cuh:33-182) on the
device, driven as the reference's queue test drives it (test/CTACommTest.cu:91-119: a filled
queue, lane 0 of each wave dequeuing) and with producers and consumers at once.

What the queue must do, checked without any oracle: every item comes out exactly once; in
exclusive mode each wave takes items in queue (FIFO) order, so its sequence is a subsequence of
the fill order; in concurrent mode items pass through a 128-slot ring many times over (wrap
around) and the counts add up."""
from __future__ import annotations

import numpy as np
import pytest

from gary_amd import _lib, ops


@pytest.mark.gpu
@pytest.mark.parametrize("n", [0, 1, 37, 128])
def test_exclusive_fill_then_dequeue(n):
    rng = np.random.default_rng(n)
    items = rng.permutation(1000)[:n].astype(np.uint32)
    out, status = ops.block_queue_exercise(ops.QUEUE_EXCLUSIVE, items)
    assert status.min() >= 0 and int(status.sum()) == n
    taken = []
    rank = {int(x): i for i, x in enumerate(items)}
    for w in range(ops.QUEUE_WAVES):
        seq = out[w * ops.QUEUE_SLOTS: w * ops.QUEUE_SLOTS + status[w]].tolist()
        order = [rank[x] for x in seq]
        assert order == sorted(order), f"wave {w} took items out of queue order"
        taken += seq
    assert sorted(taken) == sorted(items.tolist())


@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 128, 129, 5000, 200_000])
def test_concurrent_producers_consumers(n):
    items = np.random.default_rng(7).permutation(n).astype(np.uint32)
    out, status = ops.block_queue_exercise(ops.QUEUE_CONCURRENT, items)
    assert status.min() >= 0, status
    half = ops.QUEUE_WAVES // 2
    assert int(status[:half].sum()) == n  # enqueued
    assert int(status[half:].sum()) == n  # dequeued
    assert np.array_equal(out, np.ones(n, np.uint32))  # each item exactly once


def test_exercise_argument_errors():
    """Rejected on the host before any device work."""
    lib = _lib.lib
    rc = lib.gary_block_queue_exercise(2, None, 0, 1, 1, None)
    assert rc == _lib.GARY_ERR_INVALID and "mode" in _lib.last_error()
    rc = lib.gary_block_queue_exercise(0, 1, 129, 1, 1, None)
    assert rc == _lib.GARY_ERR_INVALID and "at most 128" in _lib.last_error()
    rc = lib.gary_block_queue_exercise(1, None, 5, 1, 1, None)
    assert rc == _lib.GARY_ERR_INVALID
