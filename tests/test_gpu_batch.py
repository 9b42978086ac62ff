"""Batched triangle counting: per-graph uint64 counts equal the oracle's; hipGraph replays equal
the eager launch, replay after replay."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from gary_amd import batch, graph, graph_query as gq

pytestmark = pytest.mark.gpu


def _oracle_counts(g, vbase, idx=None):
    idx = range(len(vbase) - 1) if idx is None else idx
    return {i: oracle.tri_count_range(g.indPtr, g.colIdx, int(vbase[i]), int(vbase[i + 1]))
            for i in idx}


@pytest.fixture(scope="module")
def small_batch():
    specs = batch.batch_specs(150, seed=0x5EED, n_range=(64, 3000), deg_range=(4, 40))
    g, vbase = batch.build_union(specs)
    return g, vbase, batch.Batch(g, vbase)


def test_batch_counts_equal_oracle(small_batch):
    g, vbase, b = small_batch
    want = _oracle_counts(g, vbase)
    eager_b = b
    eager_b.launch(use_graph=False)
    eager = eager_b.counts()
    assert all(int(eager[i]) == want[i] for i in want)
    assert int(eager.sum()) == oracle.tri_count(g.indPtr, g.colIdx)


def test_batch_graph_replays(small_batch):
    g, vbase, b = small_batch
    b.launch(use_graph=False)
    eager = b.counts().copy()
    for rep in range(5):
        b.launch(use_graph=True)
        got = b.counts()
        assert np.array_equal(got, eager), rep
        # eager work of the program between replays
        b.launch(use_graph=False)
        assert np.array_equal(b.counts(), eager), rep


def test_batch_cfg_changes_keep_counts(small_batch):
    g, vbase, b = small_batch
    b.launch(use_graph=False)
    eager = b.counts().copy()
    for bins in ((2, 2), (16, 512), (64, 2048)):
        b.set_cfg(gq.launch_cfg(grid=300, bin_small=bins[0], bin_hub=bins[1]))
        assert np.array_equal(b.count(), eager), bins
    b.set_cfg(None)


def test_batch_single_graph_and_empty_graphs():
    # graphs with no edges and one-vertex graphs inside a batch
    specs = [batch.GraphSpec("er", 5, 0, 1), batch.GraphSpec("er", 40, 300, 2),
             batch.GraphSpec("er", 1, 0, 3), batch.GraphSpec("rmat", 200, 1500, 4)]
    g, vbase = batch.build_union(specs)
    b = batch.Batch(g, vbase)
    want = _oracle_counts(g, vbase)
    got = b.count()
    assert [int(x) for x in got] == [want[i] for i in range(4)]


def test_batch_full_config_sampled():
    """The 2,000-graph workload: every graph's count against the GPU single-graph count of the
    union (sum), and a sample of 100 graphs against the oracle."""
    specs = batch.batch_specs(2000)
    g, vbase = batch.build_union(specs)
    b = batch.Batch(g, vbase)
    got = b.count()
    sample = list(range(0, 2000, 20))
    want = _oracle_counts(g, vbase, sample)
    assert all(int(got[i]) == want[i] for i in sample)
    total = gq.triangleCountHom(b.deviceGraph(), gq.BITMAP)
    assert int(got.sum()) == total
