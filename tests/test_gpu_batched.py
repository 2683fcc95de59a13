"""GPU parity of the batched forms: 64-source bitmask BFS against per-source
CPU BFS runs (bit-exact)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
G = pytest.importorskip("graptor_amd")


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()
    yield


def seeded_sources(index, k, seed):
    """k sources drawn from the non-isolated vertices, fixed seed."""
    cand = np.flatnonzero(np.diff(index) > 0)
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=k, replace=len(cand) < k).astype(np.uint32)


@pytest.mark.parametrize("reorder", [False, True])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_msbfs_64_sources_rmat(reorder, d):
    index, nbrs = O.rmat_graph(14)
    src = seeded_sources(index, 64, 1)
    with G.Graph.from_csr(index, nbrs) as g:
        if reorder:
            g.reorder()
        lv, info = g.msbfs(src, direction=d)
    ref = O.msbfs_levels(index, nbrs, src)
    assert np.array_equal(lv, ref)
    n = len(index) - 1
    deg = np.diff(index)
    assert info.edges == int(sum(deg[ref[j] <= n].sum() for j in range(len(src))))


def test_msbfs_odd_counts_duplicates_and_components():
    index, nbrs = O.build_csr(10, np.array([0, 1, 5, 6], np.uint32), np.array([1, 2, 6, 7], np.uint32))
    with G.Graph.from_csr(index, nbrs) as g:
        for src in ([0], [0, 5], [2, 2, 9], list(range(10)) * 6):
            lv, _ = g.msbfs(np.array(src, np.uint32))
            assert np.array_equal(lv, O.msbfs_levels(index, nbrs, src)), src
        with pytest.raises(G.GprError):
            g.msbfs(np.arange(65, dtype=np.uint32) % 10)


def test_msbfs_grid_high_diameter():
    index, nbrs = O.grid_graph(64, 64, 0.0)
    src = seeded_sources(index, 17, 3)
    with G.Graph.from_csr(index, nbrs) as g:
        lv, info = g.msbfs(src)
    assert np.array_equal(lv, O.msbfs_levels(index, nbrs, src))
    assert info.iters >= 64


# ---- batched CC ---------------------------------------------------------------------
@pytest.mark.parametrize("count,nmin,nmax", [(64, 64, 4096), (9, 1000, 3000), (3, 70000, 90000)])
def test_batch_generate_and_cc(count, nmin, nmax):
    seed = 1234
    ref_index, ref_nbrs, ref_voff = O.batch_graphs(count, nmin, nmax, seed)
    with G.Batch.generate(count, nmin, nmax, seed) as b:
        index, nbrs, voff = b.download()
        assert np.array_equal(voff, ref_voff)
        assert np.array_equal(index, ref_index) and np.array_equal(nbrs, ref_nbrs)
        ref = O.cc_labels(ref_index, ref_nbrs)
        for replay in range(3):      # hipGraph replays, each checked
            lab, info = b.cc()
            assert np.array_equal(lab, ref), replay
        assert info.edges == len(ref_nbrs)


def test_batch_from_csr_matches_per_graph_runs():
    index, nbrs, voff = O.batch_graphs(20, 100, 2000, 99)
    with G.Batch.from_csr(index, nbrs, voff) as b:
        lab, _ = b.cc()
    for g in range(20):
        lo, hi = int(voff[g]), int(voff[g + 1])
        li = (index[lo:hi + 1] - index[lo]).astype(np.uint64)
        ln = nbrs[index[lo]:index[hi]] - np.uint32(lo)
        assert np.array_equal(lab[lo:hi] - lo, O.cc_labels(li, ln)), g


def test_batch_rejects_cross_graph_edges():
    index = np.array([0, 1, 2, 2], np.uint64)
    nbrs = np.array([1, 0], np.uint32)
    with pytest.raises(G.GprError):
        G.Batch.from_csr(index, nbrs, np.array([0, 1, 3], np.uint32))
