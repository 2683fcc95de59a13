"""The workloads at their stated sizes, against the oracle.

  RMAT-24 ef16 (16.7M vertices, ~500M directed edges): CCv labels bit-exact
  against union-find; PRv FP32, 20 iterations, relative L1 <= 1e-6 against
  the OpenMP restatement (oracle.c orc_graptor_pr: FP64 row sums, FP64
  vertex scans, PRv.C:176-269) and <= 1e-6 against the FP64 power iteration.

  64-source BFS on the same RMAT-24 graph: every lane bit-exact against a
  single-source run of the OpenMP restatement from that source
  (orc_graptor_bfs, itself pinned to the textbook BFS in test_oracle.py).

  Batched CC over 4096 graphs (vertex counts log-uniform in [1024, 65536],
  even graphs RMAT edge factor 8, odd graphs Erdos-Renyi mean degree 6):
  the device generator equals the CPU generator, and the labels of one
  hipGraph launch equal union-find over the disjoint union, replay after
  replay.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")


@pytest.fixture(scope="module")
def rmat24():
    G.lib()
    g = G.Graph.rmat(24)
    index, nbrs = g.download()          # original IDs
    s = g.highest_degree_vertex()
    g.reorder()
    yield g, index, nbrs, s
    g.destroy()


def test_rmat24_cc_bit_exact(rmat24):
    g, index, nbrs, s = rmat24
    lab, info = g.ccv()
    want = O.cc_labels(index, nbrs)
    assert np.array_equal(lab, want)
    assert info.edges == len(nbrs)


def test_rmat24_bfs_bit_exact(rmat24):
    g, index, nbrs, s = rmat24
    lv, _ = g.bfslvl(s)
    want, _ = O.graptor_bfs(index, nbrs, s)
    assert np.array_equal(lv, want)


def test_rmat24_pagerank_20_iterations(rmat24):
    g, index, nbrs, s = rmat24
    x, info = g.prv(max_iter=20)
    assert info.iters == 20
    xc, it, _ = O.graptor_pr(index, nbrs, 20, 0.0)
    assert it == 20
    err = float(np.abs(x.astype(np.float64) - xc).sum() / np.abs(xc.astype(np.float64)).sum())
    assert err <= 1e-6, err
    assert abs(float(np.sum(x, dtype=np.float64)) - 1.0) < 1e-5


def test_rmat24_pagerank_vs_fp64(rmat24):
    g, index, nbrs, s = rmat24
    x, _ = g.prv(max_iter=20)
    x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
    err = float(np.abs(x - x64).sum() / x64.sum())
    assert err <= 1e-6, err


def seeded_sources(index, k, seed):
    """k sources drawn uniformly from the non-isolated vertices, fixed seed."""
    cand = np.flatnonzero(np.diff(index) > 0)
    rng = np.random.default_rng(seed)
    return rng.choice(cand, size=k, replace=False).astype(np.uint32)


def test_rmat24_msbfs_64_sources_every_lane(rmat24):
    g, index, nbrs, s = rmat24
    n = len(index) - 1
    src = seeded_sources(index, 64, 0x47524150)
    lv, info = g.msbfs(src)
    assert lv.shape == (64, n)
    edges = 0
    deg = np.diff(index).astype(np.int64)
    for j, v in enumerate(src):
        want, _ = O.graptor_bfs(index, nbrs, int(v))
        assert np.array_equal(lv[j], want), (j, int(v))
        edges += int(deg[want <= n].sum())
    assert info.edges == edges


def test_batched_cc_4096_graphs():
    G.lib()
    count, nmin, nmax, seed = 4096, 1024, 65536, 0x47524150
    with G.Batch.generate(count, nmin, nmax, seed) as b:
        index, nbrs, voff = b.download()
        assert b.count == count and len(voff) == count + 1
        sizes = np.diff(voff.astype(np.int64))
        assert sizes.min() >= nmin and sizes.max() <= nmax
        want = O.cc_labels(index, nbrs)
        for replay in range(3):
            lab, info = b.cc()
            assert np.array_equal(lab, want), replay
        assert info.edges == len(nbrs)
    # the device generator is the CPU generator
    ri, rn, rv = O.batch_graphs(count, nmin, nmax, seed)
    assert np.array_equal(rv, voff) and np.array_equal(ri, index) and np.array_equal(rn, nbrs)
