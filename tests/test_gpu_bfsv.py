"""BFSv (bench/BFSv.C:88-230): BFS parents through the setif relax.

Which frontier neighbour becomes a vertex's parent depends on the order of
the relaxes (in the reference too), so the parents are checked by the
properties that define a BFS tree, against the oracle's BFS levels:
exactly the reached vertices have a parent; the start is its own parent;
every other parent is an in-neighbour (an edge parent -> vertex exists)
one level closer to the start.  The chain lengths then equal the BFS
levels, which is the reference's own validation (BFSv.C:205-230).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()   # fails loudly when the library or the gfx950 device is missing
    yield


def named_graphs():
    out = {}
    s = np.arange(299, dtype=np.uint32)
    out["path"] = O.build_csr(300, s, s + 1)
    out["star"] = O.build_csr(2000, np.zeros(1999, np.uint32), np.arange(1, 2000, dtype=np.uint32))
    out["grid"] = O.grid_graph(40, 30, 0.01, seed=3)
    out["er"] = O.er_graph(5000, 6, seed=11)
    out["rmat10"] = O.rmat_graph(10)
    out["rmat14"] = O.rmat_graph(14)
    out["rmat16_ef8_a045"] = O.rmat_graph(16, 8, (0.45, 0.25, 0.25))
    out["single"] = (np.zeros(2, np.uint64), np.zeros(0, np.uint32))
    out["two_isolated_pairs"] = O.build_csr(6, np.array([0, 3], np.uint32), np.array([1, 4], np.uint32))
    return out


GRAPHS = named_graphs()


def sources(index):
    n = len(index) - 1
    return sorted({0, n - 1, O.highest_degree_vertex(index), n // 2})


def check_parents(index, nbrs, start, par):
    n = len(index) - 1
    lv = O.bfs_levels(index, nbrs, start)
    reached = lv <= n
    assert np.array_equal(par != G.UNSET, reached)
    assert par[start] == start
    v = np.flatnonzero(reached & (np.arange(n) != start))
    p = par[v].astype(np.int64)
    assert (p < n).all()
    assert np.array_equal(lv[p], lv[v] - 1)
    # the edge p -> v exists (rows of the CSR are sorted)
    keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(index).astype(np.int64)) * n + nbrs
    want = p * n + v
    pos = np.searchsorted(keys, want)
    assert (pos < len(keys)).all() and np.array_equal(keys[np.minimum(pos, len(keys) - 1)], want)
    return lv


@pytest.fixture(scope="module", params=sorted(GRAPHS), ids=sorted(GRAPHS))
def graph(request):
    index, nbrs = GRAPHS[request.param]
    g = G.Graph.from_csr(index, nbrs)
    yield request.param, index, nbrs, g
    g.destroy()


@pytest.mark.parametrize("d", [G.DIR_AUTO, G.DIR_PUSH, G.DIR_PULL])
@pytest.mark.parametrize("k", [0, G.FRONTIER_LIST, G.FRONTIER_BITMAP])
def test_bfsv_parents_form_a_bfs_tree(graph, d, k):
    name, index, nbrs, g = graph
    n = len(index) - 1
    for s in sources(index):
        par, info = g.bfsv(s, direction=d, out_kind=k)
        lv = check_parents(index, nbrs, s, par)
        assert info.edges == int(np.diff(index)[lv <= n].sum()), (name, s)
        assert info.iters == int(lv[lv <= n].max()) + 1 or len(nbrs) == 0


@pytest.mark.parametrize("name", ["rmat14", "grid", "er", "star"])
def test_bfsv_reordered(name):
    index, nbrs = GRAPHS[name]
    with G.Graph.from_csr(index, nbrs).reorder() as g:
        for s in sources(index):
            for d in (G.DIR_AUTO, G.DIR_PULL, G.DIR_PUSH):
                par, _ = g.bfsv(s, direction=d)
                check_parents(index, nbrs, s, par)


def test_bfsv_directed_graph():
    rng = np.random.default_rng(4)
    n = 3000
    src = rng.integers(0, n, 20000).astype(np.uint32)
    dst = rng.integers(0, n, 20000).astype(np.uint32)
    index, nbrs = O.build_csr(n, src, dst, symmetrize=False)
    with G.Graph.from_csr(index, nbrs, symmetric=False) as g:
        for d in (G.DIR_AUTO, G.DIR_PUSH, G.DIR_PULL):
            par, _ = g.bfsv(5, direction=d)
            check_parents(index, nbrs, 5, par)


def test_bfsv_then_bfslvl_same_graph():
    """The parents property is independent of other runs on the graph: a
    level run before and after gives the same levels the tree implies."""
    index, nbrs = GRAPHS["rmat16_ef8_a045"]
    with G.Graph.from_csr(index, nbrs) as g:
        s = O.highest_degree_vertex(index)
        lv0, _ = g.bfslvl(s)
        par, _ = g.bfsv(s)
        lv1, _ = g.bfslvl(s)
        assert np.array_equal(lv0, lv1)
        assert np.array_equal(check_parents(index, nbrs, s, par), lv0)


def test_bfsv_rmat20():
    with G.Graph.rmat(20) as g:
        index, nbrs = g.download()
        order = g.order()
        g.reorder()
        s = O.highest_degree_vertex(index)
        par, info = g.bfsv(s)
        check_parents(index, nbrs, s, par)


def test_bfsv_invalid_start():
    index, nbrs = GRAPHS["path"]
    with G.Graph.from_csr(index, nbrs) as g:
        with pytest.raises(G.GprError):
            g.bfsv(len(index) - 1)


def test_driver_bfsv(tmp_path):
    drv = os.path.join(os.path.dirname(G.__file__), "gpr_driver")
    index, nbrs = O.rmat_graph(12, 8)
    path = str(tmp_path / "rmat12.b2")
    O.b2_write(path, index, nbrs)
    s = O.highest_degree_vertex(index)
    r = subprocess.run([drv, "-app", "bfsv", "-b", "-s", "-rounds", "2", "-start", str(s), "-reorder", path],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    lv = O.bfs_levels(index, nbrs, s)
    n = len(index) - 1
    m = re.search(r"Longest path from \d+ \(original: (\d+)\) : (\d+)", r.stdout)
    assert int(m.group(1)) == s and int(m.group(2)) == int(lv[lv <= n].max())
    assert r.stdout.count("Number of vertices with longer chains: 0 PASS") == 2
    assert r.stdout.count("Every vertex activated at most once: PASS") == 2
    assert "Number of activated vertices: %d" % int((lv <= n).sum()) in r.stdout
