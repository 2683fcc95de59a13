"""Fused sparse push (Graptor's api::fusion) against the oracle.

A fused step relaxes the edges of every vertex it activates in the same
launch, until nothing is left (fusion.hip: the levels of the push run
inside one cooperative launch, separated by grid barriers).  The relaxations are monotone
minima, so BFS levels and CC labels must equal the level-synchronous ones
bit for bit whatever order the work queue takes.  Covered: the benchmarks'
FUSION settings (BFSLVLv.C:28-40: off by default, 1 = low-degree graphs and
after a dense step, 2 = every sparse step; CCv.C:47-54: FUSION 1 =
default_fusion_select, frontiers of at least 4096 edges), relabelled graphs,
a directed graph, a long path (200000 levels of one vertex in one launch),
a hub whose row is cut into slices for many waves, one
fused edgemap step called directly, the engine state after a fused step
(queued bits clear, so later unfused steps are exact), and the failure
path forced from the host (spin limit 0: a barrier wait gives up at once,
the step fails cleanly, every wave exits, and the next step is exact
again).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")

FUSION_DEFAULT, FUSION_OFF, FUSION_1, FUSION_2 = 0, -1, 1, 2


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()   # fails loudly when the library or the gfx950 device is missing
    yield


@pytest.fixture(autouse=True)
def restore_spin_limit():
    old = G.tune_get("fusion_spin_limit")
    yield
    G.tune_set("fusion_spin_limit", old)


def named_graphs():
    out = {}
    s = np.arange(299, dtype=np.uint32)
    out["path"] = O.build_csr(300, s, s + 1)
    out["star"] = O.build_csr(2000, np.zeros(1999, np.uint32), np.arange(1, 2000, dtype=np.uint32))
    out["grid"] = O.grid_graph(120, 100, 0.0, seed=3)
    out["grid_long"] = O.grid_graph(120, 100, 0.01, seed=3)
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


def visited_edges(index, lv):
    return int(np.diff(index)[lv <= len(index) - 1].sum())


@pytest.fixture(scope="module", params=sorted(GRAPHS), ids=sorted(GRAPHS))
def graph(request):
    index, nbrs = GRAPHS[request.param]
    g = G.Graph.from_csr(index, nbrs)
    yield request.param, index, nbrs, g
    g.destroy()


@pytest.mark.parametrize("fusion", [FUSION_1, FUSION_2])
def test_bfs_fused_bit_exact(graph, fusion):
    name, index, nbrs, g = graph
    for s in sources(index):
        ref = O.bfs_levels(index, nbrs, s)
        lv, info = g.bfslvl(s, fusion=fusion, trace=True)
        assert np.array_equal(lv, ref), (name, s, fusion)
        assert info.edges == visited_edges(index, ref), (name, s)
        tr = g.trace()
        if fusion == FUSION_2:
            # every sparse (push) step is fused; a fused step either ends the
            # traversal or hands back a level that turned dense
            pushes = [i for i, t in enumerate(tr) if t["direction"] == G.DIR_PUSH and t["nactv_in"]]
            assert all(tr[i]["fused"] == 1 for i in pushes), (name, s, tr)
            for i in pushes:
                assert (tr[i]["nactv_out"] == 0) == (i == len(tr) - 1), (name, s, tr)


@pytest.mark.parametrize("fusion", [FUSION_DEFAULT, FUSION_OFF, FUSION_1, FUSION_2])
def test_cc_fused_bit_exact(graph, fusion):
    name, index, nbrs, g = graph
    lab, _ = g.ccv(fusion=fusion)
    assert np.array_equal(lab, O.cc_labels(index, nbrs)), (name, fusion)


@pytest.mark.parametrize("name", ["rmat14", "grid_long", "er"])
def test_fused_reordered_bit_exact(name):
    index, nbrs = GRAPHS[name]
    with G.Graph.from_csr(index, nbrs).reorder() as g:
        for s in sources(index):
            lv, _ = g.bfslvl(s, fusion=FUSION_2)
            assert np.array_equal(lv, O.bfs_levels(index, nbrs, s)), (name, s)
        lab, _ = g.ccv(fusion=FUSION_2)
        assert np.array_equal(lab, O.cc_labels(index, nbrs)), name


def test_fused_directed_graph():
    rng = np.random.default_rng(4)
    n = 3000
    src = rng.integers(0, n, 20000).astype(np.uint32)
    dst = rng.integers(0, n, 20000).astype(np.uint32)
    index, nbrs = O.build_csr(n, src, dst, symmetrize=False)
    with G.Graph.from_csr(index, nbrs, symmetric=False) as g:
        for s in (5, 77, n - 1):
            lv, _ = g.bfslvl(s, fusion=FUSION_2)
            assert np.array_equal(lv, O.bfs_levels(index, nbrs, s)), s


def test_fused_long_path_and_hub():
    """A path of 200000 vertices: 200000 levels of one vertex, each ended by
    a grid barrier; a hub of 300000 leaves: its row is cut into slices of
    256 edges spread over the waves."""
    n = 200000
    s = np.arange(n - 1, dtype=np.uint32)
    index, nbrs = O.build_csr(n, s, s + 1)
    with G.Graph.from_csr(index, nbrs) as g:
        lv, info = g.bfslvl(0, fusion=FUSION_2)
        assert np.array_equal(lv, np.arange(n, dtype=np.uint32))
        lv, _ = g.bfslvl(n // 2, fusion=FUSION_2)
        assert np.array_equal(lv, O.bfs_levels(index, nbrs, n // 2))
        lab, _ = g.ccv(fusion=FUSION_2)
        assert not lab.any()
    m = 300000
    hub_src = np.zeros(m, np.uint32)
    leaves = np.arange(1, m + 1, dtype=np.uint32)
    # leaves also form pairs, so the second level pushes again
    index, nbrs = O.build_csr(m + 1, np.concatenate([hub_src, leaves[:-1:2]]),
                              np.concatenate([leaves, leaves[1::2]]))
    with G.Graph.from_csr(index, nbrs) as g:
        for st in (0, 1, m):
            lv, _ = g.bfslvl(st, fusion=FUSION_2)
            assert np.array_equal(lv, O.bfs_levels(index, nbrs, st)), st


def test_dense_level_handed_back_then_pulled():
    """With the direction left to the engine, a fused step stops at the first
    level above the push/pull threshold and returns it as its output
    frontier (a list with exact counts); the next step pulls it.  With push
    forced, one fused step runs to the fixpoint."""
    index, nbrs = GRAPHS["rmat14"]
    n = len(index) - 1
    deg = np.diff(index).astype(np.int64)
    s = int(np.flatnonzero(deg == 1)[0])
    ref = O.bfs_levels(index, nbrs, s)
    with G.Graph.from_csr(index, nbrs) as g:
        lv, info = g.bfslvl(s, fusion=FUSION_2, trace=True)
        assert np.array_equal(lv, ref)
        tr = g.trace()
        assert tr[0]["fused"] == 1 and tr[0]["nactv_out"] > 0 and tr[1]["direction"] == G.DIR_PULL, tr[:2]
        # the handed-back frontier is the first level above the threshold
        m = int(deg.sum())
        lvl = next(L for L in range(1, n) if (ref == L).sum() + deg[ref == L].sum() > m // 20)
        assert tr[0]["nactv_out"] == (ref == lvl).sum() and tr[0]["nacte_out"] == deg[ref == lvl].sum(), lvl
        lv, _ = g.bfslvl(s, fusion=FUSION_2, direction=G.DIR_PUSH, trace=True)
        assert np.array_equal(lv, ref)
        tr = g.trace()
        assert len(tr) == 1 and tr[0]["fused"] == 1 and tr[0]["nactv_out"] == 0
        # direct call: the output frontier is a list of that level
        p = g.prop(G.U32)
        lv0 = np.full(n, n + 1, np.uint32)
        lv0[s] = 0
        p.upload(lv0)
        fin = g.frontier().set_single(s)
        fout = g.frontier()
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, direction=G.DIR_AUTO, fusion=G.FUSION_ON)
        assert st["fused"] == 1 and fout.stats()[2] == G.FRONTIER_LIST
        assert np.array_equal(np.sort(fout.download()), np.flatnonzero(ref == lvl).astype(np.uint32))
        got = p.download()
        assert np.array_equal(got[ref <= lvl], ref[ref <= lvl]) and (got[ref > lvl] == n + 1).all()
        # an unfused step after it deduplicates exactly (bitmaps left clear)
        fnext = g.frontier()
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fout, fnext, direction=G.DIR_PUSH, out_kind=G.FRONTIER_LIST)
        assert st["nactv_out"] == (ref == lvl + 1).sum()
        for h in (fin, fout, fnext, p):
            h.destroy()


def test_fused_edgemap_step_direct():
    """One edgemap call with fusion on from a single source computes the
    whole BFS; its output frontier is empty and its stats say fused.  From
    the full frontier with min-relax it computes the component minima."""
    index, nbrs = GRAPHS["rmat14"]
    n = len(index) - 1
    s = O.highest_degree_vertex(index)
    ref = O.bfs_levels(index, nbrs, s)
    with G.Graph.from_csr(index, nbrs) as g:
        p = g.prop(G.U32)
        lv0 = np.full(n, n + 1, np.uint32)
        lv0[s] = 0
        p.upload(lv0)
        fin = g.frontier().set_single(s)
        fout = g.frontier()
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, direction=G.DIR_PUSH, fusion=G.FUSION_ON)
        assert np.array_equal(p.download(), ref)
        assert st["fused"] == 1 and st["nactv_out"] == 0 and st["direction"] == G.DIR_PUSH
        assert st["edges_scanned"] >= visited_edges(index, ref)
        assert fout.stats() == (0, 0, G.FRONTIER_EMPTY)
        # the default selector does not fuse a one-vertex frontier of < 4096 edges
        lo = int(np.argmin(np.where(np.diff(index) > 0, np.diff(index), 1 << 40)))
        lv1 = np.full(n, n + 1, np.uint32)
        lv1[lo] = 0
        p.upload(lv1)
        fin.set_single(lo)
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, direction=G.DIR_PUSH, fusion=G.FUSION_DEFAULT)
        nb = np.setdiff1d(nbrs[index[lo]:index[lo + 1]], [lo])
        assert st["fused"] == 0 and st["nactv_out"] == len(nb)
        # a caller's list may repeat IDs: two sources, each listed twice
        ref2 = np.minimum(ref, O.bfs_levels(index, nbrs, lo))
        lv2 = np.full(n, n + 1, np.uint32)
        lv2[[s, lo]] = 0
        p.upload(lv2)
        fin.set_list(np.array([s, lo, s, lo], np.uint32))
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, direction=G.DIR_PUSH, fusion=G.FUSION_ON)
        assert st["fused"] == 1 and np.array_equal(p.download(), ref2)
        # labels: min over each component from the identity labelling
        p.upload(np.arange(n, dtype=np.uint32))
        fin.set_all()
        st = g.edgemap(G.RELAX_MIN, p, fin, fout, direction=G.DIR_PUSH, fusion=G.FUSION_ON)
        assert st["fused"] == 1
        assert np.array_equal(p.download(), O.cc_labels(index, nbrs))
        for h in (fin, fout, p):
            h.destroy()


def test_unfused_steps_exact_after_fused_steps():
    """A fused step leaves every queued bit clear; the unfused push uses the
    same bits to deduplicate its output list."""
    index, nbrs = GRAPHS["rmat16_ef8_a045"]
    with G.Graph.from_csr(index, nbrs) as g:
        for s in sources(index):
            ref = O.bfs_levels(index, nbrs, s)
            a, _ = g.bfslvl(s, fusion=FUSION_2)
            b, _ = g.bfslvl(s, fusion=FUSION_OFF, direction=G.DIR_PUSH, out_kind=G.FRONTIER_LIST)
            assert np.array_equal(a, ref) and np.array_equal(b, ref), s
        lab, _ = g.ccv(fusion=FUSION_OFF, direction=G.DIR_PUSH, out_kind=G.FRONTIER_LIST)
        assert np.array_equal(lab, O.cc_labels(index, nbrs))


def test_forced_spin_limit_failure_then_recovery():
    """Spin limit 0: a barrier wait that does not see every block arrive on
    its first polls gives up, so a traversal of a long path (50000 barriers)
    fails with the fused-push error, and every wave leaves its loop.  The
    engine state is restored: the next fused and unfused runs are exact."""
    n = 50000
    s = np.arange(n - 1, dtype=np.uint32)
    index, nbrs = O.build_csr(n, s, s + 1)
    with G.Graph.from_csr(index, nbrs) as g:
        G.tune_set("fusion_spin_limit", 0)
        with pytest.raises(G.GprError, match="fused push"):
            g.bfslvl(0, fusion=FUSION_2)
        G.tune_set("fusion_spin_limit", 1 << 22)
        ref = np.arange(n, dtype=np.uint32)
        lv, _ = g.bfslvl(0, fusion=FUSION_2)
        assert np.array_equal(lv, ref)
        lv, _ = g.bfslvl(0, fusion=FUSION_OFF, direction=G.DIR_PUSH, out_kind=G.FRONTIER_LIST)
        assert np.array_equal(lv, ref)


@pytest.mark.parametrize("scale", [20])
def test_rmat20_fused_equals_unfused(scale):
    with G.Graph.rmat(scale) as g:
        index, _ = g.download()
        for s in (0, O.highest_degree_vertex(index), 12345):
            a, ia = g.bfslvl(s, fusion=FUSION_OFF)
            b, ib = g.bfslvl(s, fusion=FUSION_2)
            c, _ = g.bfslvl(s, fusion=FUSION_1)
            assert np.array_equal(a, b) and np.array_equal(a, c), s
            assert ia.edges == ib.edges
        la, _ = g.ccv(fusion=FUSION_OFF)
        for f in (FUSION_DEFAULT, FUSION_2):
            lb, _ = g.ccv(fusion=f)
            assert np.array_equal(la, lb), f
