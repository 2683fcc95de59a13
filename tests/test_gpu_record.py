"""Record forms of api::record on the GPU, through the C ABI.

The reference has four (api.h:739-757): no record argument (csr_sparse_no_f:
no output), the reduction (activate d when a relax changed it), the method
(BFSLVLv's DEFERRED_UPDATE default, level[d] != prev_level[d],
BFSLVLv.C:166-172) and reduction_or_method (CCv's default, CCv.C:536-547:
the method in push, the reduction in pull, api.h:1401-1416).  In push the
method is evaluated right after a relax that changed d (csr_sparse_with_f,
dsl/edgemap.h:1497-1507: relax && update); in pull it is evaluated for every
destination after its row.

Single steps are checked against these rules evaluated on the CPU, then the
benchmarks run in their DEFERRED_UPDATE=1 configuration, bit-exact against
the oracle.  Bars: bit-exact throughout.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()
    yield


def edges_of(index):
    return np.repeat(np.arange(len(index) - 1), np.diff(index).astype(np.int64))


RECORDS = [G.RECORD_NONE, G.RECORD_REDUCTION, G.RECORD_METHOD_NE, G.RECORD_REDUCTION_OR_METHOD]


def step_case(seed=9):
    """A BFS step whose frontier is all at level 3, so every relax offers 4
    and each destination changes at most once: the method read after the
    relax is then the same whatever order the relaxes run in."""
    index, nbrs = O.rmat_graph(14)
    n = len(index) - 1
    rng = np.random.default_rng(seed)
    level = np.full(n, n + 1, np.uint32)
    F = np.unique(rng.integers(0, n, 500)).astype(np.uint32)
    level[F] = 3
    visited = np.setdiff1d(np.unique(rng.integers(0, n, 2000)), F).astype(np.uint32)
    level[visited] = 2
    src = edges_of(index)
    inF = np.zeros(n, bool)
    inF[F] = True
    hit = np.zeros(n, bool)
    hit[nbrs[inF[src]]] = True
    changed = hit & (level > 4)
    final = level.copy()
    final[changed] = 4
    # prev: the levels before the step, except that a third of the changed
    # vertices already hold 4 (the method leaves them out) and some
    # untouched vertices hold 7 (a pull's method records them, a push never
    # looks at them)
    prev = level.copy()
    ch = np.flatnonzero(changed)
    prev[rng.choice(ch, len(ch) // 3, replace=False)] = 4
    quiet = np.flatnonzero(~changed)
    prev[rng.choice(quiet, 300, replace=False)] = 7
    return index, nbrs, level, F, prev, changed, final


def expected_active(record, direction, changed, final, prev):
    if record == G.RECORD_NONE:
        return np.zeros_like(changed)
    if record == G.RECORD_REDUCTION:
        return changed
    if direction == G.DIR_PUSH:   # method and reduction_or_method
        return changed & (final != prev)
    if record == G.RECORD_METHOD_NE:
        return final != prev
    return changed                # reduction_or_method pulls with the reduction


@pytest.mark.parametrize("record", RECORDS)
@pytest.mark.parametrize("direction", [G.DIR_PUSH, G.DIR_PULL, G.DIR_AUTO])
@pytest.mark.parametrize("kind", [G.FRONTIER_LIST, G.FRONTIER_BITMAP])
def test_single_step_record_forms(record, direction, kind):
    index, nbrs, level, F, prev, changed, final = step_case()
    deg = np.diff(index).astype(np.int64)
    with G.Graph.from_csr(index, nbrs) as g:
        p = g.prop(G.U32)
        p.upload(level)
        q = g.prop(G.U32)
        q.upload(prev)
        fin = g.frontier().set_list(F)
        fout = g.frontier()
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, dst_filter=G.DST_GT, dst_arg=4, direction=direction,
                       out_kind=kind, record=record, prev=q)
        ran = st["direction"]
        assert ran in (G.DIR_PUSH, G.DIR_PULL)
        if direction != G.DIR_AUTO:
            assert ran == direction
        assert np.array_equal(p.download(), final), (record, direction, kind)
        want = expected_active(record, ran, changed, final, prev)
        got = fout.download()
        assert np.array_equal(got, np.flatnonzero(want).astype(np.uint32)), (record, ran, kind)
        assert st["nactv_out"] == want.sum() and st["nacte_out"] == int(deg[want].sum())
        # the graph's dedup bitmap is clear again: a following reduction
        # step records every vertex it changes
        p.upload(level)
        f2 = g.frontier()
        g.edgemap(G.RELAX_MIN_PLUS1, p, fin, f2, dst_filter=G.DST_GT, dst_arg=4, direction=G.DIR_PUSH,
                  out_kind=G.FRONTIER_LIST, record=G.RECORD_REDUCTION)
        assert np.array_equal(f2.download(), np.flatnonzero(changed).astype(np.uint32))
        for h in (f2, fin, fout, q, p):
            h.destroy()


def test_method_records_need_prev():
    index, nbrs, level, F, prev, changed, final = step_case()
    with G.Graph.from_csr(index, nbrs) as g:
        p = g.prop(G.U32)
        p.upload(level)
        fin = g.frontier().set_list(F)
        fout = g.frontier()
        for rec in (G.RECORD_METHOD_NE, G.RECORD_REDUCTION_OR_METHOD):
            with pytest.raises(G.GprError) as e:
                g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, dst_filter=G.DST_GT, dst_arg=4, record=rec)
            assert e.value.code == G.GPR_E_INVAL
        with pytest.raises(G.GprError):
            g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, record=7)
        for h in (fin, fout, p):
            h.destroy()


@pytest.mark.parametrize("with_out", [False, True])
def test_push_without_record_leaves_engine_clean(with_out):
    """A push step with no record argument relaxes but activates nothing; a
    full BFSLVLv and CCv on the same handle afterwards stay bit-exact."""
    index, nbrs, level, F, prev, changed, final = step_case(seed=4)
    n = len(index) - 1
    s = O.highest_degree_vertex(index)
    with G.Graph.from_csr(index, nbrs) as g:
        p = g.prop(G.U32)
        p.upload(level)
        fin = g.frontier().set_list(F)
        fout = g.frontier() if with_out else None
        st = g.edgemap(G.RELAX_MIN_PLUS1, p, fin, fout, dst_filter=G.DST_GT, dst_arg=4, direction=G.DIR_PUSH,
                       out_kind=G.FRONTIER_LIST, record=G.RECORD_NONE)
        assert np.array_equal(p.download(), final)
        assert st["nactv_out"] == 0
        if fout is not None:
            assert len(fout.download()) == 0
            fout.destroy()
        for d in (G.DIR_AUTO, G.DIR_PUSH):
            lv, _ = g.bfslvl(s, direction=d)
            assert np.array_equal(lv, O.bfs_levels(index, nbrs, s))
            lab, _ = g.ccv(direction=d)
            assert np.array_equal(lab, O.cc_labels(index, nbrs))
        fin.destroy()
        p.destroy()
        assert n > 0


def named_graphs():
    out = {}
    n = 300
    s = np.arange(n - 1, dtype=np.uint32)
    out["path"] = O.build_csr(n, s, s + 1)
    out["star"] = O.build_csr(2000, np.zeros(1999, np.uint32), np.arange(1, 2000, dtype=np.uint32))
    out["grid"] = O.grid_graph(40, 30, 0.01, seed=3)
    out["er"] = O.er_graph(5000, 6, seed=11)
    out["rmat14"] = O.rmat_graph(14)
    out["rmat16_ef8_a045"] = O.rmat_graph(16, 8, (0.45, 0.25, 0.25))
    out["two_isolated_pairs"] = O.build_csr(6, np.array([0, 3], np.uint32), np.array([1, 4], np.uint32))
    return out


GRAPHS = named_graphs()


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("reorder", [False, True])
def test_benchmarks_deferred_update_bit_exact(name, reorder):
    """BFSLVLv and CCv with the reference's DEFERRED_UPDATE=1 record forms,
    under every direction and fusion setting."""
    index, nbrs = GRAPHS[name]
    want_cc = O.cc_labels(index, nbrs)
    srcs = sorted({0, O.highest_degree_vertex(index), (len(index) - 1) // 2})
    with G.Graph.from_csr(index, nbrs) as g:
        if reorder:
            g.reorder()
        for d in (G.DIR_AUTO, G.DIR_PUSH, G.DIR_PULL):
            for fu in (-1, 0, 1, 2):
                for s in srcs:
                    lv, _ = g.bfslvl(s, direction=d, fusion=fu, deferred_update=1)
                    assert np.array_equal(lv, O.bfs_levels(index, nbrs, s)), (name, d, fu, s)
                lab, _ = g.ccv(direction=d, fusion=fu, deferred_update=1)
                assert np.array_equal(lab, want_cc), (name, d, fu)


def test_deferred_update_same_steps_rmat20():
    """At RMAT-20 the two record configurations take the same BFS steps with
    the same frontier counts (they activate the same vertices).  CC pulls
    are asynchronous (a row may read labels lowered in the same step,
    CCv.C:571-577 with LEVEL_ASYNC), so its counts vary from run to run in
    either configuration; only its labels are compared."""
    with G.Graph.rmat(20) as g:
        g.reorder()
        s = g.highest_degree_vertex()
        a, ia = g.bfslvl(s, trace=True)
        ta = [(t["direction"], t["nactv_out"], t["nacte_out"]) for t in g.trace()]
        b, ib = g.bfslvl(s, trace=True, deferred_update=1)
        tb = [(t["direction"], t["nactv_out"], t["nacte_out"]) for t in g.trace()]
        assert np.array_equal(a, b) and ta == tb
        la, _ = g.ccv()
        lb, _ = g.ccv(deferred_update=1)
        assert np.array_equal(la, lb)
