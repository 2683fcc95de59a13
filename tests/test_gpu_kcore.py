"""KC_bucket (bench/KC_bucket.C:104-330): core numbers by peeling.

Core numbers are unique, so the GPU's coreness must equal the oracle's
(oracle/oracle.c orc_kcore, the textbook bucket algorithm) bit for bit, and
the largest core must equal largestCore.  The reference's own validation
(KC_bucket.C:332-460) is checked as well: no vertex has a coreness above
its degree, coreness 0 exactly for isolated vertices, and every vertex has
at least coreness-many neighbours of at least its coreness.  The one
cooperative launch is also run through ~25000 barriers (a long path) and
through its failure path, forced from the host with a zero spin limit.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()   # fails loudly when the library or the gfx950 device is missing
    yield


def complete(k, extra_path=0):
    src = [i for i in range(k) for j in range(k) if i < j] + list(range(k - 1, k - 1 + extra_path))
    dst = [j for i in range(k) for j in range(k) if i < j] + list(range(k, k + extra_path))
    return O.build_csr(k + extra_path, np.array(src, np.uint32), np.array(dst, np.uint32))


def named_graphs():
    out = {}
    s = np.arange(299, dtype=np.uint32)
    out["path"] = O.build_csr(300, s, s + 1)
    out["star"] = O.build_csr(2000, np.zeros(1999, np.uint32), np.arange(1, 2000, dtype=np.uint32))
    out["k64_tail"] = complete(64, 40)
    out["grid"] = O.grid_graph(60, 50, 0.01, seed=3)
    out["er"] = O.er_graph(5000, 6, seed=11)
    out["rmat10"] = O.rmat_graph(10)
    out["rmat14"] = O.rmat_graph(14)
    out["rmat16_ef8_a045"] = O.rmat_graph(16, 8, (0.45, 0.25, 0.25))
    out["single"] = (np.zeros(2, np.uint64), np.zeros(0, np.uint32))
    out["isolated_only"] = (np.zeros(100, np.uint64), np.zeros(0, np.uint32))
    out["two_isolated_pairs"] = O.build_csr(6, np.array([0, 3], np.uint32), np.array([1, 4], np.uint32))
    return out


GRAPHS = named_graphs()


def check_core(index, nbrs, core, kmax):
    want, wmax = O.kcore(index, nbrs)
    assert np.array_equal(core, want)
    assert kmax == wmax
    deg = np.diff(index).astype(np.int64)
    assert (core <= deg).all()
    assert np.array_equal(core == 0, deg == 0)
    n = len(index) - 1
    if len(nbrs):
        src = np.repeat(np.arange(n), deg)
        ge = np.bincount(src[core[nbrs] >= core[src]], minlength=n)
        assert (ge >= core).all()


@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_kcore_matches_oracle(name):
    index, nbrs = GRAPHS[name]
    with G.Graph.from_csr(index, nbrs) as g:
        core, kmax, info = g.kcore()
        check_core(index, nbrs, core, kmax)
        assert info.iters == len(np.unique(core))
        assert info.edges == len(nbrs)
        # a second run on the same handle reuses the lists and agrees
        core2, kmax2, _ = g.kcore()
        assert np.array_equal(core, core2) and kmax == kmax2


@pytest.mark.parametrize("name", ["rmat14", "grid", "er", "k64_tail"])
def test_kcore_reordered(name):
    index, nbrs = GRAPHS[name]
    with G.Graph.from_csr(index, nbrs).reorder() as g:
        core, kmax, _ = g.kcore()
    check_core(index, nbrs, core, kmax)


def test_kcore_after_other_apps():
    """The peeling shares the handle with the other benchmarks."""
    index, nbrs = GRAPHS["rmat16_ef8_a045"]
    with G.Graph.from_csr(index, nbrs) as g:
        s = O.highest_degree_vertex(index)
        lv, _ = g.bfslvl(s, fusion=2)
        core, kmax, _ = g.kcore()
        lab, _ = g.ccv()
        check_core(index, nbrs, core, kmax)
        assert np.array_equal(lv, O.bfs_levels(index, nbrs, s))
        assert np.array_equal(lab, O.cc_labels(index, nbrs))


def test_kcore_rmat20():
    with G.Graph.rmat(20) as g:
        index, nbrs = g.download()
        g.reorder()
        core, kmax, info = g.kcore()
    check_core(index, nbrs, core, kmax)
    assert kmax > 100


def test_kcore_rejects_directed():
    rng = np.random.default_rng(4)
    index, nbrs = O.build_csr(300, rng.integers(0, 300, 900).astype(np.uint32),
                              rng.integers(0, 300, 900).astype(np.uint32), symmetrize=False)
    with G.Graph.from_csr(index, nbrs, symmetric=False) as g:
        with pytest.raises(G.GprError):
            g.kcore()


def test_kcore_long_path_and_forced_barrier_failure():
    """A path of 50000 vertices peels one vertex from each end per round:
    about 25000 rounds (barriers) in one launch, coreness 1 everywhere.
    With the spin limit at 0 (forced from the host) a barrier wait gives up
    at once: the call fails with the k-core error, every wave leaves (the
    error flag ends every wait), and the next run on the handle is exact."""
    n = 50000
    s = np.arange(n - 1, dtype=np.uint32)
    index, nbrs = O.build_csr(n, s, s + 1)
    old = G.tune_get("fusion_spin_limit")
    try:
        with G.Graph.from_csr(index, nbrs) as g:
            core, kmax, info = g.kcore()
            assert (core == 1).all() and kmax == 1
            assert info.push_iters >= n // 2
            G.tune_set("fusion_spin_limit", 0)
            with pytest.raises(G.GprError, match="kcore"):
                g.kcore()
            G.tune_set("fusion_spin_limit", old)
            core, kmax, _ = g.kcore()
            assert (core == 1).all() and kmax == 1
    finally:
        G.tune_set("fusion_spin_limit", old)
