"""APR_vary (bench/APR_vary.C:378-640, :900-1074): delta PageRank.

The GPU run is compared with the oracle's restatement (oracle/oracle.c
orc_apr_vary) on the same graphs.  Both compute every vertex operation in
FP32 without contraction; they differ in the order of the edge sums (the
pull's FP64 or fixed-point sums, the push's FP32 atomics, against the
oracle's CSC order), and a last-bit difference can move a vertex across the
activity test |delta| > eps2*x.  The bars are therefore on the normalised
ranks (sum |x_gpu - x_oracle|, both of unit L1 norm): 2e-6 with FP32
contributions only, 2e-4 with the 16-bit stages (each truncation is 2^-9 or
2^-7 of the value), and on the iteration count (the same, or one apart when
the last L1 sum sits at the threshold).  With eps2 = 0 the iteration is the
power iteration, checked against an FP64 dense product.
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


def directed(n, m, seed):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    keep = (src % 7 != 3) & (dst % 11 != 5)   # sinks and sources
    return O.build_csr(n, src[keep], dst[keep], symmetrize=False)


def named_graphs():
    out = {}
    out["rmat10"] = (O.rmat_graph(10), True)
    out["rmat14"] = (O.rmat_graph(14), True)
    out["rmat16_ef8_a045"] = (O.rmat_graph(16, 8, (0.45, 0.25, 0.25)), True)
    out["er"] = (O.er_graph(5000, 6, seed=11), True)
    out["grid"] = (O.grid_graph(60, 50, 0.01, seed=3), True)
    out["star"] = (O.build_csr(2000, np.zeros(1999, np.uint32), np.arange(1, 2000, dtype=np.uint32)), True)
    out["isolated"] = (O.build_csr(40, np.array([0, 3, 7], np.uint32), np.array([1, 4, 8], np.uint32)), True)
    out["directed"] = (directed(3000, 24000, 4), False)
    return out


GRAPHS = named_graphs()

STAGES = {
    "2046": ((0, 1e-1), (5, 1e-7)),
    "fp32": ((0, 1e-7),),
    "bf16": ((0, 1e-1), (1, 1e-7)),
    "bf16t": ((4, 1e-7),),
}
NAMES = {0: "fp32", 1: "bf16", 4: "bf16t", 5: "scfp16"}
TOL = {"2046": 2e-4, "fp32": 2e-6, "bf16": 2e-4, "bf16t": 2e-4}


def oracle_apr(index, nbrs, sym, stages, eps2=1e-2, max_iter=100):
    if sym:
        return O.apr_vary(index, nbrs, stages=stages, eps2=eps2, max_iter=max_iter)
    cindex, cnbrs = transpose(index, nbrs)
    return O.apr_vary(index, nbrs, stages=stages, eps2=eps2, max_iter=max_iter, outdeg=O.degrees(index),
                      cindex=cindex, cnbrs=cnbrs)


def transpose(index, nbrs):
    n = len(index) - 1
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(index).astype(np.int64))
    return O.build_csr(n, nbrs.astype(np.uint32), src, symmetrize=False)


def gpu_stages(stages):
    return [(NAMES[f], e) for f, e in stages]


@pytest.fixture(scope="module", params=sorted(GRAPHS), ids=sorted(GRAPHS))
def graph(request):
    (index, nbrs), sym = GRAPHS[request.param]
    g = G.Graph.from_csr(index, nbrs, symmetric=sym)
    yield request.param, index, nbrs, sym, g
    g.destroy()


@pytest.mark.parametrize("d", [G.DIR_AUTO, G.DIR_PUSH, G.DIR_PULL])
@pytest.mark.parametrize("variant", sorted(STAGES))
def test_apr_matches_oracle(graph, variant, d):
    name, index, nbrs, sym, g = graph
    stages = STAGES[variant]
    want, wit, wl1, wedges, tr = oracle_apr(index, nbrs, sym, stages)
    x, info = g.apr_vary(gpu_stages(stages), direction=d)
    err = float(np.abs(x - want).sum())
    assert err <= TOL[variant], (name, variant, d, err)
    assert abs(info.iters - wit) <= 1, (name, variant, d, info.iters, wit)
    assert abs(x.sum() - 1.0) < 1e-9 and (x >= 0).all()
    if d == G.DIR_AUTO and info.iters == wit:
        # the same branches: push iterations where the oracle pushes (a
        # vertex at the activity threshold may move one frontier)
        assert abs(info.push_iters - int((tr[:, 0] == 3).sum())) <= 1, (name, variant)
        assert abs(int(info.edges) - int(wedges)) <= 0.01 * wedges, (name, variant)
    if d == G.DIR_PULL:
        assert info.push_iters == 0


@pytest.mark.parametrize("name", ["rmat14", "directed", "grid"])
def test_apr_eps2_zero_is_power_iteration(name):
    """All vertices stay active: after K iterations x is the K-th power
    iterate p' = d*P*p + (1-d)/n from p = 1/n (FP64 dense product)."""
    (index, nbrs), sym = GRAPHS[name]
    n = len(index) - 1
    K = 10
    with G.Graph.from_csr(index, nbrs, symmetric=sym) as g:
        x, info = g.apr_vary([("fp32", 0.0)], eps2=0.0, max_iter=K)
    assert info.iters == K
    deg = np.diff(index).astype(np.float64)
    src = np.repeat(np.arange(n), np.diff(index).astype(np.int64))
    p = np.full(n, 1.0 / n)
    for _ in range(K):
        y = np.zeros(n)
        np.add.at(y, nbrs.astype(np.int64), p[src] / deg[src])
        p = 0.85 * y + 0.15 / n
    assert np.abs(x - p / np.abs(p).sum()).sum() < 2e-6


@pytest.mark.parametrize("name", ["rmat14", "grid", "er"])
def test_apr_reordered(name):
    """Relabelling changes the internal order only: the ranks come back in
    original IDs and agree with the oracle to the same bar."""
    (index, nbrs), sym = GRAPHS[name]
    want, wit, *_ = oracle_apr(index, nbrs, sym, STAGES["2046"])
    with G.Graph.from_csr(index, nbrs).reorder() as g:
        x, info = g.apr_vary()
    assert np.abs(x - want).sum() <= TOL["2046"]
    assert abs(info.iters - wit) <= 1


def test_apr_push_pull_agree_rmat18():
    """Forcing every frontier iteration to push or to pull gives the same
    ranks to within the accumulation-order bar."""
    with G.Graph.rmat(18) as g:
        xa, ia = g.apr_vary(direction=G.DIR_AUTO)
        xp, ip = g.apr_vary(direction=G.DIR_PUSH)
        xl, il = g.apr_vary(direction=G.DIR_PULL)
    assert ia.push_iters > 0 and il.push_iters == 0
    assert np.abs(xa - xp).sum() < 2e-4
    assert np.abs(xa - xl).sum() < 2e-4


def test_apr_rmat20_default_variant():
    with G.Graph.rmat(20) as g:
        index, nbrs = g.download()
        g.reorder()
        x, info = g.apr_vary()
    want, wit, wl1, wedges, tr = O.apr_vary(index, nbrs)
    assert np.abs(x - want).sum() <= TOL["2046"]
    assert abs(info.iters - wit) <= 1
    assert info.delta < 1e-7


def test_apr_deterministic():
    (index, nbrs), _ = GRAPHS["rmat16_ef8_a045"]
    with G.Graph.from_csr(index, nbrs) as g:
        x1, _ = g.apr_vary(direction=G.DIR_PULL)
        x2, _ = g.apr_vary(direction=G.DIR_PULL)
    assert np.array_equal(x1, x2)


def test_apr_rejects_unsigned_formats():
    (index, nbrs), _ = GRAPHS["rmat10"]
    with G.Graph.from_csr(index, nbrs) as g:
        for f in ("fp16", "cfp16"):
            with pytest.raises(G.GprError):
                g.apr_vary([(f, 1e-3)])
        x, _ = g.apr_vary([("fp32", 1e-7)], max_iter=5)   # the handle still works
        assert np.isfinite(x).all()
