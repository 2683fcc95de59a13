"""PageRank with the cold-source split (pr_cold.hip) against the FP64 oracle.

The engine pulls the contributions of sources below a split point and sums
the edges from the other ("cold") sources through a binned layout with
fixed-point accumulation (split point 0: every edge binned).  The split is
set here per test (tunable pr_split_bytes = 4 bytes per pulled source;
1e12 = no split) so that test-sized graphs take the binned path, with tiles
capped by source count (32768) and, in test_tiles_closed_by_edge_count, by
edge count (2^24), several destination bins (16384 each), long rows (> 2048
in-edges), directed graphs, and (on RMAT-20) hub bins of more than 2^20
entries that are summed in chunks.

Bars: the same as the plain pull (relative L1 vs FP64: FP32 1e-6, 16-bit
formats 1e-3 on RMAT and 2e-3 on near-regular graphs, truncating bf16
1e-2); FP32 results within 5e-8 relative L1 of the unsplit engine (both
sum in at least FP64 precision; they differ by final FP32 roundings); runs
bit-identical to each other (fixed-point sums do not depend on the order of
the atomics); sum of ranks 1 within 1e-6.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")

TOL = {"fp32": 1e-6, "bf16": 1e-3, "fp16": 1e-3, "cfp16": 1e-3, "bf16t": 1e-2}
NEAR_REGULAR_FACTOR = 2.0


def graphs():
    out = {}
    out["rmat14"] = O.rmat_graph(14)
    out["rmat16_ef8_a045"] = O.rmat_graph(16, 8, (0.45, 0.25, 0.25))
    out["grid"] = O.grid_graph(200, 180, 0.01, seed=5)
    out["er"] = O.er_graph(40000, 6, seed=12)
    # hub with 5000 leaves: the hub's row is cut into chunks (> 2048 edges)
    # and every leaf is a cold source
    out["star5000"] = O.build_csr(5001, np.zeros(5000, np.uint32), np.arange(1, 5001, dtype=np.uint32))
    return out


GRAPHS = graphs()


@pytest.fixture(autouse=True)
def restore_split():
    G.lib()   # fails loudly when the library or the gfx950 device is missing
    old = G.tune_get("pr_split_bytes")
    yield
    G.tune_set("pr_split_bytes", old)


def split_at(h):
    """Pull sources [0, h); sum the rest through the binned layout."""
    G.tune_set("pr_split_bytes", 4.0 * h)


def rel_l1(a, b):
    return float(np.abs(a.astype(np.float64) - b).sum() / np.abs(b).sum())


@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("reorder", [False, True], ids=["ids", "degree_order"])
def test_split_fp32_vs_fp64_and_unsplit(name, reorder):
    index, nbrs = GRAPHS[name]
    n = len(index) - 1
    x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
    with G.Graph.from_csr(index, nbrs) as g:
        if reorder:
            g.reorder()
        G.tune_set("pr_split_bytes", 1e12)
        x0, _ = g.prv(max_iter=20)
        for h in sorted({0, 1, 7, max(1, n // 3), n - 1}):
            split_at(h)
            x1, info = g.prv(max_iter=20)
            x2, _ = g.prv(max_iter=20)
            assert np.array_equal(x1, x2), (name, h)
            assert rel_l1(x1, x64) <= 1e-6, (name, h, rel_l1(x1, x64))
            assert rel_l1(x1, x0) <= 5e-8, (name, h, rel_l1(x1, x0))
            assert abs(float(x1.sum(dtype=np.float64)) - 1.0) < 1e-6
            assert info.iters == 20


@pytest.mark.parametrize("fmt", sorted(TOL))
@pytest.mark.parametrize("name", sorted(GRAPHS))
def test_split_formats_vs_fp64(name, fmt):
    index, nbrs = GRAPHS[name]
    n = len(index) - 1
    x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
    tol = TOL[fmt] * (1.0 if name.startswith("rmat") or fmt == "fp32" else NEAR_REGULAR_FACTOR)
    with G.Graph.from_csr(index, nbrs).reorder() as g:
        for h in (0, max(1, n // 50)):
            split_at(h)
            x, info = g.pr_vary([(fmt, 0.0)], max_iter=20)
            assert rel_l1(x, x64) <= tol, (name, fmt, h, rel_l1(x, x64))
            xb, _ = g.pr_vary([(fmt, 0.0)], max_iter=20)
            assert np.array_equal(x, xb), (name, fmt, h)
            assert abs(float(x.sum(dtype=np.float64)) - 1.0) < 1e-6


def test_split_converged_stages():
    index, nbrs = GRAPHS["rmat16_ef8_a045"]
    x64, it64, _ = O.pr_fp64(index, nbrs, 1000, 1e-7)
    with G.Graph.from_csr(index, nbrs).reorder() as g:
        split_at(1000)
        x, info = g.prv(max_iter=1000, tol=1e-7)
        assert info.delta <= 1e-7 and abs(info.iters - it64) <= 2
        assert np.abs(x - x64).sum() < 1e-6
        xs, infos = g.pr_vary([("cfp16", 1e-3), ("fp32", 1e-7)], max_iter=1000)
        assert infos.delta <= 1e-7
        assert np.abs(xs - x64).sum() < 2e-6


def test_split_directed_graph():
    rng = np.random.default_rng(8)
    n = 50000
    src = rng.integers(0, n, 400000).astype(np.uint32)
    dst = (rng.pareto(1.2, 400000) * 50).astype(np.int64) % n
    index, nbrs = O.build_csr(n, src, dst.astype(np.uint32), symmetrize=False)
    ci, cn = O.build_csr(n, dst.astype(np.uint32), src, symmetrize=False)
    x64, _, _ = O.pr_fp64(ci, cn, 20, 0.0, outdeg=np.diff(index).astype(np.uint32))
    with G.Graph.from_csr(index, nbrs, symmetric=False) as g:
        for h in (0, 1, 4096, 30000):
            split_at(h)
            x, _ = g.prv(max_iter=20)
            assert rel_l1(x, x64) <= 1e-6, (h, rel_l1(x, x64))


def test_split_point_changes_and_graph_reuse():
    """The layout follows the split point; relabelling drops it."""
    index, nbrs = GRAPHS["rmat14"]
    x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
    with G.Graph.from_csr(index, nbrs) as g:
        for h in (100, 3, 0, 100, 5000):
            split_at(h)
            x, _ = g.prv(max_iter=20)
            assert rel_l1(x, x64) <= 1e-6, h
        g.reorder()
        x, _ = g.prv(max_iter=20)
        assert rel_l1(x, x64) <= 1e-6


@pytest.mark.parametrize("h", [0, 1 << 16])
def test_split_rmat20_chunked_hub_bins(h):
    """RMAT-20 relabelled: the hub bins hold several million entries and are
    summed in chunks of 2^20 whose integer partials are combined."""
    with G.Graph.rmat(20) as g:
        index, nbrs = g.download()
        x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
        g.reorder()
        G.tune_set("pr_split_bytes", 1e12)
        x0, _ = g.prv(max_iter=20)
        split_at(h)
        x1, _ = g.prv(max_iter=20)
        x2, _ = g.prv(max_iter=20)
        assert np.array_equal(x1, x2)
        assert rel_l1(x1, x64) <= 1e-6, rel_l1(x1, x64)
        assert rel_l1(x1, x0) <= 5e-8, rel_l1(x1, x0)
        xb, _ = g.pr_vary([("bf16", 0.0)], max_iter=20)
        assert rel_l1(xb, x64) <= 1e-3, rel_l1(xb, x64)


@pytest.mark.parametrize("name", ["rmat14", "er", "star5000"])
@pytest.mark.parametrize("cells", [0, 2])
def test_write_run_offsets_from_hbm(name, cells):
    """A source tile with more non-empty (tile, bin) cells than the write
    kernel stages in LDS (4096; only above 2^26 vertices) reads its run
    offsets from HBM.  Capping the staging (tunable pr_write_stage_cells)
    forces that path for every tile (0) or for the tiles with more than two
    cells (2): the values land in the same slots, so the ranks are
    bit-identical to the staged run, FP32 and 16-bit."""
    index, nbrs = GRAPHS[name]
    old = G.tune_get("pr_write_stage_cells")
    try:
        with G.Graph.from_csr(index, nbrs).reorder() as g:
            split_at(0)
            want, _ = g.prv(max_iter=20)
            wantb, _ = g.pr_vary([("bf16", 0.0)], max_iter=20)
            G.tune_set("pr_write_stage_cells", cells)
            x, _ = g.prv(max_iter=20)
            xb, _ = g.pr_vary([("bf16", 0.0)], max_iter=20)
        assert np.array_equal(x, want)
        assert np.array_equal(xb, wantb)
    finally:
        G.tune_set("pr_write_stage_cells", old)


@pytest.mark.parametrize("skip", [0, G.PR_SKIP_UNLINKED, G.PR_SKIP_RANKS, G.PR_SKIP_UNLINKED | G.PR_SKIP_RANKS])
@pytest.mark.parametrize("name", sorted(GRAPHS))
@pytest.mark.parametrize("fmt", ["fp32", "bf16", "fp16", "cfp16", "scfp16"])
def test_fused_tail_iteration_bit_identical(name, fmt, skip):
    """gpr_pr_iterate_fused equals gpr_pr_finish followed by the PR_SUM
    edgemap, bit for bit in x and y, over three iterations from the same
    state (every edge binned); the flags that leave ranks unwritten, passed
    to all but the last fused call, change neither."""
    import ctypes as C
    index, nbrs = GRAPHS[name]
    f = G.CONTRIB_NAMES[fmt]
    split_at(0)
    with G.Graph.from_csr(index, nbrs) as g:
        n = g.n
        runs = []
        for fused in (False, True):
            x, y = g.prop(G.F32), g.prop(G.F32)
            cp = g.prop(G.U32 if fmt == "fp32" else G.U16)
            x.upload(np.full(n, 1.0 / n, np.float32))
            G._check(G.lib().gpr_pr_contrib(g.h, x.h, cp.h, f, C.c_float(0.85)))
            g.edgemap(G.RELAX_PR_SUM, y, contrib=cp, contrib_fmt=f, damping=0.85)
            for it in range(3):
                if fused:
                    G._check(G.lib().gpr_pr_iterate_fused(g.h, x.h, y.h, f, C.c_float(0.85), skip if it < 2 else 0,
                                                          None))
                else:
                    G._check(G.lib().gpr_pr_finish(g.h, x.h, y.h, cp.h, f, C.c_float(0.85), None, None))
                    g.edgemap(G.RELAX_PR_SUM, y, contrib=cp, contrib_fmt=f, damping=0.85)
            runs.append((x.download().view(np.uint32).copy(), y.download().view(np.uint32).copy()))
            for p in (x, y, cp):
                p.destroy()
        assert np.array_equal(runs[0][0], runs[1][0]), (name, fmt, "x")
        assert np.array_equal(runs[0][1], runs[1][1]), (name, fmt, "y")


@pytest.mark.parametrize("name", ["rmat14", "grid", "star5000"])
def test_prv_fused_tail_equals_unfused(name):
    """PRv with a fixed iteration count (tail fused into the next write
    stage) equals PRv with a convergence test that never stops it (tails
    not fused): the same 20 iterations, bit-identical ranks."""
    index, nbrs = GRAPHS[name]
    split_at(0)
    with G.Graph.from_csr(index, nbrs) as g:
        xf, inf = g.prv(max_iter=20, tol=0.0)
        xu, inu = g.prv(max_iter=20, tol=1e-300)
        assert inf.iters == inu.iters == 20
        assert np.array_equal(xf.view(np.uint32), xu.view(np.uint32))
        assert inf.delta == pytest.approx(inu.delta, rel=1e-12, abs=1e-300)


def test_fused_tail_needs_binned_layout():
    import ctypes as C
    index, nbrs = GRAPHS["er"]
    G.tune_set("pr_split_bytes", 1e12)   # every edge pulled: not fusable
    with G.Graph.from_csr(index, nbrs) as g:
        x, y = g.prop(G.F32), g.prop(G.F32)
        rc = G.lib().gpr_pr_iterate_fused(g.h, x.h, y.h, G.CONTRIB_F32, C.c_float(0.85), 0, None)
        assert rc == G.GPR_E_INVAL
        # PRv still runs (the tail is not fused) and matches the oracle
        xr, info = g.prv(max_iter=20)
        xc, _, _ = O.graptor_pr(index, nbrs, 20, 0.0)
        assert np.abs(xr.astype(np.float64) - xc).sum() / np.abs(xc).sum() < 1e-6
        x.destroy()
        y.destroy()


def test_tiles_closed_by_edge_count():
    """Sources whose edges fill a tile before its 32768-source cap: 8192
    hubs of ~2200 edges each (18M hub edges, so a tile closes at ~7600
    hubs by the 2^24-edge cap), joined to 200000 leaves.  Binning every
    edge gives the pulled ranks (5e-8) and the FP64 ranks (1e-6), and runs
    repeat bit for bit."""
    hubs, leaves, per = 8192, 200000, 2200
    rng = np.random.default_rng(24)
    src = np.repeat(np.arange(hubs, dtype=np.uint32), per)
    dst = (hubs + rng.integers(0, leaves, size=hubs * per)).astype(np.uint32)
    index, nbrs = O.build_csr(hubs + leaves, src, dst, True)
    deg = np.diff(index)
    assert (deg[:hubs] > 0).all() and int(deg[:8000].sum()) > (1 << 24)
    x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
    with G.Graph.from_csr(index, nbrs) as g:
        g.reorder()
        G.tune_set("pr_split_bytes", 1e12)
        x0, _ = g.prv(max_iter=20)
        split_at(0)
        x1, _ = g.prv(max_iter=20)
        x2, _ = g.prv(max_iter=20)
        assert np.array_equal(x1, x2)
        assert rel_l1(x1, x0) <= 5e-8, rel_l1(x1, x0)
        assert rel_l1(x1, x64) <= 1e-6, rel_l1(x1, x64)
