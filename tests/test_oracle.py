"""CPU tests that pin the oracle (oracle/oracle.c) before it is trusted.

The reference ships no golden vectors for this path (its only tests are x86
SIMD trait tests; the bench ctest is disabled and its graphs are absent), and
it cannot be compiled here (C++23 <stdfloat> and libconfig are missing), so
each oracle function is pinned by mathematics instead:

  orc_mix64            splitmix64 published known answer (seed 0 ->
                       0xe220a8397b1dcdaf) and the avalanche of one input bit
  orc_rmat_*           degenerate probabilities (a=1: all self loops; b=1:
                       edge (0, n-1)), quadrant frequencies vs a/b/c/d
  orc_build_csr        brute-force set construction on random edge lists
  orc_b2_*             byte layout of the header, v2/v3 reads, bad headers
  orc_bfs_levels       brute-force Bellman-Ford hop distances; path closed form
  orc_cc_labels        brute-force DFS components; min-ID labels
  orc_pr_fp64          dense-matrix power iteration; cycle -> uniform 1/n;
                       complete bipartite closed form
  orc_graptor_*        equal to the textbook oracles (BFS/CC bit-exact, PR
                       within FP32 rounding)
  orc_cfp16_*          exhaustive 2^16 decode against the closed form
                       (1 + m/1024) * 2^(e-63), truncation of every encode
  orc_bf16*            truncation vs the upper half; RNE vs torch.bfloat16
  orc_scfp16_*         exhaustive 2^16 decode against the closed form
                       (-1)^s (1 + m/512) 2^(e-63) (code 0 -> 0), truncation
                       toward zero of every encode, the exponent aliasing
  orc_kcore            brute force from the definition (core number = the
                       largest k whose k-core, left by deleting vertices of
                       degree < k until none remains, holds the vertex) on
                       random, RMAT, complete and path graphs
  orc_apr_vary         with eps2 = 0 the delta iteration equals the FP64
                       dense power iteration p' = d*P*p + (1-d)/n after the
                       same number of iterations (directed graphs with sinks
                       and sources, CSR != CSC); a pure-Python FP32 brute
                       force of the frontier branches on tiny graphs, bit for
                       bit; the default stages converge near PageRank
"""
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle as O


def brute_csr(n, src, dst, sym=True):
    s = set()
    for u, v in zip(src.tolist(), dst.tolist()):
        if u == v:
            continue
        s.add((u, v))
        if sym:
            s.add((v, u))
    e = sorted(s)
    index = np.zeros(n + 1, np.uint64)
    for u, _ in e:
        index[u + 1] += 1
    index = np.cumsum(index).astype(np.uint64)
    return index, np.array([v for _, v in e], np.uint32)


def bf_levels(index, nbrs, s):
    n = len(index) - 1
    INF = n + 1
    lv = [INF] * n
    lv[s] = 0
    changed = True
    while changed:
        changed = False
        for u in range(n):
            if lv[u] == INF:
                continue
            for e in range(int(index[u]), int(index[u + 1])):
                v = int(nbrs[e])
                if lv[u] + 1 < lv[v]:
                    lv[v] = lv[u] + 1
                    changed = True
    return np.array(lv, np.uint32)


def dfs_labels(index, nbrs):
    n = len(index) - 1
    lab = [-1] * n
    for s in range(n):
        if lab[s] >= 0:
            continue
        stack, comp = [s], []
        lab[s] = s
        while stack:
            u = stack.pop()
            comp.append(u)
            for e in range(int(index[u]), int(index[u + 1])):
                v = int(nbrs[e])
                if lab[v] < 0:
                    lab[v] = s
                    stack.append(v)
        m = min(comp)
        for u in comp:
            lab[u] = m
    return np.array(lab, np.uint32)


def small_graphs():
    rng = np.random.default_rng(7)
    out = []
    for n, k in ((1, 0), (2, 1), (7, 9), (30, 40), (64, 300), (200, 150)):
        src = rng.integers(0, n, k).astype(np.uint32)
        dst = rng.integers(0, n, k).astype(np.uint32)
        out.append((n, src, dst))
    return out


def test_mix64_known_answer():
    assert O.lib().orc_mix64(0) == 0xE220A8397B1DCDAF
    # avalanche: one flipped input bit flips about half the output bits
    flips = [bin(O.lib().orc_mix64(1 << i) ^ O.lib().orc_mix64(0)).count("1") for i in range(64)]
    assert 20 < np.mean(flips) < 44


def test_rmat_thresholds():
    t = O.rmat_thresholds(0.5, 0.25, 0.125)
    assert list(t) == [1 << 31, 3 << 30, 7 << 29]


def test_rmat_degenerate():
    src, dst = O.rmat_edges(8, 4, (1.0, 0.0, 0.0))
    assert (src == 0).all() and (dst == 0).all()
    index, nbrs = O.build_csr(256, src, dst)
    assert len(nbrs) == 0 and (index == 0).all()
    src, dst = O.rmat_edges(8, 4, (0.0, 1.0, 0.0))
    assert (src == 0).all() and (dst == 255).all()
    index, nbrs = O.build_csr(256, src, dst)
    assert list(nbrs) == [255, 0]


def test_rmat_quadrant_frequencies():
    a, b, c = 0.57, 0.19, 0.19
    src, dst = O.rmat_edges(1, 100000, (a, b, c))
    q = src * 2 + dst
    freq = np.bincount(q, minlength=4) / len(q)
    for f, p in zip(freq, (a, b, c, 1 - a - b - c)):
        assert abs(f - p) < 4 * np.sqrt(p * (1 - p) / len(q))


def test_rmat_stream_is_counter_based():
    s1, d1 = O.rmat_edges(12, 16)
    s2, d2 = O.rmat_edges(12, 16, first=1000, count=500)
    assert (s1[1000:1500] == s2).all() and (d1[1000:1500] == d2).all()


@pytest.mark.parametrize("case", range(6))
def test_build_csr_brute(case):
    n, src, dst = small_graphs()[case]
    for sym in (True, False):
        i1, n1 = O.build_csr(n, src, dst, sym)
        i2, n2 = brute_csr(n, src, dst, sym)
        assert np.array_equal(i1, i2) and np.array_equal(n1, n2)
        assert O.check_csr(i1, n1) == 0


def test_check_csr_detects_errors():
    index = np.array([0, 2, 3], np.uint64)
    assert O.check_csr(index, np.array([1, 0, 0], np.uint32)) == 4   # unsorted row 0
    assert O.check_csr(index, np.array([0, 0, 0], np.uint32)) == 4   # duplicate
    assert O.check_csr(index, np.array([0, 5, 0], np.uint32)) == 3   # out of range
    assert O.check_csr(np.array([0, 2, 4], np.uint64), np.array([0, 1, 0], np.uint32)) == 1  # index[n] != m


def test_b2_layout_and_versions(tmp_path):
    index, nbrs = O.rmat_graph(8, 4)
    p = str(tmp_path / "g.b2")
    O.b2_write(p, index, nbrs)
    raw = open(p, "rb").read()
    n, m = len(index) - 1, len(nbrs)
    assert struct.unpack("<8Q", raw[:64]) == (2, 1, n, m, 4, 8, 0, 0)
    assert len(raw) == 64 + 8 * n + 4 * m   # index[n] is not stored (GraphCSx.h:2023-2024)
    i2, n2 = O.b2_read(p)
    assert np.array_equal(index, i2) and np.array_equal(nbrs, n2)
    # version 3 stores n+1 index entries (GraphCSx.h:2255-2256)
    p3 = str(tmp_path / "g3.b2")
    with open(p3, "wb") as f:
        f.write(struct.pack("<8Q", 3, 1, n, m, 4, 8, 0, 0))
        f.write(index.tobytes())
        f.write(nbrs.tobytes())
    i3, n3 = O.b2_read(p3)
    assert np.array_equal(index, i3) and np.array_equal(nbrs, n3)
    bad = str(tmp_path / "bad.b2")
    with open(bad, "wb") as f:
        f.write(struct.pack("<8Q", 4, 1, n, m, 4, 8, 0, 0))
    with pytest.raises(IOError):
        O.b2_read(bad)


def path_graph(n):
    src = np.arange(n - 1, dtype=np.uint32)
    return O.build_csr(n, src, src + 1)


@pytest.mark.parametrize("case", range(6))
def test_bfs_brute(case):
    n, src, dst = small_graphs()[case]
    index, nbrs = O.build_csr(n, src, dst)
    for s in range(0, n, max(1, n // 5)):
        ref = bf_levels(index, nbrs, s)
        assert np.array_equal(O.bfs_levels(index, nbrs, s), ref)
        lv, _ = O.graptor_bfs(index, nbrs, s, 4)
        assert np.array_equal(lv, ref)


def test_bfs_path_closed_form():
    index, nbrs = path_graph(100)
    lv = O.bfs_levels(index, nbrs, 37)
    assert np.array_equal(lv, np.abs(np.arange(100) - 37).astype(np.uint32))


def test_bfs_rmat_invariant_and_graptor():
    index, nbrs = O.rmat_graph(14)
    s = O.highest_degree_vertex(index)
    lv = O.bfs_levels(index, nbrs, s)
    n = len(index) - 1
    src = np.repeat(np.arange(n), np.diff(index).astype(np.int64))
    a, b = lv[src].astype(np.int64), lv[nbrs].astype(np.int64)
    # BFSLVLv.C:282-296 symmetric-graph validation
    assert (np.abs(a - b) <= 1).all()
    isolated = np.diff(index) == 0
    assert (lv[isolated & (np.arange(n) != s)] == n + 1).all()
    lv2, info = O.graptor_bfs(index, nbrs, s, 8)
    assert np.array_equal(lv, lv2)
    assert info.push_iters >= 1 and info.iters > info.push_iters   # both directions used


@pytest.mark.parametrize("case", range(6))
def test_cc_brute(case):
    n, src, dst = small_graphs()[case]
    index, nbrs = O.build_csr(n, src, dst)
    ref = dfs_labels(index, nbrs)
    assert np.array_equal(O.cc_labels(index, nbrs), ref)
    lab, _ = O.graptor_cc(index, nbrs, 4)
    assert np.array_equal(lab, ref)


def test_cc_rmat_graptor():
    index, nbrs = O.rmat_graph(15)
    ref = O.cc_labels(index, nbrs)
    n = len(index) - 1
    src = np.repeat(np.arange(n), np.diff(index).astype(np.int64))
    assert (ref[src] == ref[nbrs]).all()           # CCv.C:605-660 neighbour check
    assert (ref <= np.arange(n)).all()             # label is the minimum ID
    lab, _ = O.graptor_cc(index, nbrs, 8)
    assert np.array_equal(lab, ref)


def test_pr_dense_matrix():
    for n, src, dst in small_graphs()[2:]:
        index, nbrs = O.build_csr(n, src, dst)
        x, it, _ = O.pr_fp64(index, nbrs, 15, 0.0)
        xd = O.pr_dense_fp64(index, nbrs, 15)
        assert it == 15
        assert np.abs(x - xd).max() < 1e-12


def test_pr_closed_forms():
    n = 50
    src = np.arange(n, dtype=np.uint32)
    index, nbrs = O.build_csr(n, src, (src + 1) % n)      # cycle: 2-regular
    x, _, _ = O.pr_fp64(index, nbrs, 30, 0.0)
    assert np.abs(x - 1.0 / n).max() < 1e-15
    # complete bipartite K(p,q): ranks depend only on the side; the fixed
    # point of x_p = 0.15/n + 0.85*q*x_q/p ... solves to x_p/x_q below
    p, q = 3, 5
    s, d = np.meshgrid(np.arange(p), np.arange(p, p + q))
    index, nbrs = O.build_csr(p + q, s.ravel().astype(np.uint32), d.ravel().astype(np.uint32))
    x, _, _ = O.pr_fp64(index, nbrs, 400, 0.0)
    n = p + q
    # x_p = (1-d)/n + d*q*x_q/q... each p-vertex receives from all q-vertices
    # (outdeg p):  x_p = (1-d)/n + d*q*x_q/p,  x_q = (1-d)/n + d*p*x_p/q
    A = np.array([[1, -0.85 * q / p], [-0.85 * p / q, 1]])
    xp, xq = np.linalg.solve(A, [0.15 / n, 0.15 / n])
    assert abs(x[0] - xp) < 1e-12 and abs(x[-1] - xq) < 1e-12
    assert abs(x.sum() - 1) < 1e-12


def test_pr_rmat_graptor_and_sum():
    index, nbrs = O.rmat_graph(14)
    x, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
    assert abs(x.sum() - 1) < 1e-12
    xf, it, _ = O.graptor_pr(index, nbrs, 20, 0.0, 8)
    assert it == 20
    assert np.abs(xf - x).sum() / x.sum() < 1e-6


def test_pr_tolerance_stops():
    index, nbrs = O.rmat_graph(12)
    x, it, d = O.pr_fp64(index, nbrs, 1000, 1e-7)
    assert d <= 1e-7 and it < 1000


def test_cfp16_exhaustive_closed_form():
    L = O.lib()
    for h in range(1 << 16):
        e6, mant = h >> 10, h & 1023
        v = L.orc_cfp16_decode(h)
        assert v == np.ldexp(1.0 + mant / 1024.0, e6 - 63)
        assert L.orc_cfp16_encode(v) == h
    # truncation: every float in [2^-63, 2) encodes to the code just below it
    rng = np.random.default_rng(3)
    xs = np.ldexp(rng.random(20000) + 1.0, rng.integers(-63, 1, 20000)).astype(np.float32)
    for x in xs:
        h = L.orc_cfp16_encode(float(x))
        assert L.orc_cfp16_decode(h) <= x < (L.orc_cfp16_decode(h + 1) if h < 65535 else 2.0)


def test_bf16_truncation_and_rne():
    rng = np.random.default_rng(5)
    xs = (rng.standard_normal(4000) * 10.0 ** rng.integers(-30, 30, 4000)).astype(np.float32)
    bits = xs.view(np.uint32)
    for x, b in zip(xs, bits):
        assert O.bf16t_encode(float(x)) == int(b) >> 16
        assert O.bf16_decode(int(b) >> 16) == np.float32(np.uint32(int(b) & 0xFFFF0000).view(np.float32))
    torch = pytest.importorskip("torch")
    ref = torch.from_numpy(xs).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    got = np.array([O.bf16_rne_encode(float(x)) for x in xs], np.uint16)
    assert np.array_equal(got, ref)
    # ties go to even
    assert O.bf16_rne_encode(float(np.uint32(0x3F808000).view(np.float32))) == 0x3F80
    assert O.bf16_rne_encode(float(np.uint32(0x3F818000).view(np.float32))) == 0x3F82


def test_planted_mistakes_are_caught():
    """A dropped symmetrisation or an off-by-one level must fail the pins."""
    index, nbrs = O.rmat_graph(10)
    s = O.highest_degree_vertex(index)
    lv = O.bfs_levels(index, nbrs, s)
    wrong = lv.copy()
    wrong[wrong == 2] = 3
    n = len(index) - 1
    src = np.repeat(np.arange(n), np.diff(index).astype(np.int64))
    assert not (np.abs(wrong[src].astype(np.int64) - wrong[nbrs].astype(np.int64)) <= 1).all() or \
        not np.array_equal(wrong, bf_levels(index, nbrs, s))
    src_e, dst_e = O.rmat_edges(8, 4)
    i1, n1 = O.build_csr(256, src_e, dst_e, False)
    i2, n2 = brute_csr(256, src_e, dst_e, True)
    assert not (np.array_equal(i1, i2) and np.array_equal(n1, n2))


def test_batch_sizes_and_graphs():
    index, nbrs, voff = O.batch_graphs(200, 100, 5000, 7)
    nv = np.diff(voff.astype(np.int64))
    assert (nv >= 100).all() and (nv <= 5000).all()
    # log-uniform: the median of log n sits near the middle of the range
    mid = 0.5 * (np.log(100) + np.log(5000))
    assert abs(np.median(np.log(nv)) - mid) < 0.6
    assert O.check_csr(index, nbrs) == 0
    src = np.repeat(np.arange(len(index) - 1), np.diff(index).astype(np.int64))
    g_of = np.searchsorted(voff, src, side="right") - 1
    assert (nbrs >= voff[g_of]).all() and (nbrs < voff[g_of + 1]).all()    # no cross-graph edge


# ---- delta PageRank (APR_vary) ---------------------------------------------------
def scfp16_closed(h):
    if h == 0:
        return 0.0
    sg = -1.0 if h & 0x8000 else 1.0
    return sg * float(np.ldexp(1.0 + (h & 511) / 512.0, ((h >> 9) & 63) - 63))


def test_scfp16_exhaustive_closed_form():
    L = O.lib()
    for h in range(1 << 16):
        v = L.orc_scfp16_decode(h)
        assert v == scfp16_closed(h)
        if h:
            assert L.orc_scfp16_encode(v) == h
    assert L.orc_scfp16_encode(0.0) == 0
    # truncation toward zero inside the range [2^-63, 2), sign kept
    rng = np.random.default_rng(11)
    mag = np.ldexp(rng.random(20000) + 1.0, rng.integers(-63, 1, 20000))
    xs = (mag * np.where(rng.random(20000) < 0.5, -1.0, 1.0)).astype(np.float32)
    for x in xs:
        h = L.orc_scfp16_encode(float(x))
        d = L.orc_scfp16_decode(h)
        assert (d < 0) == (x < 0)
        assert abs(d) <= abs(x)
        if (h & 0x7FFF) != 0x7FFF:
            assert abs(x) < abs(L.orc_scfp16_decode(h + 1))
    # magnitudes >= 2 keep only their low exponent bits (no saturation)
    assert L.orc_scfp16_encode(2.0) == 0
    assert L.orc_scfp16_decode(L.orc_scfp16_encode(3.0)) == scfp16_closed(0x100)


def directed_graph(n, m, seed):
    """Random directed graph with sinks and sources: CSR, CSC, out-degrees."""
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)
    keep = (src % 7 != 3) & (dst % 11 != 5)   # sinks: v%7==3, sources: v%11==5
    src, dst = src[keep], dst[keep]
    index, nbrs = O.build_csr(n, src, dst, False)
    cindex, cnbrs = O.build_csr(n, dst, src, False)
    return index, nbrs, cindex, cnbrs, O.degrees(index)


def power_iterate(index, nbrs, iters, d=0.85):
    n = len(index) - 1
    deg = np.diff(index).astype(np.float64)
    P = np.zeros((n, n))
    for s in range(n):
        for e in range(int(index[s]), int(index[s + 1])):
            P[int(nbrs[e]), s] += 1.0 / deg[s]
    p = np.full(n, 1.0 / n)
    for _ in range(iters):
        p = d * (P @ p) + (1.0 - d) / n
    return p


@pytest.mark.parametrize("graph", ["rmat", "directed"])
# scustomfp<true,6,9> holds magnitudes below 2 only; from iteration 0 the
# scaled contributions 1/(n*factor*deg) lie in (1, 10] for small degrees, so
# (as in the reference's variant 2046) it runs after an FP32 stage.
@pytest.mark.parametrize("stages,tol", [(((0, 0.0),), 1e-6), (((4, 0.0),), 5e-3), (((0, 0.1), (5, 0.0)), 1e-3)])
def test_apr_equals_power_iteration(graph, stages, tol):
    if graph == "rmat":
        index, nbrs = O.rmat_graph(8)
        cindex, cnbrs, deg = index, nbrs, O.degrees(index)
    else:
        index, nbrs, cindex, cnbrs, deg = directed_graph(300, 1500, 4)
    K = 12
    x, it, l1, ed, tr = O.apr_vary(index, nbrs, stages=stages, max_iter=K, eps2=0.0, outdeg=deg,
                                   cindex=cindex, cnbrs=cnbrs, threads=2)
    assert it == K
    p = power_iterate(index, nbrs, K)
    assert np.abs(x - p / np.abs(p).sum()).sum() < tol
    # a transposed graph is a different operator: the pin sees it
    if graph == "directed":
        xt, *_ = O.apr_vary(cindex, cnbrs, stages=stages, max_iter=K, eps2=0.0, outdeg=O.degrees(cindex),
                            cindex=index, cnbrs=nbrs, threads=2)
        assert np.abs(xt - p / np.abs(p).sum()).sum() > 10 * tol


def _enc(fmt, v):
    v = np.float32(v)
    u = int(np.array([v], np.float32).view(np.uint32)[0])
    if fmt == 0:
        return v
    return ((u >> 14) & 0x7FFF) | ((u >> 16) & 0x8000)


def _dec(fmt, c):
    if fmt == 0:
        return np.float32(c)
    return np.float32(scfp16_closed(c))


def apr_brute(index, nbrs, stages, max_iter, eps2):
    """Pure-Python FP32 restatement of APR_vary.C:436-640 / :900-1074."""
    f32 = np.float32
    n = len(index) - 1
    m = int(index[-1])
    deg = [int(index[v + 1] - index[v]) for v in range(n)]
    ins = [[] for _ in range(n)]
    for s in range(n):
        for e in range(int(index[s]), int(index[s + 1])):
            ins[int(nbrs[e])].append(s)
    factor = f32(1.0)
    nnn = n
    while nnn > 1:
        factor = f32(float(factor) / 10.0)
        nnn //= 10
    dampen = f32(0.85)
    ninv = f32(1.0) / f32(n)
    cst = (f32(1.0) - dampen) * ninv
    e2 = f32(eps2)
    x = [f32(0.0)] * n
    delta = [f32(1.0 / n)] * n
    act = [True] * n
    fv, fe, it, kinds = n, m, 0, []
    for fmt, eps in stages:
        while it < max_iter:
            if it == 0:
                kind = 0
            elif (fv + fe) / m > 0.6:
                kind = 1
            else:
                kind = 3 if float(f32(fv + fe)) <= float(f32(m)) / 100.0 else 2
            allsrc = kind <= 1
            c = {}
            for v in range(n):
                if (allsrc or act[v]) and deg[v]:
                    c[v] = _enc(fmt, delta[v] * (f32(1.0) / (factor * f32(deg[v]))))
            z = []
            for d in range(n):
                if fmt == 0:
                    acc = 0.0
                    for s in sorted(ins[d]):
                        if allsrc or act[s]:
                            acc += float(_dec(fmt, c[s])) * float(factor)
                    z.append(f32(acc))
                else:
                    acc = f32(0.0)
                    for s in sorted(ins[d]):
                        if allsrc or act[s]:
                            acc = f32(acc + _dec(fmt, c[s]) * factor)
                    z.append(acc)
            l1, fv, fe, nact = 0.0, 0, 0, []
            for v in range(n):
                if kind == 0:
                    dl = dampen * z[v] + cst
                    x[v] = x[v] + dl
                    dl = dl - ninv
                    a = abs(dl) > e2 * x[v]
                else:
                    dl = z[v] * dampen
                    a = abs(dl) > e2 * x[v]
                    if a:
                        x[v] = x[v] + dl
                delta[v] = dl
                nact.append(a)
                l1 += float(abs(dl))
                fv += int(a)
                fe += deg[v] if a else 0
            act = nact
            kinds.append(kind)
            it += 1
            if l1 < eps:
                break
    ss = sum(abs(float(v)) for v in x)
    return np.array([float(v) * (1.0 / ss) for v in x]), it, kinds


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("stages", [((0, 1e-4),), ((0, 1e-1), (5, 1e-6))])
def test_apr_branches_brute(seed, stages):
    if seed % 2:
        index, nbrs = O.rmat_graph(6, 4, seed=seed + 1)
        cindex, cnbrs = index, nbrs
    else:
        index, nbrs, cindex, cnbrs, _ = directed_graph(64, 400, seed)
    want, wit, wkinds = apr_brute(index, nbrs, stages, 40, 0.05)
    x, it, l1, ed, tr = O.apr_vary(index, nbrs, stages=stages, max_iter=40, eps2=0.05, outdeg=O.degrees(index),
                                   cindex=cindex, cnbrs=cnbrs, threads=1)
    assert it == wit
    assert [int(k) for k in tr[:, 0]] == wkinds
    assert np.array_equal(x, want)
    assert {1, 2, 3} & set(wkinds)   # the frontier branches are exercised


def test_apr_default_stages_near_pagerank():
    index, nbrs = O.rmat_graph(14)
    x, it, l1, ed, tr = O.apr_vary(index, nbrs)
    assert l1 < 1e-7 and it < 100
    kinds = set(int(k) for k in tr[:, 0])
    assert kinds >= {0, 1, 3}
    xp, _, _ = O.pr_fp64(index, nbrs, 200, 1e-13)
    assert np.abs(x - xp / xp.sum()).sum() < 1e-2


def test_apr_rejects_unsigned_formats():
    index, nbrs = O.rmat_graph(6)
    for f in (2, 3):
        with pytest.raises(ValueError):
            O.apr_vary(index, nbrs, stages=((f, 1e-3),))


# ---- k-core (KC_bucket) ---------------------------------------------------------
def brute_core(index, nbrs):
    n = len(index) - 1
    adj = [set(int(u) for u in nbrs[index[v]:index[v + 1]]) for v in range(n)]
    core = [0] * n
    k = 1
    alive = set(range(n))
    while alive:
        # the k-core: delete vertices of degree < k until none is left
        changed = True
        while changed:
            changed = False
            for v in list(alive):
                if len(adj[v] & alive) < k:
                    alive.discard(v)
                    changed = True
        for v in alive:
            core[v] = k
        k += 1
    return np.array(core, np.uint32)


@pytest.mark.parametrize("case", range(8))
def test_kcore_brute(case):
    rng = np.random.default_rng(100 + case)
    if case < 4:
        n = int(rng.integers(5, 60))
        cnt = int(rng.integers(0, 4 * n))
        index, nbrs = O.build_csr(n, rng.integers(0, n, cnt).astype(np.uint32),
                                  rng.integers(0, n, cnt).astype(np.uint32))
    elif case < 6:
        index, nbrs = O.rmat_graph(7 + case - 4, 8, seed=case)
    elif case == 6:   # complete graph K_12 plus a path hanging off it
        src = [i for i in range(12) for j in range(12) if i < j] + list(range(11, 20))
        dst = [j for i in range(12) for j in range(12) if i < j] + list(range(12, 21))
        index, nbrs = O.build_csr(21, np.array(src, np.uint32), np.array(dst, np.uint32))
    else:             # isolated vertices only
        index, nbrs = np.zeros(6, np.uint64), np.zeros(0, np.uint32)
    want = brute_core(index, nbrs)
    got, kmax = O.kcore(index, nbrs)
    assert np.array_equal(got, want)
    assert kmax == (int(want.max()) if len(want) else 0)
    if case == 6:
        assert (got[:11] == 11).all() and (got[12:] == 1).all()


def test_kcore_planted_mistake_fails_brute():
    index, nbrs = O.rmat_graph(8, 8)
    got, _ = O.kcore(index, nbrs)
    wrong = got.copy()
    wrong[np.argmax(wrong)] -= 1
    assert not np.array_equal(wrong, brute_core(index, nbrs))
