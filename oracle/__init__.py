"""CPU oracle for the traversal hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package, and only as the checker (or the timed
CPU baseline).  The product (``graptor_amd``) never imports it.

The C restatement lives in ``oracle.c`` (built by ``oracle/Makefile`` into
``oracle/_build/liboracle.so``); this module wraps it with ctypes and adds a
few numpy helpers.  Every C function cites the reference file:line it
restates; the checks that pin each function are listed in
``tests/test_oracle.py``.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class RunInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("push_iters", C.c_int), ("edges", C.c_uint64)]


def build():
    """Compile the C oracle (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.orc_mix64.restype = C.c_uint64
        L.orc_mix64.argtypes = [C.c_uint64]
        L.orc_rmat_thresholds.argtypes = [C.c_double, C.c_double, C.c_double, u64p]
        L.orc_rmat_edges.argtypes = [C.c_uint32, u64p, C.c_uint64, C.c_uint64, C.c_uint64, u32p, u32p]
        L.orc_er_edges.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, u32p, u32p]
        L.orc_batch_sizes.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, u32p]
        L.orc_batch_graph.restype = C.c_uint64
        L.orc_batch_graph.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64, u64p, u32p]
        L.orc_build_csr.restype = C.c_uint64
        L.orc_build_csr.argtypes = [C.c_uint32, C.c_uint64, u32p, u32p, C.c_int, u64p, u32p]
        L.orc_b2_write.argtypes = [C.c_char_p, C.c_uint32, C.c_uint64, u64p, u32p]
        L.orc_b2_read.argtypes = [C.c_char_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p]
        L.orc_check_csr.argtypes = [C.c_uint32, C.c_uint64, u64p, u32p]
        L.orc_bfs_levels.argtypes = [C.c_uint32, u64p, u32p, C.c_uint32, u32p]
        L.orc_cc_labels.argtypes = [C.c_uint32, u64p, u32p, u32p]
        L.orc_pr_fp64.restype = C.c_int
        L.orc_pr_fp64.argtypes = [C.c_uint32, u64p, u32p, u32p, C.c_int, C.c_double, C.c_double,
                                  f64p, C.POINTER(C.c_double)]
        L.orc_graptor_bfs.argtypes = [C.c_uint32, C.c_uint64, u64p, u32p, C.c_uint32, u32p, C.c_int,
                                      C.POINTER(RunInfo)]
        L.orc_graptor_cc.argtypes = [C.c_uint32, C.c_uint64, u64p, u32p, u32p, C.c_int, C.POINTER(RunInfo)]
        L.orc_graptor_pr.restype = C.c_int
        L.orc_graptor_pr.argtypes = [C.c_uint32, u64p, u32p, u32p, C.c_int, C.c_double, f32p, C.c_int,
                                     C.POINTER(C.c_double)]
        L.orc_scfp16_encode.restype = C.c_uint16
        L.orc_scfp16_encode.argtypes = [C.c_float]
        L.orc_scfp16_decode.restype = C.c_float
        L.orc_scfp16_decode.argtypes = [C.c_uint16]
        L.orc_apr_vary.restype = C.c_int
        L.orc_apr_vary.argtypes = [C.c_uint32, C.c_uint64, u64p, u32p, u32p, C.POINTER(C.c_int),
                                   C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double, f64p, C.c_int,
                                   C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_void_p]
        L.orc_kcore.restype = C.c_uint32
        L.orc_kcore.argtypes = [C.c_uint32, u64p, u32p, u32p]
        L.orc_cfp16_encode.restype = C.c_uint16
        L.orc_cfp16_encode.argtypes = [C.c_float]
        L.orc_cfp16_decode.restype = C.c_float
        L.orc_cfp16_decode.argtypes = [C.c_uint16]
        L.orc_bf16t_encode.restype = C.c_uint16
        L.orc_bf16t_encode.argtypes = [C.c_float]
        L.orc_bf16_rne_encode.restype = C.c_uint16
        L.orc_bf16_rne_encode.argtypes = [C.c_float]
        L.orc_bf16_decode.restype = C.c_float
        L.orc_bf16_decode.argtypes = [C.c_uint16]
        _lib = L
    return _lib


# --------------------------------------------------------------------------
# Graph generation
# --------------------------------------------------------------------------
RMAT_DEFAULT = (0.57, 0.19, 0.19)
DEFAULT_SEED = 0x47524150


def rmat_thresholds(a, b, c):
    t = np.zeros(3, np.uint64)
    lib().orc_rmat_thresholds(a, b, c, t)
    return t


def rmat_edges(scale, edge_factor, abc=RMAT_DEFAULT, seed=DEFAULT_SEED, first=0, count=None):
    if count is None:
        count = edge_factor << scale
    t = rmat_thresholds(*abc)
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    lib().orc_rmat_edges(scale, t, seed, first, count, src, dst)
    return src, dst


def er_edges(n, count, seed):
    src = np.empty(count, np.uint32)
    dst = np.empty(count, np.uint32)
    lib().orc_er_edges(n, seed, count, src, dst)
    return src, dst


def build_csr(n, src, dst, symmetrize=True):
    src = np.ascontiguousarray(src, np.uint32)
    dst = np.ascontiguousarray(dst, np.uint32)
    index = np.zeros(n + 1, np.uint64)
    nbrs = np.empty(max(1, (2 if symmetrize else 1) * len(src)), np.uint32)
    m = lib().orc_build_csr(n, len(src), src, dst, 1 if symmetrize else 0, index, nbrs)
    return index, nbrs[:m].copy()


def rmat_graph(scale, edge_factor=16, abc=RMAT_DEFAULT, seed=DEFAULT_SEED):
    """Cleaned, symmetrised RMAT graph as (index[n+1], nbrs[m])."""
    src, dst = rmat_edges(scale, edge_factor, abc, seed)
    return build_csr(1 << scale, src, dst, True)


def er_graph(n, mean_degree, seed):
    src, dst = er_edges(n, (n * mean_degree) // 2, seed)
    return build_csr(n, src, dst, True)


def grid_graph(w, h, long_frac=0.0, seed=1):
    """w x h 4-neighbour grid plus round(long_frac*w*h) uniform random extra
    edges, symmetrised and cleaned."""
    n = w * h
    v = np.arange(n, dtype=np.uint64)
    x = v % w
    right = v[x + 1 < w]
    down = v[v + w < n]
    src = [right, down]
    dst = [right + 1, down + w]
    k = int(round(long_frac * n))
    if k:
        s2, d2 = er_edges(n, k, seed)
        src.append(s2.astype(np.uint64))
        dst.append(d2.astype(np.uint64))
    src = np.concatenate(src).astype(np.uint32)
    dst = np.concatenate(dst).astype(np.uint32)
    return build_csr(n, src, dst, True)


def batch_graphs(count, nmin, nmax, seed):
    """The batch of small graphs as one CSR with global IDs plus vertex
    offsets (index, nbrs, voff)."""
    nv = np.empty(count, np.uint32)
    lib().orc_batch_sizes(count, nmin, nmax, seed, nv)
    voff = np.zeros(count + 1, np.uint64)
    voff[1:] = np.cumsum(nv.astype(np.uint64))
    idx_parts, nb_parts = [np.zeros(1, np.uint64)], []
    ebase = 0
    for g in range(count):
        n = int(nv[g])
        index = np.zeros(n + 1, np.uint64)
        nbrs = np.empty(16 * n + 16, np.uint32)
        m = lib().orc_batch_graph(g, n, seed, index, nbrs)
        idx_parts.append(index[1:] + np.uint64(ebase))
        nb_parts.append(nbrs[:m] + np.uint32(voff[g]))
        ebase += m
    return (np.concatenate(idx_parts), np.concatenate(nb_parts) if nb_parts else np.zeros(0, np.uint32),
            voff.astype(np.uint32))


def b2_write(path, index, nbrs):
    n = len(index) - 1
    r = lib().orc_b2_write(path.encode(), n, len(nbrs), np.ascontiguousarray(index, np.uint64),
                           np.ascontiguousarray(nbrs, np.uint32))
    if r != 0:
        raise IOError(path)


def b2_read(path):
    n = C.c_uint32()
    m = C.c_uint64()
    r = lib().orc_b2_read(path.encode(), C.byref(n), C.byref(m), None, None)
    if r != 0:
        raise IOError("b2 header %s: %d" % (path, r))
    index = np.empty(n.value + 1, np.uint64)
    nbrs = np.empty(max(1, m.value), np.uint32)
    r = lib().orc_b2_read(path.encode(), C.byref(n), C.byref(m), index.ctypes.data, nbrs.ctypes.data)
    if r != 0:
        raise IOError("b2 body %s: %d" % (path, r))
    return index, nbrs[: m.value]


def check_csr(index, nbrs):
    return lib().orc_check_csr(len(index) - 1, len(nbrs), index, nbrs)


def degrees(index):
    return np.diff(index).astype(np.uint32)


def highest_degree_vertex(index):
    """Highest out-degree vertex, ties to the lowest ID."""
    return int(np.argmax(np.diff(index)))


# --------------------------------------------------------------------------
# Oracles
# --------------------------------------------------------------------------
def bfs_levels(index, nbrs, start):
    n = len(index) - 1
    out = np.empty(n, np.uint32)
    lib().orc_bfs_levels(n, index, nbrs, start, out)
    return out


def cc_labels(index, nbrs):
    n = len(index) - 1
    out = np.empty(n, np.uint32)
    lib().orc_cc_labels(n, index, nbrs, out)
    return out


def pr_fp64(index, nbrs, max_iter=20, tol=0.0, damping=0.85, outdeg=None):
    n = len(index) - 1
    if outdeg is None:
        outdeg = degrees(index)
    x = np.empty(n, np.float64)
    d = C.c_double()
    it = lib().orc_pr_fp64(n, index, nbrs, np.ascontiguousarray(outdeg, np.uint32), max_iter, tol, damping,
                           x, C.byref(d))
    return x, it, d.value


def default_threads():
    """OMP_NUM_THREADS when set, else the CPUs this process may run on."""
    return int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))


def graptor_bfs(index, nbrs, start, threads=None):
    n = len(index) - 1
    out = np.empty(n, np.uint32)
    info = RunInfo()
    lib().orc_graptor_bfs(n, len(nbrs), index, nbrs, start, out, threads or default_threads(), C.byref(info))
    return out, info


def graptor_cc(index, nbrs, threads=None):
    n = len(index) - 1
    out = np.empty(n, np.uint32)
    info = RunInfo()
    lib().orc_graptor_cc(n, len(nbrs), index, nbrs, out, threads or default_threads(), C.byref(info))
    return out, info


def graptor_pr(index, nbrs, max_iter=20, tol=0.0, threads=None, outdeg=None):
    n = len(index) - 1
    if outdeg is None:
        outdeg = degrees(index)
    x = np.empty(n, np.float32)
    d = C.c_double()
    it = lib().orc_graptor_pr(n, index, nbrs, np.ascontiguousarray(outdeg, np.uint32), max_iter, tol, x,
                              threads or default_threads(), C.byref(d))
    return x, it, d.value


def cfp16_encode(f):
    return lib().orc_cfp16_encode(float(f))


def cfp16_decode(h):
    return lib().orc_cfp16_decode(int(h))


def bf16t_encode(f):
    return lib().orc_bf16t_encode(float(f))


def bf16_rne_encode(f):
    return lib().orc_bf16_rne_encode(float(f))


def bf16_decode(h):
    return lib().orc_bf16_decode(int(h))


# --------------------------------------------------------------------------
# numpy restatements for small cases
# --------------------------------------------------------------------------
def scfp16_decode_c(h):
    return lib().orc_scfp16_decode(int(h))


def scfp16_encode_c(f):
    return lib().orc_scfp16_encode(float(f))


APR_STAGES_2046 = ((0, 1e-1), (5, 1e-7))


def apr_vary(index, nbrs, stages=APR_STAGES_2046, max_iter=100, eps2=1e-2, threads=None, outdeg=None,
             cindex=None, cnbrs=None):
    """Delta PageRank (APR_vary).  Returns (x normalised, iterations, final
    L1, edges traversed, trace[iters, 4] = kind, nactv, nacte, L1)."""
    n = len(index) - 1
    if outdeg is None:
        outdeg = degrees(index)
    if cindex is None:
        cindex, cnbrs = index, nbrs
    fm = (C.c_int * len(stages))(*[int(f) for f, _ in stages])
    ep = (C.c_double * len(stages))(*[float(e) for _, e in stages])
    x = np.empty(max(n, 1), np.float64)
    l1 = C.c_double()
    ed = C.c_uint64()
    tr = np.zeros((max_iter, 4), np.float64)
    it = lib().orc_apr_vary(n, len(cnbrs), np.ascontiguousarray(cindex, np.uint64),
                            np.ascontiguousarray(cnbrs, np.uint32), np.ascontiguousarray(outdeg, np.uint32),
                            fm, ep, len(stages), max_iter, eps2, x, threads or default_threads(), C.byref(l1),
                            C.byref(ed), tr.ctypes.data)
    if it < 0:
        raise ValueError("apr_vary: contribution format without a sign")
    return x[:n], it, l1.value, ed.value, tr[:it]


def kcore(index, nbrs):
    """Core numbers (KC_bucket's coreness) and the largest one."""
    n = len(index) - 1
    out = np.empty(max(n, 1), np.uint32)
    kmax = lib().orc_kcore(n, np.ascontiguousarray(index, np.uint64), np.ascontiguousarray(nbrs, np.uint32), out)
    return out[:n], int(kmax)


def pr_dense_fp64(index, nbrs, iters, damping=0.85):
    """Dense-matrix power iteration of the same update (tiny graphs only):
    P[d, s] = 1/outdeg(s) for each edge s->d; y = d*P*x; y += (1-sum y)/n;
    y /= sum y.  Independent of the CSR loop order."""
    n = len(index) - 1
    deg = np.diff(index).astype(np.float64)
    P = np.zeros((n, n))
    for s in range(n):
        for e in range(int(index[s]), int(index[s + 1])):
            P[int(nbrs[e]), s] += 1.0 / deg[s]
    x = np.full(n, 1.0 / n)
    for _ in range(iters):
        y = damping * (P @ x)
        y += (1.0 - y.sum()) / n
        y /= y.sum()
        x = y
    return x


def msbfs_levels(index, nbrs, sources):
    """Per-source BFS levels, shape (len(sources), n)."""
    return np.stack([bfs_levels(index, nbrs, int(s)) for s in sources])
