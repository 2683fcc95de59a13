"""graptor_amd -- Python bindings of the gfx950 edgemap/vertexmap engine.

The engine is the C-ABI library ``graptor_amd/libgpr.so`` (declared in
``include/gpr.h``); this module loads it with ctypes.  There is no CPU
fallback: importing works without a GPU (so that the CPU test suite can
check the exported symbols), but every compute call requires the library
and a gfx950 device and raises otherwise.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpr.so")

GPR_OK, GPR_E_INVAL, GPR_E_FORMAT, GPR_E_OOM, GPR_E_HIP, GPR_E_IO, GPR_E_NODEV = range(7)
U32, F32, F64, U64, U16 = range(5)
FRONTIER_EMPTY, FRONTIER_ALL, FRONTIER_LIST, FRONTIER_BITMAP = range(4)
RELAX_MIN_PLUS1, RELAX_MIN, RELAX_PR_SUM, RELAX_SETIF = range(4)
UNSET = 0xFFFFFFFF   # an unset BFSv parent
DST_NONE, DST_GT, DST_NE = range(3)
RECORD_NONE, RECORD_REDUCTION, RECORD_METHOD_NE, RECORD_REDUCTION_OR_METHOD = range(4)
DIR_AUTO, DIR_PUSH, DIR_PULL, DIR_POLICY = range(4)
TRACE_FOLLOW = 0x100    # gpr_policy_trace: the unforced steps follow the loaded network
TRACE_PERSIST = 0x200   # gpr_policy_trace: the forced action holds to the end of the run
CONTRIB_F32, CONTRIB_BF16, CONTRIB_F16, CONTRIB_CFP16, CONTRIB_BF16T, CONTRIB_SCFP16 = range(6)
PR_SKIP_UNLINKED = 1
PR_SKIP_RANKS = 2
VMAP_FILL, VMAP_COPY, VMAP_IOTA, VMAP_FILL_FRONTIER, VMAP_COPY_FRONTIER = range(5)
CONTRIB_NAMES = {"fp32": CONTRIB_F32, "bf16": CONTRIB_BF16, "fp16": CONTRIB_F16, "cfp16": CONTRIB_CFP16,
                 "bf16t": CONTRIB_BF16T, "scfp16": CONTRIB_SCFP16}
# APR_vary's default variant 2046: FP32 contributions down to 1e-1, then
# scustomfp<true,6,9> down to 1e-7 (bench/APR_vary.C:37-39, :976-980).
APR_STAGES_2046 = (("fp32", 1e-1), ("scfp16", 1e-7))


class GprError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("gpr error %d: %s" % (code, msg))
        self.code = code


class EmapDesc(C.Structure):
    _fields_ = [("relax", C.c_int), ("prop", C.c_void_p), ("prev", C.c_void_p), ("src_strong", C.c_int),
                ("dst_filter", C.c_int), ("dst_arg", C.c_uint32), ("record", C.c_int), ("direction", C.c_int),
                ("out_kind", C.c_int), ("threshold_pct", C.c_float), ("contrib", C.c_void_p),
                ("contrib_fmt", C.c_int), ("damping", C.c_float), ("fusion", C.c_int)]


class IterStats(C.Structure):
    _fields_ = [("nactv_in", C.c_uint32), ("nacte_in", C.c_uint64), ("nactv_out", C.c_uint32),
                ("nacte_out", C.c_uint64), ("direction", C.c_int), ("out_kind", C.c_int),
                ("edges_scanned", C.c_uint64), ("ms", C.c_double), ("fused", C.c_int), ("levels", C.c_int)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RunOpts(C.Structure):
    _fields_ = [("direction", C.c_int), ("out_kind", C.c_int), ("threshold_pct", C.c_float),
                ("max_iter", C.c_int), ("tol", C.c_double), ("trace", C.c_int), ("fusion", C.c_int),
                ("deferred_update", C.c_int)]


class RunInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("push_iters", C.c_int), ("edges", C.c_uint64), ("ms", C.c_double),
                ("delta", C.c_double), ("sum", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class PrStage(C.Structure):
    _fields_ = [("contrib", C.c_int), ("tol", C.c_double)]


_P = C.c_void_p
_PP = C.POINTER(C.c_void_p)
# name -> (restype, argtypes); every entry point declared in include/gpr.h
SIGNATURES = {
    "gpr_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
    "gpr_version": (C.c_char_p, []),
    "gpr_init": (C.c_int, [C.c_int]),
    "gpr_graph_create": (C.c_int, [_P, _P, C.c_uint32, C.c_uint64, C.c_int, _PP]),
    "gpr_graph_load_b2": (C.c_int, [C.c_char_p, C.c_int, _PP]),
    "gpr_graph_save_b2": (C.c_int, [_P, C.c_char_p]),
    "gpr_graph_generate_rmat": (C.c_int, [C.c_uint32, C.c_uint32, C.c_double, C.c_double, C.c_double,
                                          C.c_uint64, _PP]),
    "gpr_graph_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    "gpr_graph_download": (C.c_int, [_P, _P, _P]),
    "gpr_graph_reorder": (C.c_int, [_P, C.c_int]),
    "gpr_graph_order": (C.c_int, [_P, _P]),
    "gpr_graph_remap_id": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "gpr_graph_original_id": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint32)]),
    "gpr_cc_recode": (C.c_int, [_P, _P]),
    "gpr_graph_highest_degree_vertex": (C.c_int, [_P, C.POINTER(C.c_uint32)]),
    "gpr_graph_destroy": (C.c_int, [_P]),
    "gpr_prop_create": (C.c_int, [_P, C.c_int, _PP]),
    "gpr_prop_upload": (C.c_int, [_P, _P]),
    "gpr_prop_download": (C.c_int, [_P, _P]),
    "gpr_prop_destroy": (C.c_int, [_P]),
    "gpr_frontier_create": (C.c_int, [_P, _PP]),
    "gpr_frontier_set_single": (C.c_int, [_P, C.c_uint32]),
    "gpr_frontier_set_all": (C.c_int, [_P]),
    "gpr_frontier_set_list": (C.c_int, [_P, _P, C.c_uint32]),
    "gpr_frontier_set_neighbours": (C.c_int, [_P, C.c_uint32]),
    "gpr_frontier_stats": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.POINTER(C.c_int)]),
    "gpr_frontier_convert": (C.c_int, [_P, C.c_int]),
    "gpr_frontier_download": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "gpr_frontier_destroy": (C.c_int, [_P]),
    "gpr_edgemap": (C.c_int, [_P, C.POINTER(EmapDesc), _P, _P, C.POINTER(IterStats)]),
    "gpr_vertex_map": (C.c_int, [_P, C.c_int, _P, _P, _P, C.c_uint64]),
    "gpr_vertex_filter": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_int, _P]),
    "gpr_pr_iterate_fused": (C.c_int, [_P, _P, _P, C.c_int, C.c_float, C.c_int, _P]),
    "gpr_pr_contrib": (C.c_int, [_P, _P, _P, C.c_int, C.c_float]),
    "gpr_pr_finish": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_float, C.POINTER(C.c_double),
                                C.POINTER(C.c_double)]),
    "gpr_pr_scalars": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_float)]),
    "gpr_bfslvl": (C.c_int, [_P, C.c_uint32, C.POINTER(RunOpts), _P, C.POINTER(RunInfo)]),
    "gpr_bfsv": (C.c_int, [_P, C.c_uint32, C.POINTER(RunOpts), _P, C.POINTER(RunInfo)]),
    "gpr_ccv": (C.c_int, [_P, C.POINTER(RunOpts), _P, C.POINTER(RunInfo)]),
    "gpr_prv": (C.c_int, [_P, C.POINTER(RunOpts), _P, C.POINTER(RunInfo)]),
    "gpr_pr_vary": (C.c_int, [_P, C.POINTER(PrStage), C.c_int, C.POINTER(RunOpts), _P, C.POINTER(RunInfo)]),
    "gpr_kcore": (C.c_int, [_P, C.POINTER(RunOpts), _P, C.POINTER(C.c_uint32), C.POINTER(RunInfo)]),
    "gpr_apr_vary": (C.c_int, [_P, C.POINTER(PrStage), C.c_int, C.c_double, C.POINTER(RunOpts), _P,
                               C.POINTER(RunInfo)]),
    "gpr_batch_cc_noalloc": (C.c_int, [_P, _P, _P, C.c_uint32, C.c_uint32, _P, _P]),
    "gpr_msbfs": (C.c_int, [_P, _P, C.c_int, C.POINTER(RunOpts), _P, C.POINTER(RunInfo)]),
    "gpr_batch_generate": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint64, _PP]),
    "gpr_batch_create": (C.c_int, [_P, _P, _P, C.c_uint32, _PP]),
    "gpr_batch_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint64)]),
    "gpr_batch_download": (C.c_int, [_P, _P, _P, _P]),
    "gpr_batch_cc": (C.c_int, [_P, _P, C.POINTER(RunInfo)]),
    "gpr_batch_destroy": (C.c_int, [_P]),
    "gpr_mlp_param_count": (C.c_int, []),
    "gpr_policy_set": (C.c_int, [_P, _P, C.c_int]),
    "gpr_policy_reset": (C.c_int, [_P]),
    "gpr_policy_trace": (C.c_int, [_P, C.c_int, C.c_uint32, C.c_int, _P, _P, C.POINTER(C.c_int)]),
    "gpr_mlp_forward": (C.c_int, [_P, _P, C.c_int, _P]),
    "gpr_mlp_grad": (C.c_int, [_P, _P, _P, _P, C.c_int, _P, C.POINTER(C.c_float)]),
    "gpr_mlp_train": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    "gpr_trace_get": (C.c_int, [_P, C.POINTER(IterStats), C.c_int, C.POINTER(C.c_int)]),
    "gpr_tune_set": (C.c_int, [C.c_char_p, C.c_double]),
    "gpr_tune_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_double)]),
    "gpr_profile_enable": (C.c_int, [_P, C.c_int]),
    "gpr_profile_get": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "gpr_profile_reset": (C.c_int, [_P]),
}

_lib = None
_inited = False


def load_library(path=LIB_PATH):
    """Load libgpr.so (no device initialisation).  Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise GprError(GPR_E_NODEV, "%s not built: run __graft_entry__.build()" % path)
        L = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


TUNED = os.path.join(_HERE, "tuned.json")


def lib():
    """The library with the gfx950 device initialised and the autotuned
    parameters (graptor_amd/tuned.json, when present) applied."""
    global _inited
    L = load_library()
    if not _inited:
        _check(L.gpr_init(0), L)
        _inited = True
        if os.path.exists(TUNED) and not os.environ.get("GPR_NO_TUNED"):
            import json
            with open(TUNED) as f:
                for k, v in json.load(f).get("tunables", {}).items():
                    _check(L.gpr_tune_set(k.encode(), float(v)), L)
    return L


def tune_set(name, value):
    _check(lib().gpr_tune_set(name.encode(), float(value)))


def tune_get(name):
    v = C.c_double()
    _check(lib().gpr_tune_get(name.encode(), C.byref(v)))
    return v.value


def _check(rc, L=None):
    if rc != GPR_OK:
        L = L or _lib
        buf = C.create_string_buffer(2048)
        L.gpr_last_error(buf, len(buf))
        raise GprError(rc, buf.value.decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


FUSION_OFF, FUSION_ON, FUSION_DEFAULT, FUSION_SMALL = range(4)


def opts(direction=DIR_AUTO, out_kind=0, threshold_pct=0.0, max_iter=100, tol=1e-7, trace=False, fusion=0,
         deferred_update=0):
    """Run options.  fusion: 0 = the benchmark's reference default, -1 = off,
    1 = Graptor's FUSION 1 rule, 2 = every sparse step fused, 3 = small
    levels fused (GPR_FUSION_SMALL).
    deferred_update: 1 = the benchmarks' DEFERRED_UPDATE=1 record forms
    (method / reduction_or_method with a maintained previous copy)."""
    return RunOpts(direction, out_kind, threshold_pct, max_iter, tol, 1 if trace else 0, fusion,
                   1 if deferred_update else 0)


class Graph:
    """A device-resident graph (CSR + CSC + pull partitions)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        n = C.c_uint32()
        m = C.c_uint64()
        s = C.c_int()
        _check(lib().gpr_graph_info(self.h, C.byref(n), C.byref(m), C.byref(s)))
        self.n, self.m, self.symmetric = n.value, m.value, bool(s.value)

    @classmethod
    def from_csr(cls, index, nbrs, symmetric=True):
        index = np.ascontiguousarray(index, np.uint64)
        nbrs = np.ascontiguousarray(nbrs, np.uint32)
        h = C.c_void_p()
        _check(lib().gpr_graph_create(_ptr(index), _ptr(nbrs) if len(nbrs) else None, len(index) - 1, len(nbrs),
                                      1 if symmetric else 0, C.byref(h)))
        return cls(h)

    @classmethod
    def load_b2(cls, path, symmetric=True):
        h = C.c_void_p()
        _check(lib().gpr_graph_load_b2(path.encode(), 1 if symmetric else 0, C.byref(h)))
        return cls(h)

    @classmethod
    def rmat(cls, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=0x47524150):
        h = C.c_void_p()
        _check(lib().gpr_graph_generate_rmat(scale, edge_factor, a, b, c, seed, C.byref(h)))
        return cls(h)

    def save_b2(self, path):
        _check(lib().gpr_graph_save_b2(self.h, path.encode()))

    def download(self):
        """CSR in internal IDs (see order())."""
        index = np.empty(self.n + 1, np.uint64)
        nbrs = np.empty(max(1, self.m), np.uint32)
        _check(lib().gpr_graph_download(self.h, _ptr(index), _ptr(nbrs)))
        return index, nbrs[: self.m]

    def reorder(self, mode=1):
        """Relabel vertices internally by decreasing degree (IDs at the API
        stay original)."""
        _check(lib().gpr_graph_reorder(self.h, mode))
        return self

    def order(self):
        """order[i] = original ID of internal vertex i."""
        o = np.empty(max(1, self.n), np.uint32)
        _check(lib().gpr_graph_order(self.h, _ptr(o)))
        return o[: self.n]

    def remap_id(self, v):
        r = C.c_uint32()
        _check(lib().gpr_graph_remap_id(self.h, v, C.byref(r)))
        return r.value

    def original_id(self, v):
        r = C.c_uint32()
        _check(lib().gpr_graph_original_id(self.h, v, C.byref(r)))
        return r.value

    def highest_degree_vertex(self):
        v = C.c_uint32()
        _check(lib().gpr_graph_highest_degree_vertex(self.h, C.byref(v)))
        return v.value

    def destroy(self):
        if self.h:
            _check(lib().gpr_graph_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()

    # ---- applications -------------------------------------------------------
    # Each run returns its result array, or None with download=False: the
    # result then stays on the device, as a benchmark's timed run() leaves it
    # (driver.C:245-255 times run() only).
    def bfslvl(self, start, download=True, **kw):
        out = np.empty(self.n, np.uint32) if download else None
        o = opts(**kw)
        info = RunInfo()
        _check(lib().gpr_bfslvl(self.h, start, C.byref(o), _ptr(out) if download else None, C.byref(info)))
        return out, info

    def bfsv(self, start, download=True, **kw):
        """BFS parents (original IDs): the start its own parent, UNSET where
        unreached."""
        out = np.empty(self.n, np.uint32) if download else None
        o = opts(**kw)
        info = RunInfo()
        _check(lib().gpr_bfsv(self.h, start, C.byref(o), _ptr(out) if download else None, C.byref(info)))
        return out, info

    def ccv(self, download=True, **kw):
        out = np.empty(self.n, np.uint32) if download else None
        o = opts(**kw)
        info = RunInfo()
        _check(lib().gpr_ccv(self.h, C.byref(o), _ptr(out) if download else None, C.byref(info)))
        return out, info

    def prv(self, max_iter=20, tol=0.0, download=True, **kw):
        out = np.empty(self.n, np.float32) if download else None
        o = opts(max_iter=max_iter, tol=tol, **kw)
        info = RunInfo()
        _check(lib().gpr_prv(self.h, C.byref(o), _ptr(out) if download else None, C.byref(info)))
        return out, info

    def pr_vary(self, stages, max_iter=100, download=True, **kw):
        """stages: list of (format name or CONTRIB_*, tol)."""
        arr = (PrStage * len(stages))(*[PrStage(CONTRIB_NAMES.get(f, f), t) for f, t in stages])
        out = np.empty(self.n, np.float32) if download else None
        o = opts(max_iter=max_iter, **kw)
        info = RunInfo()
        _check(lib().gpr_pr_vary(self.h, arr, len(stages), C.byref(o), _ptr(out) if download else None,
                                 C.byref(info)))
        return out, info

    def apr_vary(self, stages=APR_STAGES_2046, eps2=1e-2, max_iter=100, download=True, **kw):
        """Delta PageRank (APR_vary): stages of (format name or CONTRIB_*, eps);
        returns (x as FP64 scaled to unit L1 norm, RunInfo)."""
        arr = (PrStage * len(stages))(*[PrStage(CONTRIB_NAMES.get(f, f), t) for f, t in stages])
        out = np.empty(self.n, np.float64) if download else None
        o = opts(max_iter=max_iter, **kw)
        info = RunInfo()
        _check(lib().gpr_apr_vary(self.h, arr, len(stages), eps2, C.byref(o), _ptr(out) if download else None,
                                  C.byref(info)))
        return out, info

    def kcore(self, download=True, **kw):
        """Core numbers (KC_bucket's coreness); returns (coreness, largest
        core, RunInfo)."""
        out = np.empty(self.n, np.uint32) if download else None
        o = opts(**kw)
        info = RunInfo()
        kmax = C.c_uint32()
        _check(lib().gpr_kcore(self.h, C.byref(o), _ptr(out) if download else None, C.byref(kmax), C.byref(info)))
        return out, kmax.value, info

    def msbfs(self, sources, **kw):
        """Levels from up to 64 sources in one traversal, shape (k, n)."""
        src = np.ascontiguousarray(sources, np.uint32)
        out = np.empty((len(src), self.n), np.uint32)
        o = opts(**kw)
        info = RunInfo()
        _check(lib().gpr_msbfs(self.h, _ptr(src), len(src), C.byref(o), _ptr(out), C.byref(info)))
        return out, info

    def set_policy(self, params):
        """Load the direction/representation network (None: threshold rule)."""
        if params is None:
            _check(lib().gpr_policy_set(self.h, None, 0))
            return
        p = np.ascontiguousarray(params, np.float32)
        _check(lib().gpr_policy_set(self.h, _ptr(p), len(p)))

    def policy_trace(self, algo, start=0, cap=64, follow=False, persist=False):
        """Features (k,6) of up to `cap` evenly spread states of a BFS
        (algo 0, from `start`) or CC (algo 1) run, and the ms to finish the
        run from each state under each of the four actions (k,4), every
        other step taking the rule's action (follow=True: the loaded
        network's; persist=True: the forced action holds to the end)."""
        f = np.empty((cap, 6), np.float32)
        t = np.empty((cap, 4), np.float32)
        k = C.c_int()
        flags = (TRACE_FOLLOW if follow else 0) | (TRACE_PERSIST if persist else 0)
        _check(lib().gpr_policy_trace(self.h, algo | flags, start, cap, _ptr(f), _ptr(t), C.byref(k)))
        return f[: k.value].copy(), t[: k.value].copy()

    def trace(self):
        k = C.c_int()
        _check(lib().gpr_trace_get(self.h, None, 0, C.byref(k)))
        buf = (IterStats * max(1, k.value))()
        _check(lib().gpr_trace_get(self.h, buf, k.value, C.byref(k)))
        return [buf[i].as_dict() for i in range(k.value)]

    def profile(self, on=True):
        _check(lib().gpr_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self):
        _check(lib().gpr_profile_reset(self.h))

    def profile_get(self, name):
        ms = C.c_double()
        cnt = C.c_uint64()
        _check(lib().gpr_profile_get(self.h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    # ---- engine-level access --------------------------------------------------
    def prop(self, dtype=U32):
        return Prop(self, dtype)

    def frontier(self):
        return Frontier(self)

    def edgemap(self, relax, prop, frontier_in=None, frontier_out=None, dst_filter=DST_NONE, dst_arg=0,
                record=RECORD_REDUCTION, direction=DIR_AUTO, out_kind=0, threshold_pct=0.0, prev=None,
                contrib=None, contrib_fmt=CONTRIB_F32, damping=0.85, fusion=FUSION_OFF):
        d = EmapDesc(relax, prop.h, prev.h if prev else None, 0, dst_filter, dst_arg, record, direction, out_kind,
                     threshold_pct, contrib.h if contrib else None, contrib_fmt, damping, fusion)
        st = IterStats()
        _check(lib().gpr_edgemap(self.h, C.byref(d), frontier_in.h if frontier_in else None,
                                 frontier_out.h if frontier_out else None, C.byref(st)))
        return st.as_dict()


class Batch:
    """Thousands of small graphs as one CSR with vertex offsets."""

    def __init__(self, handle):
        self.h = handle
        c, n, m = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(lib().gpr_batch_info(self.h, C.byref(c), C.byref(n), C.byref(m)))
        self.count, self.n, self.m = c.value, n.value, m.value

    @classmethod
    def generate(cls, count=4096, nmin=1024, nmax=65536, seed=0x47524150):
        h = C.c_void_p()
        _check(lib().gpr_batch_generate(count, nmin, nmax, seed, C.byref(h)))
        return cls(h)

    @classmethod
    def from_csr(cls, index, nbrs, voff):
        index = np.ascontiguousarray(index, np.uint64)
        nbrs = np.ascontiguousarray(nbrs, np.uint32)
        voff = np.ascontiguousarray(voff, np.uint32)
        h = C.c_void_p()
        _check(lib().gpr_batch_create(_ptr(index), _ptr(nbrs) if len(nbrs) else None, _ptr(voff), len(voff) - 1,
                                      C.byref(h)))
        return cls(h)

    def download(self):
        index = np.empty(self.n + 1, np.uint64)
        nbrs = np.empty(max(1, self.m), np.uint32)
        voff = np.empty(self.count + 1, np.uint32)
        _check(lib().gpr_batch_download(self.h, _ptr(index), _ptr(nbrs), _ptr(voff)))
        return index, nbrs[: self.m], voff

    def cc(self):
        out = np.empty(max(1, self.n), np.uint32)
        info = RunInfo()
        _check(lib().gpr_batch_cc(self.h, _ptr(out), C.byref(info)))
        return out[: self.n], info

    def destroy(self):
        if self.h:
            _check(lib().gpr_batch_destroy(self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()


_NP = {U32: np.uint32, F32: np.float32, F64: np.float64, U64: np.uint64, U16: np.uint16}


class Prop:
    def __init__(self, g, dtype):
        self.g, self.dtype = g, dtype
        self.h = C.c_void_p()
        _check(lib().gpr_prop_create(g.h, dtype, C.byref(self.h)))

    def upload(self, arr):
        a = np.ascontiguousarray(arr, _NP[self.dtype])
        assert a.size == self.g.n
        _check(lib().gpr_prop_upload(self.h, _ptr(a)))

    def download(self):
        a = np.empty(self.g.n, _NP[self.dtype])
        _check(lib().gpr_prop_download(self.h, _ptr(a)))
        return a

    def vertex_map(self, op, value_bits=0, other=None, frontier=None):
        _check(lib().gpr_vertex_map(self.g.h, op, self.h, other.h if other else None,
                                    frontier.h if frontier else None, value_bits))

    def destroy(self):
        if self.h:
            _check(lib().gpr_prop_destroy(self.h))
            self.h = None


class Frontier:
    def __init__(self, g):
        self.g = g
        self.h = C.c_void_p()
        _check(lib().gpr_frontier_create(g.h, C.byref(self.h)))

    def set_single(self, v):
        _check(lib().gpr_frontier_set_single(self.h, v))
        return self

    def set_all(self):
        _check(lib().gpr_frontier_set_all(self.h))
        return self

    def set_list(self, vs):
        a = np.ascontiguousarray(vs, np.uint32)
        _check(lib().gpr_frontier_set_list(self.h, _ptr(a) if len(a) else None, len(a)))
        return self

    def set_neighbours(self, v):
        _check(lib().gpr_frontier_set_neighbours(self.h, v))
        return self

    def stats(self):
        v = C.c_uint32()
        e = C.c_uint64()
        k = C.c_int()
        _check(lib().gpr_frontier_stats(self.h, C.byref(v), C.byref(e), C.byref(k)))
        return v.value, e.value, k.value

    def convert(self, kind):
        _check(lib().gpr_frontier_convert(self.h, kind))

    def filter(self, prop, value, neg=False, out=None):
        """vertex_filter: the vertices of this frontier with prop[v] == value
        (!= with neg), into `out` (a new frontier when None)."""
        out = out or Frontier(self.g)
        _check(lib().gpr_vertex_filter(self.g.h, self.h, prop.h, value, 1 if neg else 0, out.h))
        return out

    def download(self):
        nactv = self.stats()[0]
        out = np.empty(max(1, nactv), np.uint32)
        k = C.c_uint32()
        _check(lib().gpr_frontier_download(self.h, _ptr(out), C.byref(k)))
        return out[: k.value]

    def destroy(self):
        if self.h:
            _check(lib().gpr_frontier_destroy(self.h))
            self.h = None


def mlp_param_count():
    return load_library().gpr_mlp_param_count()


def mlp_forward(params, x):
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    out = np.empty((len(x), 4), np.float32)
    _check(lib().gpr_mlp_forward(_ptr(p), _ptr(x), len(x), _ptr(out)))
    return out


def mlp_grad(params, x, targets, weights=None):
    x = np.ascontiguousarray(x, np.float32)
    p = np.ascontiguousarray(params, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    g = np.empty(len(p), np.float32)
    loss = C.c_float()
    _check(lib().gpr_mlp_grad(_ptr(p), _ptr(x), _ptr(t), _ptr(w) if w is not None else None, len(x), _ptr(g),
                              C.byref(loss)))
    return g, loss.value


def mlp_train(params, x, targets, weights=None, epochs=500, lr=1e-2):
    x = np.ascontiguousarray(x, np.float32)
    p = np.array(params, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    loss = C.c_float()
    _check(lib().gpr_mlp_train(_ptr(p), _ptr(x), _ptr(t), _ptr(w) if w is not None else None, len(x), epochs, lr,
                               C.byref(loss)))
    return p, loss.value


def version():
    return load_library().gpr_version().decode()
