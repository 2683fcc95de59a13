"""gpr_batch_cc_noalloc (the entry point that allocates nothing) against
gpr_batch_cc on the same batches: labels bit-exact, on two calls in a row,
the second with a smaller batch (the static scratch keeps the first call's
contents).  The inputs and outputs are device buffers this test allocates
through the HIP runtime; the entry point itself allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DevBuf:
    def __init__(self, arr):
        self.nbytes = max(arr.nbytes, 4)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), self.nbytes) == 0
        if arr.nbytes:
            a = np.ascontiguousarray(arr)
            assert hip().hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, D2H) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()
    yield


def run_noalloc(index, nbrs, voff, poison):
    n = int(voff[-1])
    bi, bn, bv = DevBuf(index), DevBuf(nbrs), DevBuf(voff)
    bl = DevBuf(np.full(n, poison, np.uint32))
    try:
        G._check(G.lib().gpr_batch_cc_noalloc(bi.p, bn.p, bv.p, len(voff) - 1, n, bl.p, None))
        assert hip().hipDeviceSynchronize() == 0
        return bl.get(np.uint32, n)
    finally:
        for b in (bi, bn, bv, bl):
            b.free()


def run_existing(index, nbrs, voff):
    with G.Batch.from_csr(index, nbrs, voff) as b:
        lab, _ = b.cc()
    return lab


def test_noalloc_matches_batch_cc_twice():
    # first call: small graphs plus two above the LDS capacity (the scratch
    # parents path); second call: a smaller batch of small graphs only
    first = O.batch_graphs(40, 1000, 20000, 5)
    big = O.batch_graphs(2, 70000, 90000, 6)
    index = np.concatenate([first[0], big[0][1:] + first[0][-1]])
    nbrs = np.concatenate([first[1], big[1] + np.uint32(first[2][-1])])
    voff = np.concatenate([first[2], big[2][1:] + first[2][-1]]).astype(np.uint32)
    assert voff[-1] <= 1 << 21
    second = O.batch_graphs(9, 500, 3000, 7)
    for (ix, nb, vo) in ((index, nbrs, voff), second):
        want = run_existing(ix, nb, vo)
        assert np.array_equal(want, O.cc_labels(ix, nb))
        got = run_noalloc(ix, nb, vo, 0xDEADBEEF)
        assert np.array_equal(got, want)


def test_noalloc_rejects_oversized_batch():
    index = np.zeros((1 << 21) + 2, np.uint64)
    voff = np.array([0, (1 << 21) + 1], np.uint32)
    bi, bv = DevBuf(index), DevBuf(voff)
    bl = DevBuf(np.zeros(4, np.uint32))
    try:
        rc = G.lib().gpr_batch_cc_noalloc(bi.p, bv.p, bv.p, 1, (1 << 21) + 1, bl.p, None)
        assert rc == G.GPR_E_INVAL
    finally:
        for b in (bi, bv, bl):
            b.free()
