"""The GPU's contribution codecs, bit for bit against the oracle's.

The oracle's codecs (oracle/oracle.c: orc_cfp16_*, orc_bf16t_encode,
orc_bf16_rne_encode, orc_bf16_decode, orc_scfp16_*) are pinned exhaustively
in tests/test_oracle.py against the formats' mathematics (customfp.h:95-117
encode, :141-169 decode; scustomfp<true,8,7> and scustomfp<true,6,9,true,0>).
Here the engine writes contributions through gpr_pr_contrib
(contrib = damping * x / deg in FP32, PRv.C:176-187) into a 16-bit property
and the stored codes are compared with the oracle's encoders applied to the
same FP32 values; then every finite code is decoded by a PR edgemap whose
destinations each have one in-edge, and the sums compared with the oracle's
decoders.  Bars: bit-exact.

Inputs cover zero, FP32 subnormals, the customfp range edges 2^-63 and 2
(and values at or above 2, which alias as the reference's shifts do),
round-to-nearest-even ties of both parities for bf16, and 20000 random bit
patterns over [2^-70, 4) with random degrees.

Decode of scustomfp<true,6,9> code 0: the engine and the oracle read it as
0, as the reference's vector widening does for a format that may hold zero
(target/conversion.h:150-189); the scalar cvt_to_float (customfp.h:141-169
with B=0) would give 2^-63.  The delta PageRank stores 0 for inactive
vertices, so the choice is the one that keeps those sums exact.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()
    yield


def f32(bits):
    return np.asarray(bits, np.uint32).view(np.float32)


def special_values():
    v = [0x00000000, 0x00000001, 0x00000002, 0x007fffff, 0x00400000, 0x00800000,   # zero, subnormals, min normal
         0x20000000, 0x1fffffff, 0x20000001, 0x20002000, 0x20003fff,              # 2^-63 and around
         0x3f800000, 0x3fffffff, 0x3fffe000, 0x40000000, 0x40000001, 0x407fffff,  # 1, just below 2, 2, above
         0x3f808000, 0x3f818000, 0x3f807fff, 0x3f808001, 0x3f818001,              # bf16 ties (even / odd), near ties
         0x38d1b717, 0x33d6bf95, 0x2edbe6ff, 0x3c23d70a, 0x3dcccccd]              # 1e-4, 1e-7, 1e-10, 0.01, 0.1
    rng = np.random.default_rng(17)
    lo, hi = np.float32(2.0 ** -70), np.float32(4.0)
    r = rng.integers(int(lo.view(np.uint32)), int(hi.view(np.uint32)), 20000, dtype=np.uint64).astype(np.uint32)
    return np.concatenate([np.array(v, np.uint32), r])


def contrib_graph(deg):
    """A directed graph whose vertex v has out-degree deg[v] (edges to the
    next deg[v] vertices, wrapping)."""
    n = len(deg)
    index = np.zeros(n + 1, np.uint64)
    index[1:] = np.cumsum(deg)
    src = np.repeat(np.arange(n, dtype=np.int64), deg)
    off = np.arange(len(src), dtype=np.int64) - np.repeat(index[:-1].astype(np.int64), deg)
    nbrs = ((src + 1 + off) % n).astype(np.uint32)
    return index, nbrs


ENCODERS = {
    G.CONTRIB_BF16: O.bf16_rne_encode,
    G.CONTRIB_BF16T: O.bf16t_encode,
    G.CONTRIB_CFP16: O.cfp16_encode,
    G.CONTRIB_SCFP16: O.scfp16_encode_c,
}


def gpu_contrib(index, nbrs, x, fmt, damping):
    with G.Graph.from_csr(index, nbrs, symmetric=False) as g:
        xp = g.prop(G.F32)
        xp.upload(x)
        cp = g.prop(G.F32 if fmt == G.CONTRIB_F32 else G.U16)
        G._check(G.lib().gpr_pr_contrib(g.h, xp.h, cp.h, fmt, C.c_float(damping)))
        codes = cp.download()
        sc = C.c_float()
        G._check(G.lib().gpr_pr_scalars(g.h, None, None, None, C.byref(sc)))
        xp.destroy()
        cp.destroy()
        return codes, sc.value


@pytest.mark.parametrize("fmt", sorted(ENCODERS))
@pytest.mark.parametrize("signed", [False, True])
def test_encode_bit_exact_unit_degree(fmt, signed):
    """damping 1 and degree 1: the stored code is the encoder applied to x
    itself, so every planted bit pattern reaches the codec unchanged."""
    if signed and fmt == G.CONTRIB_CFP16:
        pytest.skip("customfp<6,10> has no sign; the reference stores only non-negative values in it")
    x = f32(special_values())
    if signed:
        x = np.where(np.arange(len(x)) % 2 == 1, -x, x).astype(np.float32)
    index, nbrs = contrib_graph(np.ones(len(x), np.int64))
    codes, _ = gpu_contrib(index, nbrs, x, fmt, 1.0)
    want = np.array([ENCODERS[fmt](float(v)) for v in x], np.uint16)
    bad = np.flatnonzero(codes != want)
    assert len(bad) == 0, [(hex(int(x.view(np.uint32)[i])), hex(int(codes[i])), hex(int(want[i]))) for i in bad[:8]]


@pytest.mark.parametrize("fmt", sorted(ENCODERS) + [G.CONTRIB_F32])
def test_encode_bit_exact_pagerank_values(fmt):
    """contrib = 0.85 * x / deg in FP32 for PageRank-like x and degrees from 1
    to 5000, then encoded: the same FP32 arithmetic in numpy, then the
    oracle's encoder."""
    rng = np.random.default_rng(5)
    k = 6000
    deg = np.concatenate([np.array([1, 2, 3, 7, 255, 256, 1023, 4999]), rng.integers(1, 60, k - 8)]).astype(np.int64)
    x = (rng.random(k) ** 6 * 1e-3).astype(np.float32)
    x[:50] = f32(special_values()[:50])
    x[50] = 0.0
    index, nbrs = contrib_graph(deg)
    codes, _ = gpu_contrib(index, nbrs, x, fmt, 0.85)
    cv = (np.float32(0.85) * x) / deg.astype(np.float32)
    if fmt == G.CONTRIB_F32:
        assert np.array_equal(codes.view(np.uint32), cv.view(np.uint32))
    else:
        want = np.array([ENCODERS[fmt](float(v)) for v in cv], np.uint16)
        assert np.array_equal(codes, want)


def test_encode_fp16_scaled_bit_exact():
    """FP16: the code is IEEE half (round to nearest even) of the FP32
    contribution times the iteration's power-of-two scale, chosen so that
    the largest contribution encodes at or below 2^15."""
    rng = np.random.default_rng(6)
    k = 6000
    deg = rng.integers(1, 40, k).astype(np.int64)
    x = (rng.random(k) ** 8 * 1e-4).astype(np.float32)
    x[:20] = 0.0
    x[20:30] = f32([1, 2, 0x00400000, 0x00800000, 0x20000000, 0x30000000, 0x33800000, 0x33000000, 0x337fffff, 5])
    index, nbrs = contrib_graph(deg)
    codes, scale = gpu_contrib(index, nbrs, x, G.CONTRIB_F16, 0.85)
    k2 = np.floor(np.log2(32768.0 / (0.85 * float(x.max()))))
    assert scale == 2.0 ** k2
    cv = (np.float32(0.85) * x) / deg.astype(np.float32)
    want = (cv * np.float32(scale)).astype(np.float16).view(np.uint16)
    assert np.array_equal(codes, want)
    assert (codes.view(np.float16).astype(np.float32) <= 32768).all()


DECODERS = {
    G.CONTRIB_BF16: O.bf16_decode,
    G.CONTRIB_BF16T: O.bf16_decode,
    G.CONTRIB_CFP16: O.cfp16_decode,
    G.CONTRIB_SCFP16: O.scfp16_decode_c,
}


@pytest.mark.parametrize("fmt", sorted(DECODERS))
def test_decode_every_code_bit_exact(fmt):
    """Source i carries code i and has one out-edge, to K + i: with every
    edge pulled, y[K + i] is the decoded value (a one-term FP32 sum, which
    reads -0 as +0)."""
    K = 65536
    codes = np.arange(K, dtype=np.uint32).astype(np.uint16)
    want = np.array([DECODERS[fmt](int(h)) for h in range(K)], np.float32)
    finite = np.isfinite(want)
    deg = np.zeros(2 * K, np.int64)
    deg[:K] = 1
    index, _ = contrib_graph(deg)
    nbrs = (K + np.arange(K)).astype(np.uint32)
    old = G.tune_get("pr_split_bytes")
    try:
        G.tune_set("pr_split_bytes", 1e12)   # pull every edge: y = dec(code) exactly
        with G.Graph.from_csr(index, nbrs, symmetric=False) as g:
            cp = g.prop(G.U16)
            cp.upload(np.concatenate([codes, np.zeros(K, np.uint16)]))
            y = g.prop(G.F32)
            g.edgemap(G.RELAX_PR_SUM, y, contrib=cp, contrib_fmt=fmt, damping=0.85)
            got = y.download()[K:]
            cp.destroy()
            y.destroy()
    finally:
        G.tune_set("pr_split_bytes", old)
    # a one-term sum from 0 turns -0 into +0; every other value bit for bit
    nz = finite & (want != 0)
    bad = np.flatnonzero(got.view(np.uint32)[nz] != want.view(np.uint32)[nz])
    assert len(bad) == 0, [(hex(int(h)), float(got[h]), float(want[h])) for h in np.flatnonzero(nz)[bad[:8]]]
    assert (got[finite & (want == 0)] == 0).all()
