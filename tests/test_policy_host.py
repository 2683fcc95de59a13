"""CPU tests of the learned policy's host-side pipeline (graptor_amd/policy.py):
trace labels and weights, the gate, the synthetic training grids and the
committed network file.  No GPU call is made."""
import json

import numpy as np
import pytest

import graptor_amd as G
import graptor_amd.policy as P


def test_targets_take_the_rule_unless_the_saving_is_clear():
    # rule time to the end, then actions 1..3
    times = np.array([
        [10.0, 9.0, 12.0, 11.0],     # action 1 saves 10% and 1 ms: taken
        [10.0, 9.97, 12.0, 11.0],    # 0.3% < 0.5%: the rule
        [1.0, 0.99, 1.2, 1.1],       # 1% but 10 us < 15 us: the rule
        [2.0, 2.0, 2.0, 2.0],        # a tie: the rule
        [4.0, 5.0, 3.9, 3.0],        # action 3 saves 25% and 1 ms: taken
        [0.5, 0.6, 0.48, 0.7],       # 4% and 20 us: action 2
    ], np.float32)
    T, W, lab = P.targets_from_times(times)
    assert lab.tolist() == [1, 0, 0, 0, 3, 2]
    assert np.allclose(T.sum(1), 1.0)
    assert np.allclose(T[np.arange(len(lab)), lab], 0.95)
    off = np.ones_like(T, bool)
    off[np.arange(len(lab)), lab] = False
    assert np.allclose(T[off], 0.05 / 3)
    # sample weight: the time at stake, plus a floor
    assert np.allclose(W, times.max(1) - times.min(1) + 1e-2)


def test_targets_thresholds_are_parameters():
    times = np.array([[10.0, 9.9, 11.0, 11.0]], np.float32)   # 1%, 0.1 ms
    assert P.targets_from_times(times, gain=0.005, floor_ms=0.015)[2].tolist() == [1]
    assert P.targets_from_times(times, gain=0.02, floor_ms=0.015)[2].tolist() == [0]
    assert P.targets_from_times(times, gain=0.005, floor_ms=0.2)[2].tolist() == [0]


def test_gate_needs_the_margin_over_the_rule():
    z = np.array([
        [0.0, 1.0, 0.5, 0.0],    # action 1 leads the rule by 1.0
        [1.0, 1.0, 0.0, 0.0],    # a tie with the rule: the rule
        [0.0, -1.0, -2.0, -0.5],  # the rule is the argmax
        [0.2, 0.0, 0.0, 0.7],    # action 3 leads by 0.5
    ], np.float32)
    assert P.gated(z, 0.0).tolist() == [1, 0, 0, 3]
    assert P.gated(z, 0.5).tolist() == [1, 0, 0, 3]
    assert P.gated(z, 0.6).tolist() == [1, 0, 0, 0]
    assert P.gated(z, 2.0).tolist() == [0, 0, 0, 0]


def csr_edges(index, nbrs):
    src = np.repeat(np.arange(len(index) - 1), np.diff(index.astype(np.int64)))
    return set(zip(src.tolist(), nbrs.tolist()))


def test_grid_without_rewiring_is_the_4_neighbour_grid():
    w, h = 5, 4
    index, nbrs = P.grid_csr(w, h)
    n = w * h
    assert len(index) == n + 1 and index[0] == 0 and index[-1] == len(nbrs)
    assert len(nbrs) == 2 * ((w - 1) * h + w * (h - 1))
    want = set()
    for y in range(h):
        for x in range(w):
            v = y * w + x
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                if 0 <= x + dx < w and 0 <= y + dy < h:
                    want.add((v, (y + dy) * w + x + dx))
    assert csr_edges(index, nbrs) == want
    deg = np.diff(index.astype(np.int64)).reshape(h, w)
    assert deg[0, 0] == deg[-1, -1] == 2 and deg[1, 1] == 4 and deg[0, 2] == 3
    for v in range(n):   # rows sorted
        row = nbrs[index[v]:index[v + 1]]
        assert np.all(np.diff(row.astype(np.int64)) > 0)


@pytest.mark.parametrize("rewire,long_frac", [(0.3, 0.0), (0.0, 0.5), (0.1, 0.2)])
def test_grid_with_random_edges_stays_a_simple_symmetric_graph(rewire, long_frac):
    w, h = 16, 12
    index, nbrs = P.grid_csr(w, h, rewire=rewire, long_frac=long_frac, seed=5)
    e = csr_edges(index, nbrs)
    assert all((b, a) in e for a, b in e)
    assert all(a != b for a, b in e)
    assert len(e) == len(nbrs)   # no duplicates
    base = csr_edges(*P.grid_csr(w, h))
    if rewire == 0.0:
        assert base <= e and len(e) > len(base)
    else:
        assert e != base
    again = P.grid_csr(w, h, rewire=rewire, long_frac=long_frac, seed=5)
    assert np.array_equal(again[0], index) and np.array_equal(again[1], nbrs)


def test_rmat_parameters_keep_the_benchmark_proportions():
    for a in (0.45, 0.57, 0.65):
        pa, pb, pc = P.rmat_abc(a)
        assert pa == a and pb == pc
        assert abs((1 - a - pb - pc) / pb - 0.05 / 0.19) < 1e-12


def test_initial_parameters_are_seeded_and_scaled():
    k = G.mlp_param_count()
    p0, p0b, p1 = P.init_params(0), P.init_params(0), P.init_params(1)
    assert p0.shape == (k,) and np.array_equal(p0, p0b) and not np.array_equal(p0, p1)
    # W1 (6 x 32), b1, W2 (32 x 32), b2, W3 (32 x 4), b3; biases start at 0
    o = 0
    for fan_in, fan_out in ((6, 32), (32, 32), (32, 4)):
        wts = p0[o:o + fan_in * fan_out]
        assert 0.5 * np.sqrt(2.0 / fan_in) < wts.std() < 1.5 * np.sqrt(2.0 / fan_in)
        o += fan_in * fan_out
        assert np.all(p0[o:o + fan_out] == 0)
        o += fan_out
    assert o == k


def test_committed_network_and_its_record():
    w = P.load()
    assert w.shape == (G.mlp_param_count() + 1,) and np.all(np.isfinite(w)) and w[-1] >= 0
    with open(P.WEIGHTS) as fh:
        meta = json.load(fh)["meta"]
    # the training run's own held-out evaluation met the bar the GPU test
    # asserts, with every output equal to the rule's
    ho = meta["held_out"]
    assert ho["geomean_speedup"] >= 1.10 and ho["worst_speedup"] >= 0.97 and ho["all_bit_exact"]
    assert meta["margin"] == w[-1]
    assert sum(meta["labels"]) == meta["samples"]
    assert set(meta["validation"]) == {x for x, _ in P.training_graphs() if x.endswith("/val")}
