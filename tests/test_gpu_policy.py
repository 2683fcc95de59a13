"""GPU tests of the learned direction/representation policy.

  * the MLP forward pass and gradients computed by the HIP kernels equal a
    float64 numpy restatement (the CPU oracle of the network), and a central
    finite difference of the loss confirms that restatement;
  * Adam training on the GPU lowers the loss;
  * BFS and CC stay bit-exact whatever a (random) network chooses, with the
    choice made by the fused in-kernel inference; forced networks are
    obeyed, and the gate keeps the rule's step below the margin;
  * traces record the time to finish the run under each of the four
    actions.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
G = pytest.importorskip("graptor_amd")

IN, H, OUT = 6, 32, 4


@pytest.fixture(scope="module", autouse=True)
def device():
    G.lib()
    yield


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * OUT].reshape(H, OUT); o += H * OUT
    b3 = p[o:o + OUT]; o += OUT
    assert o == len(p)
    return W1, b1, W2, b2, W3, b3


def np_forward(p, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    A1 = X @ W1 + b1
    H1 = np.maximum(A1, 0)
    A2 = H1 @ W2 + b2
    H2 = np.maximum(A2, 0)
    return H1, H2, H2 @ W3 + b3


def np_loss_grad(p, X, T, w):
    X = X.astype(np.float64)
    H1, H2, Z = np_forward(p, X)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    Zs = Z - Z.max(1, keepdims=True)
    P = np.exp(Zs) / np.exp(Zs).sum(1, keepdims=True)
    lse = np.log(np.exp(Zs).sum(1, keepdims=True)) + Z.max(1, keepdims=True)
    ws = w.sum()
    loss = float((w[:, None] * -T * (Z - lse)).sum() / ws)
    dZ = w[:, None] * (P - T) / ws
    gW3 = H2.T @ dZ
    gb3 = dZ.sum(0)
    dH2 = (dZ @ W3.T) * (H2 > 0)
    gW2 = H1.T @ dH2
    gb2 = dH2.sum(0)
    dH1 = (dH2 @ W2.T) * (H1 > 0)
    gW1 = X.T @ dH1
    gb1 = dH1.sum(0)
    return loss, np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3])


def random_problem(n=300, seed=0):
    rng = np.random.default_rng(seed)
    k = G.mlp_param_count()
    p = (rng.standard_normal(k) * 0.3).astype(np.float32)
    X = rng.random((n, IN)).astype(np.float32)
    T = rng.random((n, OUT)).astype(np.float32)
    T /= T.sum(1, keepdims=True)
    w = (rng.random(n) + 0.1).astype(np.float32)
    return p, X, T, w


def test_param_count():
    assert G.mlp_param_count() == IN * H + H + H * H + H + H * OUT + OUT


def test_forward_matches_numpy():
    p, X, _, _ = random_problem()
    Z = G.mlp_forward(p, X)
    _, _, Zr = np_forward(p, X.astype(np.float64))
    assert np.allclose(Z, Zr, rtol=1e-5, atol=1e-5)


def test_gradient_matches_numpy_and_finite_differences():
    p, X, T, w = random_problem(n=517)   # not a multiple of the workgroup size
    g, loss = G.mlp_grad(p, X, T, w)
    lr, gr = np_loss_grad(p.astype(np.float64), X, T.astype(np.float64), w.astype(np.float64))
    assert abs(loss - lr) < 1e-4 * max(1.0, abs(lr))
    assert np.allclose(g, gr, rtol=1e-3, atol=2e-6)
    # the numpy restatement itself against central differences
    rng = np.random.default_rng(1)
    pd = p.astype(np.float64)
    for i in rng.choice(len(p), 25, replace=False):
        e = np.zeros_like(pd)
        e[i] = 1e-6
        fd = (np_loss_grad(pd + e, X, T.astype(np.float64), w.astype(np.float64))[0]
              - np_loss_grad(pd - e, X, T.astype(np.float64), w.astype(np.float64))[0]) / 2e-6
        assert abs(fd - gr[i]) < 1e-5 + 1e-3 * abs(gr[i])


def test_training_lowers_loss():
    p, X, T, w = random_problem(n=800, seed=3)
    _, l0 = G.mlp_grad(p, X, T, w)
    p2, l1 = G.mlp_train(p, X, T, w, epochs=300, lr=1e-2)
    _, l2 = G.mlp_grad(p2, X, T, w)
    assert l2 < l0 and abs(l1 - l2) < 0.05 * max(l2, 1e-3) + 1e-3


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_policy_keeps_results_bit_exact(seed):
    p, _, _, _ = random_problem(seed=seed)
    for name, (index, nbrs) in (("rmat14", O.rmat_graph(14)), ("grid", O.grid_graph(50, 40, 0.01, seed=2))):
        s = O.highest_degree_vertex(index)
        with G.Graph.from_csr(index, nbrs) as g:
            g.set_policy(p)
            lv, info = g.bfslvl(s, direction=G.DIR_POLICY, trace=True)
            assert np.array_equal(lv, O.bfs_levels(index, nbrs, s)), name
            tr = g.trace()
            assert len(tr) == info.iters
            lab, _ = g.ccv(direction=G.DIR_POLICY)
            assert np.array_equal(lab, O.cc_labels(index, nbrs)), name
            g.set_policy(None)     # back to the threshold rule
            lv2, _ = g.bfslvl(s, direction=G.DIR_POLICY)
            assert np.array_equal(lv2, lv)


def pushes_where_allowed(tr, rule, m):
    """Action 1's steps: an unfused push where the rule pushes or the
    frontier's vertices + out-edges are at most twice the rule's threshold
    (default_threshold: m / 20), the rule's pull elsewhere."""
    assert len(tr) == len(rule)
    for t, r in zip(tr, rule):
        if r[0] == G.DIR_PUSH or t["nactv_in"] + t["nacte_in"] <= 2 * (m // 20):
            assert t["direction"] == G.DIR_PUSH and not t["fused"], (t, r)
        else:
            assert t["direction"] == G.DIR_PULL, (t, r)


def test_policy_choice_follows_network():
    """A network whose output bias forces one action is obeyed each step:
    0 the threshold rule's own step, 1 unfused pushes (within twice the
    rule's threshold), 2 pushes fused while their levels stay small, 3
    pulls writing lists."""
    k = G.mlp_param_count()
    index, nbrs = O.rmat_graph(12)
    ref = O.bfs_levels(index, nbrs, 0)
    with G.Graph.from_csr(index, nbrs) as g:
        lv, _ = g.bfslvl(0, trace=True)
        rule = [(t["direction"], t["out_kind"]) for t in g.trace()]
        for forced in range(4):
            p = np.zeros(k, np.float32)
            p[k - OUT + forced] = 10.0
            g.set_policy(p)
            lv, _ = g.bfslvl(0, direction=G.DIR_POLICY, trace=True)
            tr = g.trace()
            assert np.array_equal(lv, ref), forced
            if forced == 0:
                assert [(t["direction"], t["out_kind"]) for t in tr] == rule
            elif forced == 1:
                pushes_where_allowed(tr, rule, len(nbrs))
            elif forced == 2:
                assert all(t["direction"] == G.DIR_PUSH for t in tr) and tr[0]["fused"]
            else:
                # a list-writing pull where the rule pulls, the rule's step
                # where it pushes
                assert [t["direction"] for t in tr] == [d for d, _ in rule]
                assert all(t["out_kind"] in (G.FRONTIER_LIST, G.FRONTIER_EMPTY) for t in tr
                           if t["direction"] == G.DIR_PULL)
        g.set_policy(None)


def test_policy_gate_margin():
    """An action other than the rule's is taken only with a logit lead of at
    least the margin loaded after the parameters."""
    k = G.mlp_param_count()
    index, nbrs = O.rmat_graph(12)
    with G.Graph.from_csr(index, nbrs) as g:
        lv, _ = g.bfslvl(0, trace=True)
        rule = [(t["direction"], t["out_kind"], t["fused"]) for t in g.trace()]
        p = np.zeros(k, np.float32)
        p[k - OUT + 1] = 1.0            # an unfused push leads the rule's action by 1
        for margin, follows in ((0.5, True), (2.0, False)):
            g.set_policy(np.concatenate([p, [margin]]).astype(np.float32))
            lv2, _ = g.bfslvl(0, direction=G.DIR_POLICY, trace=True)
            tr = g.trace()
            assert np.array_equal(lv2, lv)
            if follows:
                pushes_where_allowed(tr, rule, len(nbrs))
            else:
                assert [(t["direction"], t["out_kind"], t["fused"]) for t in tr] == rule
        with pytest.raises(G.GprError):
            g.set_policy(np.concatenate([p, [-1.0]]).astype(np.float32))
        g.set_policy(None)


def test_trace_records_four_actions():
    """Each sample times runs with that step forced to each action, from the
    start of that step to the end of the run; action 0 is the rule's own
    step, so its times are what the rule takes from there on: the first
    sample's is a whole run of the rule, and later ones are shorter."""
    index, nbrs = O.rmat_graph(12)
    with G.Graph.from_csr(index, nbrs) as g:
        s = O.highest_degree_vertex(index)
        f, t = g.policy_trace(0, s)
        lv, info = g.bfslvl(s)
        assert f.shape[0] == t.shape[0] >= 3 and f.shape[1] == 6 and t.shape[1] == 4
        assert (t > 0).all() and np.isfinite(f).all()
        assert f.shape[0] <= info.iters + 1
        # sample 0 is step 0: a whole run of the rule, close to the
        # benchmark's own run time; the last sample starts at the last step
        assert t[0, 0] < 3 * info.ms and info.ms < 3 * t[0, 0]
        assert t[-1, 0] < t[0, 0]
        f2, t2 = g.policy_trace(1)
        assert len(f2) >= 2
        f3, t3 = g.policy_trace(0, s, cap=2)
        assert len(f3) <= 2


def test_trace_follows_network():
    """With follow=True the unforced steps are the loaded network's: the
    sampled states are those of the network's own run (one per step while
    the run has at most `cap` steps), not the rule's; without a network the
    call is refused."""
    from graptor_amd.policy import grid_csr
    k = G.mlp_param_count()
    index, nbrs = grid_csr(24, 24)   # BFS from a corner: 47 small levels
    with G.Graph.from_csr(index, nbrs) as g:
        s = 0
        want, _ = g.bfslvl(s, trace=True)
        rule_steps = len(g.trace())
        assert rule_steps == 47
        p = np.zeros(k, np.float32)
        p[k - OUT + 2] = 10.0   # always a push fused while small
        g.set_policy(p)
        lv, _ = g.bfslvl(s, direction=G.DIR_POLICY, trace=True)
        net_steps = len(g.trace())
        assert np.array_equal(lv, want)
        assert net_steps < rule_steps
        f, t = g.policy_trace(0, s, follow=True)
        assert len(f) == net_steps and (t > 0).all()
        fr, _ = g.policy_trace(0, s)
        assert len(fr) == rule_steps
        assert np.array_equal(f[0], fr[0])   # the same first state
        lv2, _ = g.bfslvl(s, direction=G.DIR_POLICY)   # the network stays loaded
        assert np.array_equal(lv2, want)
        # persist: the forced action holds to the end; the sampled states
        # are still those of the unforced run
        fp, tp = g.policy_trace(0, s, persist=True)
        assert np.array_equal(fp, fr) and (tp > 0).all()
        fpf, _ = g.policy_trace(0, s, follow=True, persist=True)
        assert np.array_equal(fpf, f)
        lv3, _ = g.bfslvl(s, direction=G.DIR_POLICY)
        assert np.array_equal(lv3, want)
        g.set_policy(None)
        with pytest.raises(G.GprError):
            g.policy_trace(0, s, follow=True)


def test_pull_list_falls_back_to_bitmap():
    """Action 3 (a pull writing a list) keeps the bitmap it also writes
    when the list would pass max(65536, n/32) entries: large outputs come
    back as bitmaps, small ones as lists, and the labels stay exact."""
    k = G.mlp_param_count()
    index, nbrs = O.rmat_graph(18)
    n = len(index) - 1
    limit = max(65536, n // 32)
    with G.Graph.from_csr(index, nbrs) as g:
        want, _ = g.ccv()
        p = np.zeros(k, np.float32)
        p[k - OUT + 3] = 10.0
        g.set_policy(p)
        lab, _ = g.ccv(direction=G.DIR_POLICY, trace=True)
        tr = g.trace()
        g.set_policy(None)
        assert np.array_equal(lab, want)
        pulls = [t for t in tr if t["direction"] == G.DIR_PULL and t["nactv_out"]]
        assert any(t["nactv_out"] > limit for t in pulls), tr
        for t in pulls:
            assert t["out_kind"] == (G.FRONTIER_BITMAP if t["nactv_out"] > limit else G.FRONTIER_LIST), t
