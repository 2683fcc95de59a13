"""The committed learned policy on the held-out graphs, against the rule.

Graphs (never seen in training or gate calibration): RMAT scale 25, edge
factor 16, (a, b, c) = (0.57, 0.19, 0.19); a 4096 x 4096 grid with 1% random
long edges.  On each: BFSLVLv from three sources (the highest-degree vertex
and two seeded ones) and CCv, each timed as the median of nine runs after a
warm one, under the threshold rule (default_threshold::is_sparse,
api.h:150-153, and CCv's frac_threshold) and under the network
(graptor_amd/policy_weights.json, its gate margin included), the rule's
runs and the network's alternating.

Checked: every output of the network's runs equals the rule's; one BFS per
graph and the grid's CC also equal the oracle; the network is never more
than 3% slower than the rule on any run, and at least 10% faster on the
geometric mean (the committed network measures about 1.13x).
The measured figures are printed, and written as policy_heldout.json into
the folder that GPR_RESULTS_DIR names, when it is set.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")
P = pytest.importorskip("graptor_amd.policy")

def test_committed_weights_load():
    w = P.load()
    assert w.shape == (G.mlp_param_count() + 1,)
    assert np.all(np.isfinite(w)) and w[-1] >= 0


def test_heldout_speedup_and_bit_exact():
    G.lib()
    W = P.load()
    rows = []
    for name, make in P.held_out_graphs():
        g = make()
        try:
            index, nbrs = g.download()
            g.reorder()
            srcs = P.sources(g)
            for k, s in enumerate(srcs + [None]):
                algo = "bfs" if s is not None else "cc"
                tr, tl, same = P.run_pair(g, algo, s, W, reps=9)
                rows.append((name, algo, s, tr, tl))
                g.set_policy(W)
                _, _ = g.bfslvl(s, direction=G.DIR_POLICY, trace=True) if algo == "bfs" else \
                    g.ccv(direction=G.DIR_POLICY, trace=True)
                steps = " ".join("%s%s:%d" % ("P" if t["direction"] == G.DIR_PUSH else "L", "f" if t["fused"] else "",
                                              t["nactv_in"]) for t in g.trace())
                print("%s %s %s: rule %.3f ms, learned %.3f ms, x%.3f | %s" % (name, algo, s, tr, tl, tr / tl, steps),
                      flush=True)
                assert same, (name, algo, s)
                # the network's result against the oracle
                if algo == "bfs" and k == 0:
                    lv, _ = g.bfslvl(s, direction=G.DIR_POLICY)
                    want, _ = O.graptor_bfs(index, nbrs, s)
                    assert np.array_equal(lv, want), (name, s)
                if algo == "cc" and name.startswith("grid"):
                    lab, _ = g.ccv(direction=G.DIR_POLICY)
                    assert np.array_equal(lab, O.cc_labels(index, nbrs)), name
                g.set_policy(None)
            del index, nbrs
        finally:
            g.destroy()
    sp = np.array([r[3] / r[4] for r in rows])
    geo, worst = float(np.exp(np.log(sp).mean())), float(sp.min())
    print("held-out geomean x%.3f, worst x%.3f" % (geo, worst), flush=True)
    out = os.environ.get("GPR_RESULTS_DIR")
    if out and os.path.isdir(out):
        with open(os.path.join(out, "policy_heldout.json"), "w") as fh:
            json.dump({"rows": rows, "geomean": geo, "worst": worst}, fh)
    assert worst >= 0.97
    assert geo >= 1.10
