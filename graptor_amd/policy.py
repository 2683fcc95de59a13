"""Training and evaluation of the learned direction/representation policy.

Pipeline (all on the GPU through the C ABI):
  1. training graphs: RMAT scale 18-22 with a in [0.45, 0.65] (b = c keep
     the 0.19:0.19:0.05 proportions of the remaining mass) generated on the
     device, and 2-D grids whose edges are rewired with probability p;
     every graph is relabelled by degree as the benchmarks are;
  2. traces: gpr_policy_trace times the four choices of every BFS and CC
     iteration from a snapshot (each plus the better next step);
  3. training: soft targets softmax(-(t/t_min - 1)/tau) over the four
     measured times, sample weight = time at stake (mean - min, ms), full-
     batch Adam on the GPU (gpr_mlp_train);
  4. evaluation on held-out graphs (RMAT-25 ef16, 4096x4096 grid with 1%
     random long edges) against the threshold rule, which stays the
     fallback when no network is loaded.

Run as ``python -m graptor_amd.policy train|eval``.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import DIR_AUTO, DIR_POLICY, Graph, mlp_param_count, mlp_train

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "policy_weights.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rmat_abc(a):
    rest = 1.0 - a
    return a, rest * 0.19 / 0.43, rest * 0.19 / 0.43


def grid_csr(w, h, rewire=0.0, long_frac=0.0, seed=0):
    """4-neighbour w x h grid; each grid edge has its far endpoint replaced
    by a uniform vertex with probability `rewire`; plus round(long_frac*n)
    uniform extra edges.  Symmetrised, deduplicated, no self loops."""
    rng = np.random.default_rng(seed)
    n = w * h
    v = np.arange(n, dtype=np.int64)
    right = v[(v % w) + 1 < w]
    down = v[v + w < n]
    src = np.concatenate([right, down])
    dst = np.concatenate([right + 1, down + w])
    if rewire > 0:
        m = rng.random(len(dst)) < rewire
        dst[m] = rng.integers(0, n, int(m.sum()))
    k = int(round(long_frac * n))
    if k:
        src = np.concatenate([src, rng.integers(0, n, k)])
        dst = np.concatenate([dst, rng.integers(0, n, k)])
    keep = src != dst
    src, dst = src[keep], dst[keep]
    keys = np.unique(np.concatenate([(src << 32) | dst, (dst << 32) | src]))
    u = (keys >> 32).astype(np.int64)
    index = np.zeros(n + 1, np.uint64)
    np.cumsum(np.bincount(u, minlength=n), out=index[1:])
    return index, (keys & 0xFFFFFFFF).astype(np.uint32)


def training_graphs(quick=False):
    """(name, constructor) pairs of the training set."""
    out = []
    scales = [18, 19, 20] if quick else [18, 19, 20, 21, 22]
    avals = [0.45, 0.57, 0.65] if quick else [0.45, 0.5, 0.55, 0.6, 0.65]
    for i, s in enumerate(scales):
        for j, a in enumerate(avals):
            if (i + j) % 2 and not quick:
                continue
            out.append(("rmat%d_a%.2f" % (s, a), lambda s=s, a=a, i=i, j=j: Graph.rmat(s, 16, *rmat_abc(a),
                                                                                       seed=1000 + 31 * i + j)))
    for side, p in ((512, 0.001), (512, 0.01), (1024, 0.001), (1024, 0.01), (1024, 0.05)):
        if quick and side == 1024 and p != 0.01:
            continue
        out.append(("grid%d_p%g" % (side, p), lambda side=side, p=p: Graph.from_csr(*grid_csr(side, side, rewire=p,
                                                                                                seed=side + int(p * 1e4)))))
    return out


def heuristic_choice(f, n, m):
    """The threshold rule's choice for a sample (api.h:150-153 and the
    engine's default output kinds), recovered from the features."""
    nactv = np.exp2(f[:, 0] * 32.0) - 1.0
    nacte = np.exp2(f[:, 1] * 40.0) - 1.0
    sparse = nactv + nacte <= m / 20.0
    return np.where(sparse, np.where(nacte > n, 1, 0), 3)


def collect(quick=False, sources=3):
    feats, times, tags, heur = [], [], [], []
    for name, make in training_graphs(quick):
        t0 = time.time()
        g = make()
        g.reorder()
        rng = np.random.default_rng(len(name))
        srcs = [g.highest_degree_vertex()] + list(rng.integers(0, g.n, sources - 1))
        for s in srcs:
            f, t = g.policy_trace(0, int(s))
            feats.append(f)
            times.append(t)
            heur.append(heuristic_choice(f, g.n, g.m))
            tags += ["%s/bfs" % name] * len(f)
        f, t = g.policy_trace(1)
        feats.append(f)
        times.append(t)
        heur.append(heuristic_choice(f, g.n, g.m))
        tags += ["%s/cc" % name] * len(f)
        log("traced %s: n=%d m=%d (%.1fs)" % (name, g.n, g.m, time.time() - t0))
        g.destroy()
    return np.concatenate(feats), np.concatenate(times), tags, np.concatenate(heur)


def targets_from_times(times, heur=None, tau=0.05, margin=0.03):
    """Soft targets over the four measured times; where the threshold rule's
    own choice is within `margin` of the fastest, the target is that choice
    (the network departs from the rule only for a measured gain)."""
    t = np.asarray(times, np.float64)
    tmin = t.min(1, keepdims=True)
    z = -(t / tmin - 1.0) / tau
    z -= z.max(1, keepdims=True)
    p = np.exp(z)
    p /= p.sum(1, keepdims=True)
    if heur is not None:
        th = t[np.arange(len(t)), heur]
        keep = th <= (1.0 + margin) * tmin[:, 0]
        p[keep] = 0.0
        p[keep, heur[keep]] = 1.0
    w = t.mean(1) - t.min(1) + 1e-3
    return p.astype(np.float32), w.astype(np.float32)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    k = mlp_param_count()
    p = np.zeros(k, np.float32)
    o = 0
    for fan_in, fan_out in ((6, 32), (32, 32), (32, 4)):
        p[o:o + fan_in * fan_out] = rng.standard_normal(fan_in * fan_out) * np.sqrt(2.0 / fan_in)
        o += fan_in * fan_out + fan_out
    return p


def train(quick=False, epochs=3000, lr=3e-3, out=WEIGHTS):
    f, t, tags, heur = collect(quick)
    T, W = targets_from_times(t, heur)
    p, loss = mlp_train(init_params(), f, T, W, epochs=epochs, lr=lr)
    pred = np.argmax(p_logits(p, f), 1)
    acc = float((pred == np.argmin(t, 1)).mean())
    # measured time of the chosen actions on the training traces
    rows = np.arange(len(t))
    cost = {"learned": float(t[rows, pred].sum()), "threshold": float(t[rows, heur].sum()),
            "oracle": float(t.min(1).sum())}
    meta = {"samples": int(len(f)), "loss": float(loss), "train_argmin_accuracy": acc, "trace_cost_ms": cost,
            "graphs": sorted(set(x.split("/")[0] for x in tags)), "epochs": epochs, "lr": lr}
    with open(out, "w") as fh:
        json.dump({"params": [float(x) for x in p], "meta": meta}, fh)
    log("trained on %d samples: loss %.4f, argmin accuracy %.3f" % (len(f), loss, acc))
    return p, meta


def p_logits(p, f):
    from . import mlp_forward
    return mlp_forward(p, f)


def load(path=WEIGHTS):
    with open(path) as fh:
        return np.array(json.load(fh)["params"], np.float32)


def held_out_graphs():
    return [("rmat25_ef16", lambda: Graph.rmat(25, 16, 0.57, 0.19, 0.19, seed=0x5EED25)),
            ("grid4096_long1pct", lambda: Graph.from_csr(*grid_csr(4096, 4096, long_frac=0.01, seed=4096)))]


def evaluate(params, reps=3, graphs=None):
    """Median run time of BFS (3 sources) and CC with the threshold rule and
    with the network; outputs are compared for equality."""
    rows = []
    for name, make in graphs or held_out_graphs():
        g = make()
        g.reorder()
        rng = np.random.default_rng(7)
        srcs = [g.highest_degree_vertex()] + [int(x) for x in rng.integers(0, g.n, 2)]
        for algo in ["bfs"] * len(srcs) + ["cc"]:
            s = srcs.pop(0) if algo == "bfs" else None
            res = {}
            for mode in ("threshold", "learned"):
                g.set_policy(params if mode == "learned" else None)
                d = DIR_POLICY if mode == "learned" else DIR_AUTO
                ts, out = [], None
                for _ in range(reps + 1):
                    o, info = g.bfslvl(s, direction=d) if algo == "bfs" else g.ccv(direction=d)
                    ts.append(info.ms)
                    out = o
                res[mode] = (float(np.median(ts[1:])), out)
            same = bool(np.array_equal(res["threshold"][1], res["learned"][1]))
            rows.append({"graph": name, "algo": algo, "source": s, "threshold_ms": res["threshold"][0],
                         "learned_ms": res["learned"][0], "speedup": res["threshold"][0] / res["learned"][0],
                         "bit_exact": same})
            log("%s %s: threshold %.3f ms, learned %.3f ms, x%.3f, equal=%s" % (
                name, algo, res["threshold"][0], res["learned"][0], rows[-1]["speedup"], same))
        g.set_policy(None)
        g.destroy()
    sp = np.array([r["speedup"] for r in rows])
    return {"rows": rows, "geomean_speedup": float(np.exp(np.log(sp).mean())), "worst_speedup": float(sp.min()),
            "all_bit_exact": all(r["bit_exact"] for r in rows)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "eval"])
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=3000)
    ap.add_argument("--out", default=WEIGHTS)
    args = ap.parse_args()
    if args.cmd == "train":
        train(args.quick, args.epochs, out=args.out)
    else:
        print(json.dumps(evaluate(load(args.out))))


if __name__ == "__main__":
    main()
