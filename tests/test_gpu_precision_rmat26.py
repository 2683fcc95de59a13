"""PageRank precision at full size: RMAT-26 (67M vertices, 2.1G edges),
relabelled by degree, every PR_vary contribution format against the FP64
oracle (OpenMP C restatement of PRv's power iteration).

Bars: relative L1 vs FP64 after 20 iterations -- FP32 1e-6, BF16 / FP16 /
customfp<6,10> 1e-3, truncating BF16 1e-2; FP32 run to an L1 residual of
1e-7 matches the oracle's converged ranks within 1e-6 (L1); the staged
PR_vary run (customfp<6,10> to 1e-3, then FP32 to 1e-7) also converges.
Run with -s to print the table (rel. L1, rel. L-inf, ms per iteration).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")

BARS = {"fp32": 1e-6, "bf16": 1e-3, "fp16": 1e-3, "cfp16": 1e-3, "bf16t": 1e-2}


def test_pr_formats_rmat26_vs_fp64():
    G.lib()   # fails loudly when the library or the gfx950 device is missing
    with G.Graph.rmat(26) as g:
        g.reorder()
        index, nbrs = g.download()   # the relabelled CSR; results come back in original IDs
        order = g.order()
        print("\nRMAT-26: n=%d m=%d" % (len(index) - 1, len(nbrs)), flush=True)
        x64, _, _ = O.pr_fp64(index, nbrs, 20, 0.0)
        den1, deninf = np.abs(x64).sum(), np.abs(x64).max()
        # the one-time edgemap layout is built here, outside the timings
        g.pr_vary([("fp32", 0.0)], max_iter=1, download=False)
        rows = {}
        for fmt, bar in BARS.items():
            x, info = g.pr_vary([(fmt, 0.0)], max_iter=20)
            d = x[order].astype(np.float64) - x64
            l1, linf = float(np.abs(d).sum() / den1), float(np.abs(d).max() / deninf)
            rows[fmt] = (l1, linf, info.ms / 20)
            print("  %-6s rel L1 %.2e  rel Linf %.2e  %.2f ms/iter" % (fmt, l1, linf, info.ms / 20), flush=True)
            assert l1 <= bar, (fmt, l1)
            assert abs(float(x.sum(dtype=np.float64)) - 1.0) < 1e-5
        # to convergence
        xc64, it64, _ = O.pr_fp64(index, nbrs, 1000, 1e-7)
        x, info = g.prv(max_iter=1000, tol=1e-7)
        x = x[order]
        print("  fp32 to 1e-7: %d iterations (oracle %d), L1 vs oracle %.2e" % (
            info.iters, it64, np.abs(x - xc64).sum()), flush=True)
        assert info.delta <= 1e-7 and np.abs(x - xc64).sum() < 1e-6
        xs, infos = g.pr_vary([("cfp16", 1e-3), ("fp32", 1e-7)], max_iter=1000)
        xs = xs[order]
        print("  cfp16 -> fp32 staged: %d iterations, residual %.2e, L1 vs oracle %.2e" % (
            infos.iters, infos.delta, np.abs(xs - xc64).sum()), flush=True)
        assert infos.delta <= 1e-7 and np.abs(xs - xc64).sum() < 2e-6
