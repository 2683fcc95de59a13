"""The command-line harness graptor_amd/gpr_driver (the bench/driver.C
shape) on the GPU: a b2 file written by the oracle is read back through the
driver's -b path, and the files it writes (-cc:outfile, -pr:outfile) and the
validation lines it prints are checked against the oracle.

Bars: CC labels bit-exact; BFS validation lines PASS with the oracle's
longest level and reached count; PR FP32 relative L1 <= 1e-6 vs the FP64
oracle (20 iterations); PR_vary (cfp16 to 1e-3, then fp32 to 1e-7) L1 < 2e-6 vs the
FP64 oracle run to 1e-7 (sum of ranks 1), with the reported delta <= 1e-7.
APR_vary ranks within 2e-4 (L1) of the oracle's restatement (2e-6 with FP32
contributions only); KC_bucket coreness bit-exact.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

G = pytest.importorskip("graptor_amd")
DRIVER = os.path.join(os.path.dirname(G.__file__), "gpr_driver")


@pytest.fixture(scope="module")
def graph(tmp_path_factory):
    G.lib()   # fails loudly when the library or the gfx950 device is missing
    assert os.path.exists(DRIVER), "gpr_driver is not built"
    index, nbrs = O.rmat_graph(12, 8)
    path = str(tmp_path_factory.mktemp("drv") / "rmat12.b2")
    O.b2_write(path, index, nbrs)
    return path, index, nbrs


def drive(*args):
    print("driver:", " ".join(args), flush=True)
    r = subprocess.run([DRIVER] + list(args), capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout, r.stderr


def running_times(out, rounds):
    t = [float(x) for x in re.findall(r"^Running : (\S+)$", out, re.M)]
    assert len(t) == rounds and all(x > 0 for x in t)
    avg = float(re.search(r"^Average : (\S+)$", out, re.M).group(1))
    assert avg == pytest.approx(sum(t) / rounds, rel=1e-4)


@pytest.mark.parametrize("reorder", [False, True])
def test_driver_bfslvl(graph, reorder):
    path, index, nbrs = graph
    start = O.highest_degree_vertex(index)
    out, err = drive("-app", "bfslvl", "-b", "-s", "-rounds", "2", "-start", str(start), "-itimes",
                     *(["-reorder"] if reorder else []), path)
    running_times(out, 2)
    lv = O.bfs_levels(index, nbrs, start)
    n = len(index) - 1
    reached = lv <= n
    m = re.search(r"Longest path from \d+ \(original: (\d+)\) : (\d+) at: \d+ \((\d+)\)", out)
    assert int(m.group(1)) == start
    assert int(m.group(2)) == int(lv[reached].max())
    assert lv[int(m.group(3))] == lv[reached].max()
    assert out.count("not activated in time: 0 PASS") == 2
    assert "Number of activated vertices: %d" % int(reached.sum()) in out
    # -itimes: one line per iteration of each round
    it = re.findall(r"^Iteration (\d+): \S+ (sparse|bit) density: ", err, re.M)
    assert len(it) >= 2 * int(lv[reached].max())


@pytest.mark.parametrize("reorder", [False, True])
def test_driver_cc_outfile(graph, tmp_path, reorder):
    path, index, nbrs = graph
    outfile = str(tmp_path / "cc.txt")
    out, err = drive("-app", "cc", "-b", "-s", "-rounds", "1", "-cc:outfile", outfile,
                     *(["-reorder"] if reorder else []), path)
    running_times(out, 1)
    got = np.loadtxt(outfile, dtype=np.int64)
    want = O.cc_labels(index, nbrs).astype(np.int64)
    assert np.array_equal(got, want)
    assert "Unique values: %d" % len(np.unique(want)) in out
    assert "All vertices accounted for? PASS" in out
    assert "Neighbour check: PASS" in err


def test_driver_pr_outfile(graph, tmp_path):
    path, index, nbrs = graph
    outfile = str(tmp_path / "pr.txt")
    out, _ = drive("-app", "pr", "-b", "-s", "-rounds", "2", "-pr:maxiter", "20", "-pr:outfile", outfile, "-reorder",
                   path)
    running_times(out, 2)
    got = np.loadtxt(outfile, dtype=np.float64)
    want, _, _ = O.pr_fp64(index, nbrs, max_iter=20)
    assert np.abs(got - want).sum() / np.abs(want).sum() <= 1e-6
    assert float(re.search(r"^PR\[0\]: (\S+)$", out, re.M).group(1)) == pytest.approx(want[0], rel=1e-5)


def test_driver_pr_vary_stages(graph, tmp_path):
    path, index, nbrs = graph
    outfile = str(tmp_path / "prv.txt")
    out, _ = drive("-app", "pr_vary", "-b", "-s", "-rounds", "1", "-pr:maxiter", "1000", "-pr:outfile", outfile,
                   path)
    running_times(out, 1)
    got = np.loadtxt(outfile, dtype=np.float64)
    want, _, _ = O.pr_fp64(index, nbrs, max_iter=1000, tol=1e-7)
    assert np.abs(got - want).sum() < 2e-6
    assert float(re.search(r"^delta: (\S+)", out, re.M).group(1)) <= 1e-7


def test_driver_rmat_generation():
    out, err = drive("-app", "cc", "-rmat", "12", "-ef", "8", "-rounds", "1")
    index, nbrs = O.rmat_graph(12, 8)
    assert "graph: n=%d m=%d symmetric" % (len(index) - 1, len(nbrs)) in err
    assert "Unique values: %d" % len(np.unique(O.cc_labels(index, nbrs))) in out


@pytest.mark.parametrize("app", ["bfslvl", "cc"])
@pytest.mark.parametrize("fusion", ["off", "1", "2"])
def test_driver_fusion(tmp_path, app, fusion):
    """-fusion selects the benchmarks' FUSION setting; on a grid (high
    diameter, sparse steps) the fused runs give the same answers."""
    index, nbrs = O.grid_graph(60, 50, 0.0, seed=2)
    path = str(tmp_path / "grid.b2")
    O.b2_write(path, index, nbrs)
    outfile = str(tmp_path / "cc.txt")
    extra = ["-cc:outfile", outfile] if app == "cc" else ["-start", "0"]
    out, _ = drive("-app", app, "-b", "-s", "-rounds", "2", "-fusion", fusion, *extra, path)
    running_times(out, 2)
    if app == "cc":
        assert np.array_equal(np.loadtxt(outfile, dtype=np.int64), O.cc_labels(index, nbrs).astype(np.int64))
        assert "Neighbour check: PASS" in _
    else:
        lv = O.bfs_levels(index, nbrs, 0)
        assert re.search(r"Longest path from 0 \(original: 0\) : %d " % int(lv.max()), out), out
        assert out.count("not activated in time: 0 PASS") == 2


def test_driver_fusion_rejects_unknown_setting(graph):
    path, _, _ = graph
    r = subprocess.run([DRIVER, "-app", "cc", "-b", "-s", "-fusion", "3", path], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0 and "unknown -fusion 3" in (r.stdout + r.stderr)


@pytest.mark.parametrize("reorder", [False, True])
def test_driver_apr_outfile(graph, tmp_path, reorder):
    """-app apr (APR_vary, variant 2046 by default): the ranks written to
    -pr:outfile equal the oracle's to the 16-bit-stage bar."""
    path, index, nbrs = graph
    outfile = str(tmp_path / "apr.txt")
    extra = ["-reorder"] if reorder else []
    out, _ = drive("-app", "apr", "-b", "-s", "-rounds", "2", "-pr:outfile", outfile, *extra, path)
    running_times(out, 2)
    got = np.loadtxt(outfile, dtype=np.float64)
    want, it, l1, ed, tr = O.apr_vary(index, nbrs)
    assert np.abs(got - want).sum() <= 2e-4
    assert float(re.search(r"^delta: (\S+)", out, re.M).group(1)) < 1e-7
    assert float(re.search(r"^PR\[0\]: (\S+)$", out, re.M).group(1)) == pytest.approx(want[0], rel=1e-3)


def test_driver_apr_stages_and_eps2(graph, tmp_path):
    path, index, nbrs = graph
    outfile = str(tmp_path / "apr32.txt")
    out, _ = drive("-app", "apr", "-b", "-s", "-rounds", "1", "-pr:stages", "f32:1e-7", "-apr:eps2", "0.05",
                   "-pr:outfile", outfile, path)
    got = np.loadtxt(outfile, dtype=np.float64)
    want, *_ = O.apr_vary(index, nbrs, stages=((0, 1e-7),), eps2=0.05)
    assert np.abs(got - want).sum() <= 2e-6


@pytest.mark.parametrize("reorder", [False, True])
def test_driver_kc_outfile(graph, tmp_path, reorder):
    """-app kc (KC_bucket): coreness written to -kc:outfile equals the
    oracle's bit for bit; Largest core and the validation line."""
    path, index, nbrs = graph
    outfile = str(tmp_path / "kc.txt")
    extra = ["-reorder"] if reorder else []
    out, err = drive("-app", "kc", "-b", "-s", "-rounds", "2", "-kc:buckets", "127", "-kc:outfile", outfile, *extra,
                     path)
    running_times(out, 2)
    want, kmax = O.kcore(index, nbrs)
    assert np.array_equal(np.loadtxt(outfile, dtype=np.int64), want.astype(np.int64))
    assert out.count("Largest core: %d" % kmax) == 2
    assert err.count("Validation successful") == 2
