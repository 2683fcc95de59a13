"""Build the gfx950 shared library graptor_amd/libgpr.so with hipcc.

Every HIP translation unit is compiled for gfx950 only, in-tree, into
graptor_amd/_build/, then linked into graptor_amd/libgpr.so.  Packed FP32
instructions are disabled (-packed-fp32-ops) for all device code.
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpr.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
DRIVER_SRC = os.path.join(HERE, "driver", "gpr_driver.cpp")
DRIVER = os.path.join(HERE, "gpr_driver")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", "-Wno-unused-result",
    "-I" + os.path.join(ROOT, "include"),
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def headers():
    return (glob.glob(os.path.join(CSRC, "*.hpp")) + glob.glob(os.path.join(ROOT, "include", "*.h"))
            + glob.glob(os.path.join(ROOT, "include", "gpr", "*.hpp")))


def _obj(src):
    return os.path.join(OUT, os.path.basename(src) + ".o")


def _stale(src, newest_header):
    o = _obj(src)
    return not os.path.exists(o) or os.path.getmtime(o) < max(os.path.getmtime(src), newest_header)


def _compile(src):
    o = _obj(src)
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", o + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    os.replace(o + ".tmp", o)
    return src


def build(force=True, verbose=False, jobs=None):
    """Compile every source (force=True) or only stale ones, then link."""
    os.makedirs(OUT, exist_ok=True)
    srcs = sources()
    hdr = max(os.path.getmtime(h) for h in headers())
    todo = srcs if force else [s for s in srcs if _stale(s, hdr)]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        for s in ex.map(_compile, todo):
            if verbose:
                print("compiled", os.path.basename(s), flush=True)
    if todo or not os.path.exists(LIB):
        cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-o", LIB + ".tmp"] + [_obj(s) for s in srcs]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
        os.replace(LIB + ".tmp", LIB)
    if todo or not os.path.exists(DRIVER) or os.path.getmtime(DRIVER) < os.path.getmtime(DRIVER_SRC):
        build_driver()
    return LIB


def build_driver():
    """Link the command-line harness graptor_amd/gpr_driver against libgpr.so
    (host code only, so the system C++ compiler builds it)."""
    cmd = [CXX, "-O2", "-std=c++17", "-fopenmp", "-Wall", "-I" + os.path.join(ROOT, "include"), DRIVER_SRC,
           "-L" + HERE, "-lgpr", "-Wl,-rpath,$ORIGIN", "-o", DRIVER + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    build(force="--force" in sys.argv, verbose=True)
