"""Build liba2c_fe.so (HIP, gfx950) in-tree with hipcc.

Every .hip/.cpp under a2c_amd/csrc is compiled to an object and linked into
a2c_amd/liba2c_fe.so.  Device code is built for gfx950 only, with packed FP32
instructions disabled and no fast-math.
"""
import os
import shlex
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "_build")
LIB = os.path.join(HERE, "liba2c_fe.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd[1:1] = ["-x", "hip"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(map(shlex.quote, cmd)), r.stdout, r.stderr))
    return obj


HOST = os.path.join(HERE, "host")
HOST_BINS = {"a2c_dmd": "a2c_dmd.cpp"}


def build_host(verbose=False):
    """C++ host programs over the library (g++, rpath to this directory)."""
    out = []
    for name, src in HOST_BINS.items():
        exe = os.path.join(HERE, name)
        cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + HOST,
               os.path.join(HOST, src), "-L" + HERE, "-la2c_fe", "-Wl,-rpath,$ORIGIN", "-o", exe + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("host build failed: %s\n%s" % (" ".join(cmd), r.stderr))
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    if verbose:
        print("built", " ".join(out))
    return out


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
