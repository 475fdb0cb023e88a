"""Builds npme_amd/libnpme_amd.so (HIP kernels for gfx950 + C ABI) in-tree,
then the command-line tools in npme_amd/bin/.

    python -m npme_amd.build
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnpme_amd.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["geometry.hip", "fft.hip", "particles.hip", "setup.cpp", "host_math.cpp", "capi.cpp",
           "devmem.cpp", "server.cpp", "fileio.cpp", "noalloc.hip", "genfunc.cpp", "twobox.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NPME_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         # packed FP32 VALU forms are disabled for every HIP translation unit
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]


def _compile(src):
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for _, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_tools()
    return LIB


# command-line tools (host C++ over the C++ header and the C ABI), linked
# against the in-tree library with an rpath to it
APPS = os.path.join(HERE, "apps")
BIN = os.path.join(HERE, "bin")
TOOLS = ["npme", "laplaceDM", "RalphaDM", "helmholtzDM", "compareV", "makeRandomBox",
         "userDefineFunc"]


def _tool(name):
    out = os.path.join(BIN, name)
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-I", APPS,
           os.path.join(APPS, name + ".cpp"), "-o", out + ".tmp", "-L", HERE, "-lnpme_amd",
           "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("tool build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(out + ".tmp", out)
    return out


def build_tools():
    os.makedirs(BIN, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(TOOLS)) as ex:
        return list(ex.map(_tool, TOOLS))


if __name__ == "__main__":
    print(build(verbose=True))
