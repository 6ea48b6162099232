"""Builds libsablesparse_hip.so in place with hipcc for gfx950.

Every ``csrc/*.hip`` file is compiled to an object in ``csrc/_build/`` (in
parallel) and linked into ``sciwasm/sparse_amd/libsablesparse_hip.so``.
Packed FP32 instructions are disabled (``-packed-fp32-ops``) for all device
code.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsablesparse_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    "-fvisibility=hidden", "-Wall", "-Wno-unused-result",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-I" + os.path.join(HERE, "..", "..", "include"),
]


def _compile(src, obj):
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return obj


def build(verbose=False):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    bdir = os.path.join(CSRC, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = [os.path.join(bdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        list(ex.map(_compile, srcs, objs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT, file=sys.stderr)
    return OUT


NODE_DIR = os.path.join(HERE, "node")
ADDON = os.path.join(NODE_DIR, "sable_addon.node")
NODE_INCLUDE = "/usr/include/node"


def build_addon(verbose=False):
    """The Node-API addon (g++; host code only) linked against the HIP
    library with an $ORIGIN rpath.  Skipped, with a message, when the
    Node-API headers are absent."""
    if not os.path.exists(os.path.join(NODE_INCLUDE, "node_api.h")):
        print("node_api.h not found: Node-API addon not built", file=sys.stderr)
        return None
    tmp = ADDON + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-DNAPI_VERSION=8",
           "-I" + NODE_INCLUDE, os.path.join(NODE_DIR, "sable_addon.cc"), "-o", tmp,
           "-L" + HERE, "-l:libsablesparse_hip.so", "-Wl,-rpath,$ORIGIN/..", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("addon build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, ADDON)
    if verbose:
        print("built", ADDON, file=sys.stderr)
    return ADDON


if __name__ == "__main__":
    build(verbose=True)
    build_addon(verbose=True)
