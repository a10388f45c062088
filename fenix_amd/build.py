"""Builds libfenix.so (HIP, gfx950) in-tree: fenix_amd/libfenix.so.

Usage: python -m fenix_amd.build  (or fenix_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfenix.so")
SOURCES = ["fenix_kernels.hip", "fenix_api.cpp", "fenix_diff.cpp", "fenix_static.cpp"]

HIPCC_FLAGS = [
    "-O3",
    "-std=c++17",
    "--offload-arch=gfx950",
    "-fPIC",
    # Reference arithmetic order: no FMA contraction.
    "-ffp-contract=off",
    # Keep the compiler from emitting packed FP32 instructions.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    """Compiles every source to an object and links libfenix.so. Always rebuilds."""
    hipcc = _hipcc()
    objdir = os.path.join(HERE, "build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            cmd[1:1] = ["-x", "hip"]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        raise RuntimeError("\n".join(f"hipcc failed on {s}:\n{t}" for s, t in failed))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose=False):
    """Builds the CPU oracle (test infrastructure) with its Makefile."""
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    return os.path.join(ROOT, "oracle", "build", "liboracle.so")


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))


def build_node_addon(verbose=False):
    """Builds the Node-API addon of the JS host (fenix_amd/js/fenix_napi.node) with the
    system Node headers, linked against libfenix.so (rpath: the package directory).
    Returns the path, or None when no Node headers are installed."""
    inc = next((d for d in ("/usr/include/node", "/usr/local/include/node") if os.path.exists(os.path.join(d, "node_api.h"))), None)
    if inc is None:
        return None
    out = os.path.join(HERE, "js", "fenix_napi.node")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-Wall", f"-I{inc}", "-DNODE_GYP_MODULE_NAME=fenix_napi",
           os.path.join(HERE, "js", "fenix_napi.cpp"), "-o", out + ".tmp", f"-L{HERE}", "-lfenix",
           "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out
