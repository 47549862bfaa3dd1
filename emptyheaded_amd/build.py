"""Builds libehx.so (HIP for gfx950) in-tree with hipcc.

    python -m emptyheaded_amd.build          # or: python emptyheaded_amd/build.py

Objects go to emptyheaded_amd/_build/, the library to emptyheaded_amd/libehx.so.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libehx.so")
ARCH = os.environ.get("EHX_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    "ehx_api.cpp",
    "ehx_graph.cpp",
    "ehx_refdb.cpp",
    "ehx_prims.hip",
    "ehx_trie.hip",
    "ehx_join.hip",
    "ehx_tuned.hip",
    "ehx_selector.hip",
    "ehx_mlp.hip",
    "ehx_cycles.hip",
    "ehx_ghd.hip",
    "ehx_closure.hip",
    "ehx_rank.hip",
]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-result"]
# Packed FP32 VALU forms are disabled for every device pass.
DEVICE = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OUT_DIR, src + ".o")
    cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(_compile, srcs))
    if verbose:
        for _, err in results:
            if err.strip():
                sys.stderr.write(err)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_dropin()
    return LIB


ENGINE = os.path.join(HERE, "engine")
DROPIN_LIB = os.path.join(ENGINE, "libehx_dropin.so")


def build_dropin():
    """The C++ drop-in API runtime plus the restated generated queries, linked
    against libehx.so (found next to the package through the rpath)."""
    srcs = [os.path.join(ENGINE, "src", "dropin.cpp")] + sorted(
        os.path.join(ENGINE, "queries", f) for f in os.listdir(os.path.join(ENGINE, "queries")) if f.endswith(".cpp"))
    cxx = shutil.which("g++") or "g++"
    tmp = DROPIN_LIB + ".tmp"
    # no fp contraction: the queries' float arithmetic is exactly what their source says
    cmd = [cxx, "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wno-unused-variable",
           "-I", os.path.join(ENGINE, "include"), "-o", tmp] + srcs + [
        "-L", HERE, "-lehx", "-Wl,-rpath,$ORIGIN/..", "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, DROPIN_LIB)
    return DROPIN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
