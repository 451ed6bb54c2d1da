"""Build libbsq.so (HIP kernels for gfx950 + host C++) in-tree with hipcc.

Usage: python -m breseq_amd.build   (or build() from __graft_entry__).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbsq.so")
OBJ_DIR = os.path.join(HERE, "_build")

SOURCES = [
    "error_count.hip",
    "identify_columns.hip",
    "score_candidates.hip",
    "kernels.hip",
    "session.hip",
    "extra.hip",
    "selftest.hip",
    "coverage.hip",
    "sample.cpp",
    "host_util.cpp",
    "evidence.cpp",
    "covfit.cpp",
    "dp_regions.cpp",
]

# -ffp-contract=off: the FP64 scorer must round every multiply and add separately, as the
# reference's x86-64 build does, to reproduce its results bit for bit.
# -packed-fp32-ops off: packed FP32 forms have produced wrong results on this hardware.
COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
HIP_ONLY = ["--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    os.makedirs(OBJ_DIR, exist_ok=True)
    cc = hipcc()
    objs = []
    procs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(OBJ_DIR, src + ".o")
        objs.append(obj)
        cmd = [cc] + COMMON
        if src.endswith(".hip"):
            cmd += HIP_ONLY + ["-x", "hip"]
        cmd += ["-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((src, text))
        elif verbose and text.strip():
            print(text)
    if failed:
        for src, text in failed:
            sys.stderr.write("---- %s ----\n%s\n" % (src, text))
        raise RuntimeError("compilation failed: " + ", ".join(s for s, _ in failed))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
