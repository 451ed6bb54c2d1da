"""Build libbsq.so (HIP kernels for gfx950 + host C++) in-tree with hipcc.

Usage: python -m breseq_amd.build   (or build() from __graft_entry__).
"""
import os
import re
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
    "gd_emit.hip",
    "sample.cpp",
    "host_util.cpp",
    "evidence.cpp",
    "covfit.cpp",
    "dp_regions.cpp",
    "simreads.cpp",
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


_LABEL = re.compile(r"^\.LBB\w+:|^; %bb\.\d+:")
_FUNC = re.compile(r"^(_Z\w+):")


def spill_before_exec_restore(asm_text):
    """Places in gfx950 assembly where a basic block stores a VGPR spill before the block's own
    exec-mask restore (the s_or_b64 exec, exec, ... that closes a divergent branch).  Such a store
    runs with the branch's narrowed mask, so the lanes outside it keep whatever the spill slot held
    before, and the reload after the restore hands those lanes another value.  Returns
    [(function, line of the block label, spill instruction)]."""
    hits = []
    func, block_line, pending, exec_written = None, 0, [], False
    for i, line in enumerate(asm_text.split("\n")):
        m = _FUNC.match(line)
        if m:
            func = m.group(1)
        if _LABEL.match(line):
            block_line, pending, exec_written = i + 1, [], False
            continue
        s = line.split(";")[0].strip() if not line.strip().startswith(";") else ""
        if not s:
            continue
        if "Folded Spill" in line and s.startswith(("scratch_store", "buffer_store")) and not exec_written:
            pending.append(s)
        if re.match(r"s_\w+\s+exec\b", s):
            if s.startswith("s_or_b64 exec, exec") and pending:
                hits.extend((func, block_line, sp) for sp in pending)
            exec_written, pending = True, []
        if s.startswith(("s_branch", "s_cbranch", "s_setpc", "s_endpgm")):
            pending = []
    return hits


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
        if src.endswith(".hip"):
            # the device assembly, checked below for spill placement
            asm = [cc] + COMMON + HIP_ONLY + ["-x", "hip", "--cuda-device-only", "-S", path, "-o",
                                              os.path.join(OBJ_DIR, src + ".s")]
            procs.append((src + " (assembly)", subprocess.Popen(asm, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
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
    bad = []
    for src in SOURCES:
        if src.endswith(".hip"):
            with open(os.path.join(OBJ_DIR, src + ".s")) as f:
                bad += [(src,) + h for h in spill_before_exec_restore(f.read())]
    if bad:
        for src, func, line, spill in bad:
            sys.stderr.write("%s: %s, block at assembly line %d: %s\n" % (src, func, line, spill))
        raise RuntimeError("the compiler placed VGPR spills before an exec-mask restore (lanes outside "
                           "the branch would reload stale values); restructure the named kernels")
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
