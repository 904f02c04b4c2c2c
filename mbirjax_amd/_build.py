"""In-tree build of libmbx.so for gfx950 (hipcc): one object per translation
unit, compiled in parallel, then linked."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

_HERE = os.path.dirname(os.path.abspath(__file__))
SOURCES = ["csrc/mbx_back_cone.hip", "csrc/mbx_back_parallel.hip", "csrc/mbx_fwd.hip", "csrc/mbx_project.hip",
           "csrc/mbx_back.hip", "csrc/mbx_vcd.hip", "csrc/mbx_denoise.hip", "csrc/mbx_noalloc.hip"]
OUTPUT = os.path.join(_HERE, "libmbx.so")
OBJDIR = os.path.join(_HERE, "_build")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay disabled in every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    return res.returncode, errs


def build(verbose=False, jobs=None):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, os.path.basename(s).replace(".hip", ".o")) for s in SOURCES]
    cmds = [[hipcc] + HIPCC_FLAGS + ["-c", os.path.join(_HERE, s), "-o", o] for s, o in zip(SOURCES, objs)]
    with ThreadPoolExecutor(max_workers=jobs or min(len(cmds), os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda c: _run(c, verbose), cmds))
    for (rc, errs), s in zip(results, SOURCES):
        if rc != 0:
            sys.stderr.write(errs + "\n")
            raise RuntimeError("hipcc failed compiling %s" % s)
    tmp = OUTPUT + ".tmp"
    rc, errs = _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    if rc != 0:
        sys.stderr.write(errs + "\n")
        raise RuntimeError("hipcc failed linking libmbx.so")
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
