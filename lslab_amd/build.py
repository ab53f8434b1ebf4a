"""Builds liblslab.so (HIP kernels + C ABI) in-tree for gfx950 with hipcc.

The kernels of the six table layouts are separate translation units
(csrc/launch.hip compiled once per layout), compiled in parallel with the C
ABI (csrc/lslab.hip), the workload generators, the host-side KVCG split
(csrc/kvcg.cpp) and the wide-key tables (csrc/wide.hip), then linked into
one shared library.

python -m lslab_amd.build [--verbose]
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LAYOUTS = ["L_M3232", "L_M3264", "L_M6432", "L_M6464", "L_S32", "L_S64"]
UNITS = ([("lslab", "csrc/lslab.hip", []), ("workload", "csrc/workload.hip", []),
          ("kvcg", "csrc/kvcg.cpp", []), ("wide", "csrc/wide.hip", [])] +
         [(f"launch_{l.lower()}", "csrc/launch.hip", [f"-DLSLAB_LAYOUT={l}"]) for l in LAYOUTS])
HEADERS = ["csrc/slab.hpp", "csrc/engine.hpp", "csrc/runs.hpp", "csrc/kernels.hpp"]
OUT = os.path.join(HERE, "liblslab.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 instructions are kept out of every code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{' '.join(cmd)}\n{msgs}")
    if verbose and msgs.strip():
        print(msgs, flush=True)


def build(verbose=False, jobs=None):
    jobs = jobs or min(len(UNITS), os.cpu_count() or 1, 16)
    with tempfile.TemporaryDirectory(prefix="lslab_build_") as tmpdir:
        objs = []
        cmds = []
        for name, src, defs in UNITS:
            obj = os.path.join(tmpdir, name + ".o")
            objs.append(obj)
            cmds.append([HIPCC] + FLAGS + defs + ["-c", "-o", obj, os.path.join(HERE, src)])
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for f in [ex.submit(_run, c, verbose) for c in cmds]:
                f.result()
        tmp = OUT + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-pthread", "-o", tmp] + objs,
             verbose)
        os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
