"""Builds libflowpro_hip.so in-tree with hipcc for gfx950 (no JIT, no torch extension)."""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libflowpro_hip.so")
# fp_quad_av.hip and fp_mixed_av.hip compile the damped halves of fp_quad.hip and fp_mixed.hip
SOURCES = ["fp_quad.hip", "fp_mixed.hip", "fp_quad_av.hip", "fp_mixed_av.hip", "fp_linalg.hip", "fp_element.hip", "fp_team.hip",
           "fp_team_kw.hip", "fp_tet.hip", "fp_host.cpp"]
# per-source optimisation level (default -O3): the k-omega SST team kernels are built at -O1, where they
# agree with the CPU restatement (at -O2 and -O3 their Jacobian rows did not; see fp_team_kw.hip)
OPT_LEVEL = {"fp_team_kw.hip": "-O1"}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FLOWPRO_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # packed FP32 instructions off for every device pass
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-result", "-Wno-unused-command-line-argument"]


def build(verbose: bool = False) -> str:
    """Compiles the sources concurrently (one hipcc per source), then links."""
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        flags = [OPT_LEVEL.get(src, f) if f == "-O3" else f for f in FLAGS]
        cmds.append([HIPCC, *flags, *lang, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        for cmd, res in zip(cmds, pool.map(lambda c: subprocess.run(c), cmds)):
            if res.returncode != 0:
                raise subprocess.CalledProcessError(res.returncode, cmd)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
