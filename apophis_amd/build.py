"""Build the native library in-tree: apophis_amd/_build/libapophis_hydro.so (gfx950).

Every translation unit is compiled with -ffp-contract=off (no fused multiply-add, so the
FP64 results keep the reference's rounding) and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
SOURCES = ["aph_fused.hip", "aph_fused_wide.hip", "aph_fused_amr.hip", "aph_adjoint.hip",
           "aph_adjoint_wide.hip", "aph_plain.hip", "aph_runtime.hip", "aph_amr.hip",
           "aph_amr_tree.hip", "aph_gravity.hip", "aph_poisson.hip"]
ARCH = os.environ.get("APH_OFFLOAD_ARCH", "gfx950")
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False, jobs=7):
    os.makedirs(OUT, exist_ok=True)
    objs, procs = [], []
    for s in SOURCES:
        o = os.path.join(OUT, s.replace(".hip", ".o"))
        objs.append(o)
        cmd = ["hipcc", *FLAGS, "-c", os.path.join(SRC, s), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT))
        if len(procs) >= jobs:
            _wait(procs)
            procs = []
    _wait(procs)
    lib = os.path.join(OUT, "libapophis_hydro.so")
    tmp = lib + ".tmp"
    subprocess.run(["hipcc", *FLAGS, "-shared", "-o", tmp, *objs], check=True)
    os.replace(tmp, lib)
    # C++ host program using the reference-shaped adapter (host code only)
    exe = os.path.join(OUT, "apophis_run")
    subprocess.run(["hipcc", "-O2", "-std=c++17", os.path.join(SRC, "host", "apophis_run.cpp"),
                    "-o", exe + ".tmp", "-L" + OUT, "-lapophis_hydro",
                    "-Wl,-rpath,$ORIGIN"], check=True)
    os.replace(exe + ".tmp", exe)
    # host build of the point/adjoint functions for the CPU adjoint checks (tests/native)
    tdir = os.path.join(os.path.dirname(HERE), "tests", "native")
    if os.path.isdir(tdir):
        os.makedirs(os.path.join(tdir, "_build"), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fPIC", "-shared",
                        os.path.join(tdir, "adj_check.cpp"), "-o",
                        os.path.join(tdir, "_build", "libadj_check.so")], check=True)
    return lib


def _wait(procs):
    for p in procs:
        out, _ = p.communicate()
        txt = "\n".join(l for l in out.decode().splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            sys.stderr.write(txt + "\n")
            raise RuntimeError("hipcc failed: " + " ".join(map(str, p.args)))


if __name__ == "__main__":
    print(build(verbose=True))
