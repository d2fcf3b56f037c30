"""In-tree build of the gfx950 engine (libhitleiden.so) and the workload
generators (libhl_gen.so).  `python -m hit_leiden_amd.build` or build_all()."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ENGINE_SO = os.path.join(HERE, "libhitleiden.so")
GEN_SO = os.path.join(HERE, "libhl_gen.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -ffp-contract=off: every a*b+c rounds twice, as the reference (Rust) does.
# -Xclang -target-feature -Xclang -packed-fp32-ops: no packed fp32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-command-line-argument",
]
ENGINE_SOURCES = ["hl_kernels.hip", "hl_move.hip", "hl_engine.hip"]


def _run_parallel(cmds):
    procs = [subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT) for c in cmds]
    failed = []
    for c, p in zip(cmds, procs):
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((c, out.decode(errors="replace")))
    if failed:
        msg = "\n".join(" ".join(c) + "\n" + o[-4000:] for c, o in failed)
        raise RuntimeError("build failed:\n" + msg)


def build_engine(verbose=False):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in ENGINE_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    _run_parallel(cmds)
    tmp = ENGINE_SO + ".tmp"
    _run_parallel([[HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]])
    os.replace(tmp, ENGINE_SO)
    if verbose:
        print("built", ENGINE_SO)


def build_gen(verbose=False):
    tmp = GEN_SO + ".tmp"
    _run_parallel([["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-I", CSRC,
                    os.path.join(CSRC, "hl_gen.cpp"), "-o", tmp]])
    os.replace(tmp, GEN_SO)
    if verbose:
        print("built", GEN_SO)


PROBE_EXE = os.path.join(HERE, "hl_gather_probe")


def build_probe(verbose=False):
    """The random-gather ceiling probe (bench.py --gather-probe)."""
    tmp = PROBE_EXE + ".tmp"
    _run_parallel([[HIPCC, *HIP_FLAGS, os.path.join(CSRC, "hl_gather_probe.hip"), "-o", tmp]])
    os.replace(tmp, PROBE_EXE)
    if verbose:
        print("built", PROBE_EXE)


def build_all(verbose=False):
    build_gen(verbose)
    build_engine(verbose)
    build_probe(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
