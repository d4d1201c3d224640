"""Build libtopopt_hip.so (gfx950) in-tree with hipcc.

Usage: python topopt.jl_amd/build.py   (or __graft_entry__.build()).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtopopt_hip.so")
OBJDIR = os.path.join(HERE, "_build")

SOURCES = ["api.hip", "k_matvec.hip", "k_matvec_zm.hip", "k_pcg.hip", "k_energy.hip", "k_filter.hip", "k_oc.hip",
           "k_conv.hip", "nn_api.hip", "k_mg.hip", "k_beso.hip", "k_heat.hip", "heat_api.hip", "k_csr.hip", "k_mma.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=fast",
    # packed FP32 arithmetic stays off: every FP32 operation is a scalar VALU op
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
]


def sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = False, jobs: int = 8) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    procs = []
    objs = []
    for s in sources():
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJDIR, s.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs)
            procs = []
    _wait(procs)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout.decode())
    os.replace(tmp, LIB)
    return LIB


def _wait(procs):
    errs = []
    for name, p in procs:
        out, _ = p.communicate()
        text = out.decode()
        if p.returncode != 0:
            errs.append(f"--- {name} ---\n{text}")
        elif "warning" in text and os.environ.get("TOHIP_SHOW_WARNINGS"):
            print(text)
    if errs:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errs))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
