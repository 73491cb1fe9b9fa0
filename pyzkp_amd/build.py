"""Build the gfx950 HIP library ``libzkhip.so`` in-tree.

``python -m pyzkp_amd.build`` (or ``build()``) compiles every ``csrc/*.hip``
for gfx950 with hipcc, links them into ``pyzkp_amd/libzkhip.so``, and
builds ``libzkhost.so``, a CPU-only build of the same field / curve
arithmetic headers used by the host-logic tests.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libzkhip.so")
HOSTLIB = os.path.join(HERE, "libzkhost.so")

ARCH = os.environ.get("PYZKP_AMD_ARCH", "gfx950")
# -packed-fp32-ops: no v_pk_*_f32 generation (all arithmetic here is integer)
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str]) -> None:
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")


# Host-only translation units, built with the host compiler and their own
# code generation flags: the CPU feature check (none) and the host G2 MSM's
# AVX-512 IFMA accumulation, entered only after that check.
HOST_SRCS = {
    "host_cpu.cpp": [],
    "g2_simd.cpp": ["-mavx512f", "-mavx512ifma"],
}


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = []
    cmds = []
    for f in srcs:
        obj = os.path.join(BUILD, f[:-4] + ".o")
        objs.append(obj)
        cmds.append([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, f), "-o", obj])
    for f, extra in HOST_SRCS.items():
        obj = os.path.join(BUILD, f.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmds.append(["g++", "-O3", "-std=c++17", "-fPIC", *extra, "-c", os.path.join(CSRC, f), "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([hipcc, *HIP_FLAGS, "-shared", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    # CPU build of the shared arithmetic for host-logic tests
    tmp = HOSTLIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared",
          os.path.join(CSRC, "hostcheck.cpp"), "-o", tmp])
    os.replace(tmp, HOSTLIB)
    if verbose:
        print("built", LIB, HOSTLIB, file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
