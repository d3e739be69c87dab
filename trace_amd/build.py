"""Builds the gfx950 shared library ``trace_amd/libtrace_hip.so`` in-tree.

hipcc flags:
* ``--offload-arch=gfx950``                 CDNA4 only;
* ``-ffp-contract=off``                     the traversal and the update rules
  reproduce the reference's float expressions, which were built without FMA
  contraction (SIRT Makefile:10-11);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``  no packed-fp32 VALU
  instructions are generated.
"""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libtrace_hip.so")
SOURCES = ["kernels.hip", "capi.hip", "batch.hip"]
HEADERS = ["traverse.h", "internal.h", "handle.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
    ]


FORBIDDEN_OPS = ("v_pk_add_f32", "v_pk_mul_f32", "v_pk_fma_f32")


def _check_isa(cc: str, src: str, odir: str, verbose: bool) -> None:
    """Writes the gfx950 assembly of `src` to _build/ and refuses packed-fp32
    VALU instructions in it."""
    asm = os.path.join(odir, src + ".gfx950.s")
    cmd = [cc, *flags(), "--cuda-device-only", "-S", os.path.join(CSRC, src), "-o", asm]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, stderr=subprocess.DEVNULL)
    text = open(asm).read()
    bad = [op for op in FORBIDDEN_OPS if op in text]
    if bad:
        raise RuntimeError(f"{src}: packed-fp32 instructions in device code: {bad}")


def build(verbose: bool = False) -> str:
    """Compiles every HIP source into LIB (always recompiles); the object
    and assembly passes of all sources run in parallel."""
    from concurrent.futures import ThreadPoolExecutor
    cc = hipcc()
    odir = os.path.join(HERE, "_build")
    os.makedirs(odir, exist_ok=True)
    objs = [os.path.join(odir, src + ".o") for src in SOURCES]

    def compile_one(i):
        cmd = [cc, *flags(), "-c", os.path.join(CSRC, SOURCES[i]), "-o", objs[i]]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    with ThreadPoolExecutor(max_workers=2 * len(SOURCES)) as ex:
        jobs = [ex.submit(compile_one, i) for i in range(len(SOURCES))]
        jobs += [ex.submit(_check_isa, cc, src, odir, verbose) for src in SOURCES]
        for j in jobs:
            j.result()  # re-raises a failed compile or a refused instruction
    tmp = LIB + ".tmp"
    cmd = [cc, "-shared", "-fPIC", "--offload-arch=gfx950", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
