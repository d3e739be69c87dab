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
SOURCES = ["kernels.hip", "ordered.hip", "staged.hip", "capi.hip", "batch.hip"]
HEADERS = ["traverse.h", "internal.h", "handle.h", "staged.h"]


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
    build_programs(verbose)
    return LIB


# HDF5 I/O and the reconstruction programs.  The I/O library is plain C so
# that it needs libhdf5 (serial 1.10, installed under /opt/conda) and no
# C++ runtime: loading it never pulls the older libstdc++ of that prefix
# into a process.  Its DT_RPATH names the HDF5 directory; the programs find
# both libraries next to them through $ORIGIN.
HDF5_PREFIX = os.environ.get("TRACE_HDF5_PREFIX", "/opt/conda")
H5LIB = os.path.join(HERE, "libtrace_h5io.so")
BIN = os.path.join(HERE, "bin")
PROGRAMS = {"sirt": 0, "mlem": 1, "pml": 2, "apmlr": 3, "trace": 4}
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_programs(verbose: bool = False) -> None:
    from concurrent.futures import ThreadPoolExecutor

    def run(cmd):
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    tmp = H5LIB + ".tmp"
    run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-I", INCLUDE,
         "-I", os.path.join(HDF5_PREFIX, "include"), os.path.join(CSRC, "h5io.c"),
         "-o", tmp, "-L", os.path.join(HDF5_PREFIX, "lib"), "-lhdf5",
         "-Wl,--disable-new-dtags,-rpath," + os.path.join(HDF5_PREFIX, "lib")])
    os.replace(tmp, H5LIB)
    os.makedirs(BIN, exist_ok=True)

    def program(name):
        out = os.path.join(BIN, name)
        run(["g++", "-O2", "-std=c++17", "-Wall", f"-DTRACE_CLI_APP={PROGRAMS[name]}",
             "-D__HIP_PLATFORM_AMD__", "-I", INCLUDE, "-I", os.path.join(ROCM, "include"),
             os.path.join(CSRC, "trace_cli.cpp"), "-o", out + ".tmp", "-L", HERE,
             "-ltrace_hip", "-ltrace_h5io", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
             "-Wl,-rpath,$ORIGIN/..," + "-rpath," + os.path.join(ROCM, "lib")])
        os.replace(out + ".tmp", out)

    with ThreadPoolExecutor(max_workers=len(PROGRAMS)) as ex:
        for j in [ex.submit(program, n) for n in PROGRAMS]:
            j.result()


if __name__ == "__main__":
    print(build(verbose=True))
