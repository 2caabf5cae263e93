"""Compile the gfx950 HIP library and its Cython binding in-tree.

    python -m qlattice_amd.build        # -> qlattice_amd/libqlat_dwf_hip.so
                                        #    qlattice_amd/_cdwf.<abi>.so

The library is built with hipcc directly (no JIT cache), so the shared
object sits next to this file and is what every loader in the package uses.
The Cython module (_cdwf.pyx, declared against include/qlat_dwf.h) is
translated with cython and compiled with the host C compiler; it links the
HIP library with an $ORIGIN run path.
"""

import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqlat_dwf_hip.so")
SOURCES = ["dwf_kernels.hip", "dwf_three_phase.hip", "dwf_api.hip"]
HEADERS = ["dwf_common.h", "dwf_internal.h", os.path.join("..", "..", "include", "qlat_dwf.h")]

INCLUDE = os.path.join(HERE, "..", "include")
PYX = os.path.join(HERE, "_cdwf.pyx")
EXT = os.path.join(HERE, "_cdwf" + sysconfig.get_config_var("EXT_SUFFIX"))

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are kept out of every kernel of this library
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-result",
]


def build(verbose=True, jobs=3):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd), src))
        objs.append(obj)
        if len(procs) >= jobs:
            p, s = procs.pop(0)
            if p.wait() != 0:
                raise RuntimeError(f"hipcc failed on {s}")
    for p, s in procs:
        if p.wait() != 0:
            raise RuntimeError(f"hipcc failed on {s}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    build_binding(verbose)
    return LIB


def build_binding(verbose=True):
    """cython _cdwf.pyx -> C, then a Python extension linked to the HIP library."""
    csrc = os.path.join(HERE, "_cdwf.c")
    cmd = [sys.executable, "-m", "cython", "-3", "-I", INCLUDE, PYX, "-o", csrc]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    cc = os.environ.get("CC", "gcc")
    tmp = EXT + ".tmp"
    cmd = [cc, "-O2", "-shared", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unreachable-code",
           f"-I{sysconfig.get_paths()['include']}", f"-I{INCLUDE}", csrc, "-o", tmp,
           f"-L{HERE}", "-l:libqlat_dwf_hip.so", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, EXT)
    os.remove(csrc)
    return EXT


if __name__ == "__main__":
    if "--binding" in sys.argv:
        build_binding(verbose="-q" not in sys.argv)
    else:
        build(verbose="-q" not in sys.argv)
