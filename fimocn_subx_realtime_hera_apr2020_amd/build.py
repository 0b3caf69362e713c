"""Build the gfx950 shared library libfimhip.so in-tree with hipcc.

The FP64 kernels are compiled with fp contraction off so that their results
follow the reference's operation order; packed FP32 instructions are disabled
for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libfimhip.so")

SOURCES = ["fimhip_api.hip", "kernels_plain.hip", "transp3d.hip", "hybgen.hip", "ens_api.hip", "kernels_ens.hip",
           "transp3d_ens.hip", "fct_tiles.hip", "hycom_cnuity.hip",
           "hycom_edgvar.hip", "hycom_momtum.hip", "hycom_transp.hip", "hycom_hybgen.hip"]
HOST_SOURCES = ["grid.cpp", "lsq_svd.cpp"]      # host-only C++ (g++)
HEADERS = ["fimhip_internal.h", "fim_pow.h", "fim_pow_tables.h", "synth.h", "host_common.h", "ens_internal.h",
           "hycom_internal.h"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FIMHIP_ARCH", "gfx950")

COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE, "-I" + CSRC,
]


def _objs_dir():
    d = os.path.join(PKG, "_build")
    os.makedirs(d, exist_ok=True)
    return d


def build(force: bool = True, verbose: bool = False) -> str:
    """Compile every source and link libfimhip.so (always recompiles when force)."""
    od = _objs_dir()
    objs, jobs = [], []
    hdr_mtime = max(os.path.getmtime(os.path.join(CSRC, h)) for h in HEADERS)
    hdr_mtime = max(hdr_mtime, os.path.getmtime(os.path.join(INCLUDE, "fimhip.h")))
    for src in SOURCES + HOST_SOURCES:
        sp = os.path.join(CSRC, src)
        op = os.path.join(od, os.path.splitext(src)[0] + ".o")
        objs.append(op)
        if not force and os.path.exists(op) and os.path.getmtime(op) > max(os.path.getmtime(sp), hdr_mtime):
            continue
        if src.endswith(".hip"):
            cmd = [HIPCC, *COMMON, "-c", sp, "-o", op]
        else:
            cmd = ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-I" + INCLUDE,
                   "-c", sp, "-o", op]
        jobs.append(cmd)
    procs = []
    for cmd in jobs:                      # the sources compile independently
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        procs.append((cmd, subprocess.Popen(cmd)))
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


FC = os.environ.get("FC_AMD", "/opt/rocm/lib/llvm/bin/amdflang")
FORTRAN_BIND = os.path.join(PKG, "fortran", "fimhip_bind.F90")
CALLER_SRC = os.path.join(ROOT, "tests", "fortran", "fimhip_caller.F90")
CALLER = os.path.join(ROOT, "tests", "fortran", "fimhip_caller")
OCN_CALLER_SRC = os.path.join(ROOT, "tests", "fortran", "ocn_caller.F90")
OCN_CALLER = os.path.join(ROOT, "tests", "fortran", "ocn_caller")


def build_fortran(verbose: bool = False):
    """The Fortran binding module (fortran/fimhip_bind.F90) and the test
    drivers that call the library through it (tests/fortran/fimhip_caller for
    the atmosphere, tests/fortran/ocn_caller for the HYCOM ocean), with
    amdflang; None when no Fortran compiler is installed."""
    if not os.path.exists(FC):
        return None
    od = os.path.join(_objs_dir(), "fortran")
    os.makedirs(od, exist_ok=True)
    cmds = [[FC, "-c", "-O2", "-J" + od, FORTRAN_BIND, "-o", os.path.join(od, "fimhip_bind.o")],
            [FC, "-c", "-O2", "-I" + od, "-J" + od, CALLER_SRC, "-o", os.path.join(od, "fimhip_caller.o")],
            [FC, os.path.join(od, "fimhip_caller.o"), os.path.join(od, "fimhip_bind.o"), "-L" + PKG, "-lfimhip",
             "-Wl,-rpath,$ORIGIN/../../" + os.path.basename(PKG), "-o", CALLER + ".tmp"],
            [FC, "-c", "-O2", "-I" + od, "-J" + od, OCN_CALLER_SRC, "-o", os.path.join(od, "ocn_caller.o")],
            [FC, os.path.join(od, "ocn_caller.o"), os.path.join(od, "fimhip_bind.o"), "-L" + PKG, "-lfimhip",
             "-Wl,-rpath,$ORIGIN/../../" + os.path.basename(PKG), "-o", OCN_CALLER + ".tmp"]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), file=sys.stderr)
        subprocess.check_call(cmd, cwd=od)
    os.replace(CALLER + ".tmp", CALLER)
    os.replace(OCN_CALLER + ".tmp", OCN_CALLER)
    return CALLER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
