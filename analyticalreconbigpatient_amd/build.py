"""Build recipe for the gfx950 library (libarecon.so) and the CPU oracle.

``build_library()`` compiles every ``csrc/*.hip`` translation unit with hipcc for
gfx950 (one process per file, in parallel) and links ``libarecon.so`` next to
this file.  ``build_oracle()`` compiles ``oracle/arecon_oracle.c`` with gcc into
``oracle/liboracle.so`` (test infrastructure; the library never links it).

Run ``python -m analyticalreconbigpatient_amd.build`` to build both.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = ROOT / "_build"
LIB = PKG / "libarecon.so"
ORACLE_SRC = ROOT / "oracle" / "arecon_oracle.c"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

ARCH = "gfx950"
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # Packed FP32 instructions stay off: hipcc must not form v_pk_{add,mul,fma}_f32.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


# Per-file device-compiler options.  The backprojectors keep per-slice
# accumulators in a register array that uniform branches update one element at a
# time; promoted to a single vector value, every branch join would copy the whole
# array.  Left as an alloca, the array is split into independent scalars after
# unrolling.
FILE_FLAGS = {
    "backproject.hip": ["-mllvm", "-disable-promote-alloca-to-vector"],
}


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libarecon.so)")


def _run_parallel(cmds: list[list[str]]) -> None:
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the GPU-only feature flag: harmless
        text = "\n".join(l for l in text.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append((c, text))
        elif text.strip():
            sys.stderr.write(text + "\n")
    if failed:
        msg = "\n\n".join(" ".join(c) + "\n" + o for c, o in failed)
        raise RuntimeError(f"compilation failed:\n{msg}")


def build_library(verbose: bool = False) -> Path:
    """Compile csrc/*.hip for gfx950 and link libarecon.so (always rebuilds)."""
    hipcc = _hipcc()
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    objs = [BUILD / (s.stem + ".o") for s in sources]
    cmds = [[hipcc, *HIPCC_FLAGS, *FILE_FLAGS.get(s.name, []), "-c", str(s), "-o", str(o)]
            for s, o in zip(sources, objs)]
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _run_parallel(cmds)
    tmp = LIB.with_suffix(".so.tmp")
    link = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    _run_parallel([link])
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) into oracle/liboracle.so."""
    cc = os.environ.get("CC", "gcc")
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    cmd = [cc, "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
           "-std=c99", "-D_GNU_SOURCE", str(ORACLE_SRC), "-o", str(tmp), "-lm"]
    if verbose:
        print(" ".join(cmd))
    _run_parallel([cmd])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


MEX = PKG / "mex"
GATEWAYS = ("ParallelRebinningCBCurve", "reWeigAdFiltr_GPU", "Parallel_FDK_Helical_3DWeighting")


def build_mex(verbose: bool = False) -> list[Path]:
    """Build the mx API shim (libmxshim.so) and the three MEX gateways, each a
    shared library exporting mexFunction, linked against libarecon.so.  With a
    MATLAB installation, compile the same gateway sources with `mex` instead."""
    cxx = os.environ.get("CXX", "g++")
    shim = MEX / "libmxshim.so"
    cmds = [[cxx, "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{MEX / 'shim'}",
             str(MEX / "shim" / "mx_shim.cpp"), "-o", str(shim)]]
    if verbose:
        print(" ".join(cmds[0]))
    _run_parallel(cmds)
    outs, cmds = [], []
    for g in GATEWAYS:
        out = MEX / f"{g}.so"
        outs.append(out)
        cmds.append([cxx, "-O2", "-fPIC", "-shared", "-std=c++17", f"-I{MEX / 'shim'}",
                     f"-I{MEX}", f"-I{INCLUDE}", str(MEX / f"{g}.cpp"), "-o", str(out),
                     f"-L{PKG}", "-larecon", f"-L{MEX}", "-lmxshim",
                     "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        for c in cmds:
            print(" ".join(c))
    _run_parallel(cmds)
    return outs


REF_SRC = Path("/root/reference/src")
REF_RECIPE = ROOT / "oracle" / "build_ref.sh"


def build_ref(verbose: bool = False) -> Path | None:
    """Compile the reference's own CPU functions (test infrastructure; the
    oracle's bit-for-bit check and the timed CPU baseline) with
    oracle/build_ref.sh into oracle/_ref/, when the reference sources are
    present.  Elsewhere the library built beside them is used as it is."""
    if not REF_SRC.is_dir():
        return None
    if verbose:
        print(f"sh {REF_RECIPE}")
    _run_parallel([["sh", str(REF_RECIPE)]])
    return ROOT / "oracle" / "_ref" / "libarecon_ref.so"


def build_all(verbose: bool = False) -> None:
    build_oracle(verbose)
    build_ref(verbose)
    build_library(verbose)
    build_mex(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
