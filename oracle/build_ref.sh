#!/bin/sh
# Builds oracle/_ref/libarecon_ref.so: the reference's own CPU functions,
# compiled from the source files where they lie (default /root/reference/src),
# for the oracle's parity tests and the timed CPU baseline.
#
#   src/Parallel_FDK_Helical_3DWeighting.cpp:15-16    TYPE and the file-level pi
#   src/Parallel_FDK_Helical_3DWeighting.cpp:42-113   ParallelFDKHelical3DWeightingCPU
#   src/ParallelRebinningCBCurve.cpp:21-97            ParallelRebinningCBCurve_CPU<T>
#
# These line ranges need only <cmath>, so they are fed to g++ on its standard
# input, preceded by that include and followed by ref_exports.inc (this
# project's extern "C" wrappers that instantiate them).  Nothing of the
# reference is written to disk; the only output is the shared library.  The rest
# of those translation units (the MEX gateways, which need MATLAB's <mex.h>, and
# the CUDA calls) is not compiled.
#
# No -march / -mfma: x86-64 SSE2 arithmetic, so no FMA contraction; the
# reference's float rebin then evaluates float expressions in float and its
# unqualified libm calls in double, as in its MATLAB-built MEX file.
set -eu
HERE=$(cd "$(dirname "$0")" && pwd)
REF=${AREF_REFERENCE_SRC:-/root/reference/src}
OUT="$HERE/_ref"
BP="$REF/Parallel_FDK_Helical_3DWeighting.cpp"
RB="$REF/ParallelRebinningCBCurve.cpp"
[ -r "$BP" ] && [ -r "$RB" ] || { echo "build_ref.sh: reference sources not found under $REF" >&2; exit 3; }
mkdir -p "$OUT"
{
  printf '#include <cmath>\n'
  printf '#line 15 "%s"\n' "$BP"; sed -n '15,16p' "$BP"
  printf '#line 42 "%s"\n' "$BP"; sed -n '42,113p' "$BP"
  printf '#line 21 "%s"\n' "$RB"; sed -n '21,97p' "$RB"
  printf '#line 1 "%s"\n' "$HERE/ref_exports.inc"; cat "$HERE/ref_exports.inc"
} | ${CXX:-g++} -std=c++11 -O2 -fPIC -shared -x c++ - -o "$OUT/libarecon_ref.so.tmp"
mv "$OUT/libarecon_ref.so.tmp" "$OUT/libarecon_ref.so"
