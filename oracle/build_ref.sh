#!/bin/bash
# Builds the part of the GeoFLOW reference that compiles from its own source
# files: the hand-unrolled FP64 contraction kernels dmxm1..dmxm24 of
# src/cdg/blas/dmxmp.F (what dmxm dispatches to, cff_blas.F:846-943), with
# amdflang.  Output goes only to oracle/_ref/ (git-ignored).  Nothing else of
# the reference builds here: cff_blas.F is rejected by flang's semantics
# checks (implicit-typing order, cff_blas.F:642-646) and the C++ sources need
# the CMake-generated configure.hpp plus Boost and MPI.
set -euo pipefail
REF=${1:-/root/reference}
HERE="$(cd "$(dirname "$0")" && pwd)"
OUT="$HERE/_ref"
FC=${FC:-/opt/rocm/llvm/bin/amdflang}
if [ ! -f "$REF/src/cdg/blas/dmxmp.F" ]; then
  echo "build_ref: reference sources not present; skipping" >&2
  exit 0
fi
mkdir -p "$OUT"
"$FC" -O2 -fPIC -c "$REF/src/cdg/blas/dmxmp.F" -o "$OUT/dmxmp.o"
"$FC" -shared -fPIC "$OUT/dmxmp.o" -o "$OUT/libref_dmxm.so"
echo "$OUT/libref_dmxm.so"
