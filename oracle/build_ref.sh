#!/usr/bin/env bash
# Builds the checker libraries (test infrastructure, never linked into the product).
#   oracle/liboracle.so          the C++ restatement (strict IEEE, -ffp-contract=off)
#   oracle/_ref/libyafel_ref.so  the reference's own Eigen-free sources (strict flags, for goldens)
#   oracle/_ref/libyafel_ref_fast.so  same sources with the reference's Release flags
#                                (-O3 -ffast-math, OpenMP), the timed CPU baseline
# The reference libraries are only built where the reference tree is present.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
REF="${YAFEL_REFERENCE_DIR:-/root/reference}"
# x86-64-v3 rather than -march=native: the built files run on another host's CPU.
ARCH="-march=x86-64-v3"
g++ -std=c++17 -O2 -ffp-contract=off -fPIC -shared "$here/oracle.cpp" -o "$here/liboracle.so"
g++ -std=c++17 -O3 -ffast-math $ARCH -fopenmp -fPIC -shared "$here/oracle.cpp" -o "$here/liboracle_fast.so"
if [ -d "$REF/include/old_handmade_linalg" ]; then
  mkdir -p "$here/_ref"
  SRC="$here/ref_harness.cpp $REF/src/quadrature/QuadratureRule.cpp"
  g++ -std=c++17 -O2 -ffp-contract=off -fPIC -shared -I"$REF/include" $SRC \
      -o "$here/_ref/libyafel_ref.so"
  g++ -std=c++17 -O3 -ffast-math $ARCH -fopenmp -fPIC -shared -I"$REF/include" $SRC \
      -o "$here/_ref/libyafel_ref_fast.so"
fi
