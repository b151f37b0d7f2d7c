#!/usr/bin/env bash
# Compiles the reference's serial Sparse BLAS and conversion code from its own
# sources under /root/reference (read in place, never copied) together with
# oracle/ref_driver.cc into oracle/_ref/libref_driver.so.  Used only to pin the
# oracle and as the benchmark's CPU baseline.  GCC with -O2 -ffp-contract=off
# (no FMA contraction), as the reference's serial code is built on x86-64.
set -euo pipefail
REF=${ASC_REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/include/asc" ]; then
  echo "reference sources not found at $REF; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$OUT"
CXXFLAGS="-std=c++20 -O2 -ffp-contract=off -fPIC -I$REF/include -DASC_CORE_BUILDING -DASC_SPARSE_BUILDING"
SRCS="$REF/src/core/status.cc $REF/src/core/memory.cc $REF/src/core/execution.cc $REF/src/core/contracts.cc $REF/src/sparse/reference_blas.cc $REF/src/sparse/standard_blas.cc"
g++ $CXXFLAGS -shared -o "$OUT/libref_driver.so.tmp" $SRCS "$HERE/ref_driver.cc"
mv "$OUT/libref_driver.so.tmp" "$OUT/libref_driver.so"
echo "built $OUT/libref_driver.so"
