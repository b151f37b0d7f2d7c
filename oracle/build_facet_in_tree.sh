#!/usr/bin/env bash
# Builds the C++ facet's test program inside the reference's world: the
# facet headers compiled against the reference's own public headers
# (ASC_SPARSE_HIP_IN_ASC_TREE), its compiled part (asc-cpp_amd/csrc/facet)
# against the reference's src/core/execution_internal.h, linked with the
# reference's core library compiled from its own sources where they lie under
# /root/reference (read in place, never copied) and with this repository's
# libasc_sparse_hip.so.  Output: oracle/_ref/facet_in_tree (test
# infrastructure: it shows the facet drops into the reference's core types).
set -euo pipefail
REF=${ASC_REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/.." && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/include/asc" ]; then
  echo "reference sources not found at $REF; skipping oracle/_ref/facet_in_tree" >&2
  exit 0
fi
LIB="$ROOT/asc-cpp_amd/lib"
if [ ! -f "$LIB/libasc_sparse_hip.so" ]; then
  echo "build asc-cpp_amd first (libasc_sparse_hip.so is missing)" >&2
  exit 1
fi
mkdir -p "$OUT/facet_obj"
FLAGS="-std=c++20 -O2 -fPIC -DASC_SPARSE_HIP_IN_ASC_TREE -DASC_CORE_BUILDING -I$REF/include -I$REF/src/core -I$ROOT/include"
for f in status memory execution contracts; do
  g++ $FLAGS -c "$REF/src/core/$f.cc" -o "$OUT/facet_obj/core_$f.o"
done
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
for f in core_hip sparse_context; do
  "$HIPCC" $FLAGS -c "$ROOT/asc-cpp_amd/csrc/facet/$f.cc" -o "$OUT/facet_obj/$f.o"
done
"$HIPCC" $FLAGS -Wno-missing-field-initializers -c "$ROOT/tests/cpp/facet_test.cc" -o "$OUT/facet_obj/facet_test.o"
"$HIPCC" -o "$OUT/facet_in_tree.tmp" "$OUT"/facet_obj/*.o -L"$LIB" -lasc_sparse_hip -Wl,-rpath,"$LIB" -Wl,-rpath,'$ORIGIN/../../asc-cpp_amd/lib'
mv "$OUT/facet_in_tree.tmp" "$OUT/facet_in_tree"
echo "built $OUT/facet_in_tree"
