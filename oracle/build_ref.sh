#!/bin/bash
# Compiles the reference's header-only path from its sources where they lie
# (/root/reference) into oracle/_ref/: one shared object per translation unit
# (the reference defines non-inline globals in its headers, and cc.cpp needs
# __DISABLE_SPMSPV__).  Flags are the reference's own x86 flags
# (CMakeLists.txt:22) plus -O2 and -fPIC.  Only run where /root/reference
# exists; the GPU box uses the prebuilt files.
set -euo pipefail
REF=${GBK_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
    echo "reference sources not found at $REF; skipping" >&2
    exit 0
fi
mkdir -p "$OUT"
FLAGS="-std=c++17 -O2 -fopenmp -fpermissive -fPIC -shared -w -D MAX_CORES=8 -I$REF"
g++ $FLAGS "$HERE/ref_bridge.cpp" -o "$OUT/libgbkref.so"
g++ $FLAGS -DREF_DISABLE_SPMSPV "$HERE/ref_bridge.cpp" -o "$OUT/libgbkref_cc.so"
echo "built $OUT/libgbkref.so $OUT/libgbkref_cc.so"
