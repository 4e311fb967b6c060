#!/bin/sh
# Builds oracle/_ref/libref.so from the reference's CPU engine sources where they lie (read only), plus the C shim
# of this folder. Test infrastructure only: runs in the development container, where the reference is mounted; the
# library is git-ignored and only the tests load it. No CUDA part is compiled (OPTLAB_HAVE_CUDA stays undefined).
set -e
REF=${OPTLAB_REFERENCE:-/root/reference}/packages/native/cuda
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -d "$REF/src" ]; then
    echo "reference sources not found at $REF: skipping oracle/_ref" >&2
    exit 0
fi
mkdir -p "$HERE/_ref"
${CXX:-g++} -O2 -std=c++17 -fPIC -pthread -ffp-contract=off -shared -I"$REF/include" \
    "$REF/src/paths_cpu.cpp" "$REF/src/leverage.cpp" "$REF/src/cpu_mc.cpp" "$REF/src/closed_forms.cpp" \
    "$REF/src/requests.cpp" "$HERE/ref_shim.cpp" -o "$HERE/_ref/libref.so.tmp"
mv "$HERE/_ref/libref.so.tmp" "$HERE/_ref/libref.so"
echo "built $HERE/_ref/libref.so"
