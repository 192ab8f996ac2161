#!/usr/bin/env bash
# Builds oracle/_ref/libdss_ref.so from the reference's own headers, in place
# under the reference tree, plus oracle/ref_wrap.cpp.  Test infrastructure
# only; needs the reference checkout (default /root/reference) and g++.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
R="${DSS_REFERENCE:-/root/reference}"
T="$R/NeuKaDiS/KaDiS/external/RBC/external/tlx"
if [ ! -d "$R/src" ]; then
  echo "reference tree not found at $R; skipping oracle/_ref" >&2
  exit 0
fi
mkdir -p "$here/_ref"
g++ -std=c++17 -O2 -DNDEBUG -fPIC -shared -w \
    -I "$R/src" -I "$T" \
    "$here/ref_wrap.cpp" -o "$here/_ref/libdss_ref.so.tmp"
mv "$here/_ref/libdss_ref.so.tmp" "$here/_ref/libdss_ref.so"
echo "built $here/_ref/libdss_ref.so"
