#!/bin/bash
# Builds the Fortran binding (module + object) with amdflang and links the
# known-answer and device-vector Newton test programs against the in-tree
# libchidg_hip.so.
set -euo pipefail
HERE="$(cd "$(dirname "$0")" && pwd)"
ROOT="$(cd "$HERE/../.." && pwd)"
FC="${FC:-amdflang}"
if ! command -v "$FC" >/dev/null 2>&1; then
    FC=/opt/rocm/bin/amdflang
fi
OUT="$HERE/_build"
mkdir -p "$OUT"
"$FC" -O2 -c "$HERE/mod_chidg_hip.f90" -J "$OUT" -o "$OUT/mod_chidg_hip.o"
"$FC" -O2 -I "$OUT" "$ROOT/tests/fortran/known_answers.f90" "$OUT/mod_chidg_hip.o" \
    -L "$ROOT/chidg_amd" -lchidg_hip -Wl,-rpath,'$ORIGIN/../../chidg_amd' \
    -o "$ROOT/tests/fortran/known_answers"
"$FC" -O2 -I "$OUT" "$ROOT/tests/fortran/newton_device.f90" "$OUT/mod_chidg_hip.o" \
    -L "$ROOT/chidg_amd" -lchidg_hip -Wl,-rpath,'$ORIGIN/../../chidg_amd' \
    -o "$ROOT/tests/fortran/newton_device"
