#!/usr/bin/env bash
# Builds the reference CPU simplifier (ParaFROST src/cpu, read in place under
# /root/reference) together with oracle/ref_harness.cpp into oracle/_ref/pfref.
# Test infrastructure only: the binary validates the oracle restatement,
# writes the golden fixtures and times the reference CPU simplifier.
# The reference's own build system (install.sh) is not run.  Its one
# generated header, version.h, is the reference's template
# templates/version.in copied verbatim (install.sh:255-262 appends only
# optional #defines that src/cpu/version.cpp defaults itself).
set -euo pipefail
REF=${PF_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src/cpu" ]; then
  echo "build_ref: $REF/src/cpu not present; skipping the reference build" >&2
  exit 0
fi
mkdir -p "$OUT/gen" "$OUT/obj"
cp "$REF/templates/version.in" "$OUT/gen/version.h"
CXX=${CXX:-g++}
FLAGS="-std=c++20 -O3 -DNDEBUG -m64 -w"
objs=()
pids=()
for src in "$REF"/src/cpu/*.cpp; do
  base=$(basename "$src" .cpp)
  [ "$base" = "main" ] && continue
  obj="$OUT/obj/$base.o"
  objs+=("$obj")
  if [ ! -f "$obj" ] || [ "$src" -nt "$obj" ]; then
    $CXX $FLAGS -I"$OUT/gen" -I"$REF/src/cpu" -c "$src" -o "$obj" &
    pids+=($!)
  fi
  if [ ${#pids[@]} -ge 8 ]; then wait "${pids[0]}"; pids=("${pids[@]:1}"); fi
done
for p in "${pids[@]}"; do wait "$p"; done
$CXX $FLAGS -I"$OUT/gen" -I"$REF/src/cpu" "$HERE/ref_harness.cpp" "${objs[@]}" -o "$OUT/pfref"
echo "build_ref: built $OUT/pfref"
