#!/bin/bash
# Builds libhammerhead_mi355x.so (gfx950) beside the package sources.
set -euo pipefail
HERE="$(cd "$(dirname "${BASH_SOURCE[0]}")" && pwd)"
OUT="$HERE/../libhammerhead_mi355x.so"
OBJ="$HERE/../_obj"
mkdir -p "$OBJ"
FLAGS=(-O3 -std=c++17 -fPIC --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -Wall -Wno-unused-function)
pids=()
for f in k_correlate k_corr_fast k_corr_mfma k_corr_mfma96 k_deform k_dewarp k_validate k_mask k_uq k_adjoint hh_api hh_engine; do
  src="$HERE/$f.hip"; [ -f "$src" ] || src="$HERE/$f.cpp"
  extra=()
  # matrix-core accumulators in VGPRs: no AGPR round trips, higher occupancy
  [ "$f" = k_corr_mfma ] || [ "$f" = k_corr_fast ] && extra=(-mllvm -amdgpu-mfma-vgpr-form=1)
  hipcc "${FLAGS[@]}" "${extra[@]}" -x hip -c "$src" -o "$OBJ/$f.o" 2> "$OBJ/$f.log" &
  pids+=($!)
done
rc=0
for p in "${pids[@]}"; do wait "$p" || rc=1; done
for f in "$OBJ"/*.log; do grep -v "unknown target feature\|packed-fp32-ops\|^1 warning generated\|^ *$" "$f" || true; done
[ $rc -eq 0 ] || { echo "compile failed" >&2; exit 1; }
hipcc --offload-arch=gfx950 -shared -fPIC -o "$OUT.tmp" "$OBJ"/*.o
mv "$OUT.tmp" "$OUT"
echo "built $OUT"
