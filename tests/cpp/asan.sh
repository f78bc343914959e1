#!/usr/bin/env bash
# Host-code AddressSanitizer + UndefinedBehaviorSanitizer build of the engine
# library and the two C++ test binaries, into tests/cpp/_asan/.  Only the host
# side of the HIP sources is instrumented (-Xarch_host); the gfx950 code
# objects are the normal ones.
#
#   bash tests/cpp/asan.sh            # build (no GPU needed)
#   bash tests/cpp/asan.sh run        # build, then run both binaries on the GPU
#
# Leak reports for the ROCm runtime's own one-time allocations (allocated
# inside libhsa-runtime64 / libamdhip64) are suppressed by lsan.supp; any
# other leak, memory error or undefined behaviour fails the run.
set -euo pipefail
ROOT="$(cd "$(dirname "${BASH_SOURCE[0]}")/../.." && pwd)"
OUT="$ROOT/tests/cpp/_asan"
HIPCC="${HIPCC:-/opt/rocm/bin/hipcc}"
CLANGXX="${CLANGXX:-/opt/rocm/llvm/bin/clang++}"
mkdir -p "$OUT"
SAN=(-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-sanitize-recover=undefined
     -Xarch_host -fno-omit-frame-pointer)
HF=(--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -ffp-contract=off "-I$ROOT/include"
    -Xclang -target-feature -Xclang -packed-fp32-ops)
CS="$ROOT/optimisation_solver_amd/csrc"
"$HIPCC" "${HF[@]}" "${SAN[@]}" -c "$CS/engine.hip" -o "$OUT/engine.o" &
"$HIPCC" "${HF[@]}" "${SAN[@]}" -c "$CS/mlp.hip" -o "$OUT/mlp.o" &
"$HIPCC" -O1 -g -std=c++17 -fPIC "-I$ROOT/include" "${SAN[@]}" -x c++ -c "$CS/pdlp_api.cpp" -o "$OUT/pdlp_api.o" &
wait
"$HIPCC" --offload-arch=gfx950 -shared "${SAN[@]}" -o "$OUT/libpdlp_amd.so" "$OUT/engine.o" "$OUT/pdlp_api.o" "$OUT/mlp.o"
for t in test_dropin test_abi_paths; do
  "$CLANGXX" -O1 -g -std=c++17 "-I$ROOT/include" -fsanitize=address,undefined -fno-sanitize-recover=undefined \
    -fno-omit-frame-pointer \
    -o "$OUT/$t" "$ROOT/tests/cpp/$t.cpp" -L"$OUT" -lpdlp_amd -L"$ROOT/optimisation_solver_amd" -llpgen \
    -Wl,-rpath,"$OUT" -Wl,-rpath,"$ROOT/optimisation_solver_amd" -pthread
done
rm -f "$OUT"/*.o
if [ "${1:-}" = run ]; then
  cd "$ROOT"
  export ASAN_OPTIONS=detect_leaks=1:halt_on_error=1
  export LSAN_OPTIONS="suppressions=$ROOT/tests/cpp/lsan.supp"
  for t in test_dropin test_abi_paths; do
    echo "== $t (host code under AddressSanitizer and UndefinedBehaviorSanitizer)"
    timeout -k 10 300 "$OUT/$t"
  done
fi
