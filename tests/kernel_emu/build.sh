#!/bin/sh
# Builds the CPU-emulation driver: the library's own sources compiled as
# host C++ against tests/kernel_emu/hip/hip_runtime.h, with AddressSanitizer
# and UBSan.  Output: tests/kernel_emu/_build/emu_driver
# The host compiler is ROCm's clang (the kernels use __bf16 and
# ext_vector_type, which g++ lacks); EMU_CXX overrides it.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="${EMU_OUT:-$HERE/_build}"
CXX="${EMU_CXX:-/opt/rocm/llvm/bin/clang++}"
mkdir -p "$OUT"
rm -f "$OUT"/*.o
SAN=${EMU_SAN:--fsanitize=address,undefined}
for src in "$ROOT"/hermite-ugks_amd/csrc/*.hip; do
  obj="$OUT/$(basename "$src" .hip).o"
  $CXX -std=c++20 -O1 -g $SAN -fno-omit-frame-pointer -x c++ -I "$HERE" -I "$ROOT/include" \
      -I "$ROOT/hermite-ugks_amd/csrc" -c "$src" -o "$obj" &
done
$CXX -std=c++20 -O1 -g $SAN -I "$HERE" -I "$ROOT/include" -c "$HERE/emu_runtime.cpp" -o "$OUT/emu_runtime.o" &
$CXX -std=c++20 -O1 -g $SAN -I "$ROOT/include" -c "$HERE/emu_driver.cpp" -o "$OUT/emu_driver.o" &
wait
$CXX $SAN -o "$OUT/emu_driver" "$OUT"/*.o -lpthread
echo "$OUT/emu_driver"
# shared-library form (no sanitizer) for ctypes: HUGKS_LIB=<path> makes
# hugks.capi load it instead of the HIP build
if [ -n "$EMU_SO" ]; then
  for src in "$ROOT"/hermite-ugks_amd/csrc/*.hip "$HERE/emu_runtime.cpp"; do
    b=$(basename "$src"); b=${b%.*}
    $CXX -std=c++20 -O2 -fPIC -x c++ -I "$HERE" -I "$ROOT/include" -I "$ROOT/hermite-ugks_amd/csrc" \
        -c "$src" -o "$OUT/so_$b.o" &
  done
  wait
  $CXX -shared -o "$OUT/libhugks_emu.so" "$OUT"/so_*.o -lpthread
  echo "$OUT/libhugks_emu.so"
fi
