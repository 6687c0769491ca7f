#!/bin/sh
# Builds the CPU-emulation driver: the library's own sources compiled with
# g++ against tests/kernel_emu/hip/hip_runtime.h, with AddressSanitizer and
# UBSan.  Output: tests/kernel_emu/_build/emu_driver
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="${EMU_OUT:-$HERE/_build}"
mkdir -p "$OUT"
rm -f "$OUT"/*.o
SAN=${EMU_SAN:--fsanitize=address,undefined}
for src in "$ROOT"/hermite-ugks_amd/csrc/*.hip; do
  obj="$OUT/$(basename "$src" .hip).o"
  g++ -std=c++20 -O1 -g $SAN -fno-omit-frame-pointer -x c++ -I "$HERE" -I "$ROOT/include" \
      -I "$ROOT/hermite-ugks_amd/csrc" -c "$src" -o "$obj" &
done
g++ -std=c++20 -O1 -g $SAN -I "$HERE" -I "$ROOT/include" -c "$HERE/emu_runtime.cpp" -o "$OUT/emu_runtime.o" &
g++ -std=c++20 -O1 -g $SAN -I "$ROOT/include" -c "$HERE/emu_driver.cpp" -o "$OUT/emu_driver.o" &
wait
g++ $SAN -o "$OUT/emu_driver" "$OUT"/*.o -lpthread
echo "$OUT/emu_driver"
# shared-library form (no sanitizer) for ctypes: HUGKS_LIB=<path> makes
# hugks.capi load it instead of the HIP build
if [ -n "$EMU_SO" ]; then
  for src in "$ROOT"/hermite-ugks_amd/csrc/*.hip "$HERE/emu_runtime.cpp"; do
    b=$(basename "$src"); b=${b%.*}
    g++ -std=c++20 -O2 -fPIC -x c++ -I "$HERE" -I "$ROOT/include" -I "$ROOT/hermite-ugks_amd/csrc" \
        -c "$src" -o "$OUT/so_$b.o" &
  done
  wait
  g++ -shared -o "$OUT/libhugks_emu.so" "$OUT"/so_*.o -lpthread
  echo "$OUT/libhugks_emu.so"
fi
