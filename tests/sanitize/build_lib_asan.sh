#!/bin/bash
# Builds libceos_gpu with AddressSanitizer + UndefinedBehaviorSanitizer on its host code (the
# device code is the ordinary build: -Xarch_host keeps the sanitizers off the GPU pass) and the
# host-only driver tests/sanitize/lib_driver.cpp against it, into tests/sanitize/_build/.
# Usage: build_lib_asan.sh [jobs]
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/../.." && pwd)
SRC=$REPO/ceos-eems_amd/csrc
OUT=$HERE/_build
ROCM=${ROCM_PATH:-/opt/rocm}
mkdir -p "$OUT"
SAN="-Xarch_host -fsanitize=address -Xarch_host -fsanitize=undefined -Xarch_host -fno-omit-frame-pointer"
FLAGS="--offload-arch=gfx950 -O1 -g -std=c++17 -fPIC -Xclang -target-feature -Xclang -packed-fp32-ops -I$REPO/include"
pids=""
for s in ceos_gpu rve mlp surrogate; do
  "$ROCM/bin/hipcc" $FLAGS $SAN -c "$SRC/$s.hip" -o "$OUT/$s.o" &
  pids="$pids $!"
done
for p in $pids; do wait $p; done
"$ROCM/bin/hipcc" --offload-arch=gfx950 -shared -fPIC "$OUT/ceos_gpu.o" "$OUT/rve.o" "$OUT/mlp.o" "$OUT/surrogate.o" \
  -o "$OUT/libceos_gpu_asan.so" -lrocsolver -lrocblas
# the driver is host C++ only: the sanitizer runtime comes with its link (clang's, as the library's)
"$ROCM/llvm/bin/clang++" -O1 -g -std=c++17 -fsanitize=address,undefined -fno-omit-frame-pointer \
  -I"$REPO/include" "$HERE/lib_driver.cpp" -L"$OUT" -lceos_gpu_asan -L"$ROCM/lib" -lamdhip64 \
  -Wl,-rpath,'$ORIGIN' -Wl,-rpath,"$ROCM/lib" -o "$OUT/lib_driver"
rm -f "$OUT"/*.o
echo "built $OUT/lib_driver"
