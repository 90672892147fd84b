#!/bin/sh
# Host-side AddressSanitizer build of the C ABI and its driver (CPU only).
#
# The host-heavy translation units (qj_api.hip: plans, composition, error
# strings; qj_ensemble.hip: recording and grouping) are compiled with the
# sanitizer on their host pass only; the kernel translation units are linked
# from the ordinary build in quantum-jax_amd/_build (build.py must have run).
# The driver, tests/asan/abi_driver.cpp, exercises the argument-checking and
# error paths without a GPU.  Outputs go to tests/asan/_build (git-ignored).
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
OUT="$HERE/_build"
OBJ="$ROOT/quantum-jax_amd/_build"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
CLANGXX=${CLANGXX:-/opt/rocm/llvm/bin/clang++}
mkdir -p "$OUT"
INC="-I $ROOT/quantum-jax_amd/csrc -I $ROOT/include"
for f in qj_api qj_ensemble; do
  "$HIPCC" -O1 -g -std=c++17 -fPIC --offload-arch=gfx950 \
    -Xclang -target-feature -Xclang -packed-fp32-ops -Wno-unused-command-line-argument \
    -Xarch_host -fsanitize=address -Xarch_host -fno-omit-frame-pointer \
    $INC -c "$ROOT/quantum-jax_amd/csrc/$f.hip" -o "$OUT/$f.o"
done
KOBJ=""
for o in "$OBJ"/*.o; do
  case "$(basename "$o")" in qj_api.o|qj_ensemble.o) ;; *) KOBJ="$KOBJ $o" ;; esac
done
"$HIPCC" -shared -fPIC --offload-arch=gfx950 -Xarch_host -fsanitize=address \
  "$OUT/qj_api.o" "$OUT/qj_ensemble.o" $KOBJ -o "$OUT/libqjamd_asan.so"
"$CLANGXX" -O1 -g -Xarch_host -fsanitize=address -fno-omit-frame-pointer -I "$ROOT/include" \
  "$HERE/abi_driver.cpp" -L "$OUT" -lqjamd_asan -Wl,-rpath,"$OUT" -o "$OUT/abi_driver"
ASAN_OPTIONS=detect_leaks=1:abort_on_error=0 "$OUT/abi_driver"
