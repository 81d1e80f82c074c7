#!/bin/bash
# The CPU test suite under AddressSanitizer and UndefinedBehaviorSanitizer: the oracle built with
# them (clang), the library's host code built with them (hipcc, host pass only), both loaded into
# one Python process with clang's ASan runtime preloaded. Runs on the CPU; no GPU is touched.
#   bash tests/sanitize.sh [pytest args]
set -euo pipefail
ROOT=$(cd "$(dirname "$0")/.." && pwd)
CLANG=/opt/rocm/llvm/bin/clang
RT=$($CLANG -print-file-name=libclang_rt.asan-x86_64.so)
if [ -z "${SKIP_BUILD:-}" ]; then  # SKIP_BUILD=1: reuse the sanitized builds already made
    make -s -C "$ROOT/oracle" san CC="$CLANG"
    (cd "$ROOT" && python -m tessera_amd.build --sanitize)
fi
cd "$ROOT"
export ASAN_OPTIONS=detect_leaks=0:halt_on_error=1:abort_on_error=1
export UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1
TSR_LIB="$ROOT/tessera_amd/_san/libtessera_mi355x.so" TSR_ORACLE_LIB="$ROOT/oracle/_san/liboracle.so" \
    LD_LIBRARY_PATH="$(dirname "$RT"):${LD_LIBRARY_PATH:-}" LD_PRELOAD="$RT" python -m pytest tests -m "not gpu" -x -q -p no:cacheprovider "$@"
