#!/bin/bash
# Runs the oracle's CPU tests against its AddressSanitizer + UBSan build
# (host code only): make -C oracle asan, then pytest with libasan preloaded
# into the (uninstrumented) Python interpreter.
set -e
cd "$(dirname "$0")/.."
make -C oracle asan
ASAN_LIB=$(g++ -print-file-name=libasan.so)
UBSAN_LIB=$(g++ -print-file-name=libubsan.so)
if [ $# -eq 0 ]; then
  set -- tests/test_oracle_structure.py tests/test_oracle_basis.py tests/test_oracle_b2.py \
    tests/test_oracle_b3.py tests/test_oracle_gp.py
fi
ORACLE_SANITIZED=1 ASAN_OPTIONS=detect_leaks=0:halt_on_error=1 UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
  LD_PRELOAD="$ASAN_LIB $UBSAN_LIB" python -m pytest -q -p no:cacheprovider "$@"
