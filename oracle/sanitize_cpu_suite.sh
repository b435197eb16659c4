#!/bin/bash
# Runs the CPU test suite with the oracle and the library's host mirror built under
# AddressSanitizer + UndefinedBehaviorSanitizer (g++): any out-of-bounds access, use after free or
# undefined behaviour in that code stops the run with a report.  The device kernels are not
# instrumented (they run only in the GPU suite).
#   oracle/sanitize_cpu_suite.sh [extra pytest arguments]
set -e
ROOT="$(cd "$(dirname "$0")/.." && pwd)"
make -C "$ROOT/oracle" -s sanitize
make -C "$ROOT/zq_smokesimulation_amd/csrc" -s -j8 sanitize
export ZQ3D_ORACLE_LIB="$ROOT/oracle/liboracle_zq3d_san.so"
export ZQ3D_LIB="$ROOT/zq_smokesimulation_amd/lib/san/libzq_smoke3d.so"
export ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:alloc_dealloc_mismatch=0:detect_odr_violation=0"
export UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1"
export LD_PRELOAD="$(gcc -print-file-name=libasan.so) $(gcc -print-file-name=libubsan.so)"
cd "$ROOT"
exec python -m pytest tests -m "not gpu" -q -p no:cacheprovider "$@"
