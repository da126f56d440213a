#!/bin/sh
# Builds the CPU oracle (test infrastructure) as oracle/liboracle.so.
# Flags follow the reference build (CMakeLists.txt:5,15: -O3, OpenMP, no
# -march => x86-64 baseline SSE2, no FMA) plus -ffp-contract=off.
set -e
cd "$(dirname "$0")"
gcc -O3 -fopenmp -ffp-contract=off -fno-fast-math -fPIC -shared -std=c99 \
    -o liboracle.so fr_oracle.c -lm
