#!/bin/sh
# Builds the CPU oracle (test infrastructure) into oracle/liboracle.so.
# The reference's own path cannot be built here (CUDA + cuStateVec + a
# pennylane-lightning checkout fetched at configure time), so there is no
# oracle/_ref build: see DESIGN.md.
set -e
here="$(cd "$(dirname "$0")" && pwd)"
gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off -fcx-limited-range \
    -fPIC -shared -std=gnu11 -Wall -Wextra -Wno-unused-parameter \
    -o "$here/liboracle.so.tmp" "$here/oracle.c" -lm
mv "$here/liboracle.so.tmp" "$here/liboracle.so"
