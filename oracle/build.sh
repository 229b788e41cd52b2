#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/:
#   liboracle.so       strict IEEE build: the parity checker
#   liboracle_fast.so  reference-like release flags (-O3 -ffast-math, OpenMP over
#                      copies, per-ISA clones): the CPU timing baseline only
set -e
cd "$(dirname "$0")"
CC=${CC:-gcc}
$CC -std=c11 -O2 -fno-fast-math -ffp-contract=off -fPIC -shared -Wall -Wno-unused-function \
    -o liboracle.so oracle.c -lm
$CC -std=c11 -O3 -ffast-math -fopenmp -DORC_FAST -fPIC -shared -Wall -Wno-unused-function \
    -o liboracle_fast.so oracle.c -lm
