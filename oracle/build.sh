#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/liboracle.so.
# No FMA contraction and no fast-math, so float and double runs keep the
# reference's operation order.
set -e
cd "$(dirname "$0")"
g++ -O2 -std=c++17 -fPIC -shared -ffp-contract=off -fno-fast-math -Wall -Wno-unused-function \
    -o liboracle.so oracle.cpp spaces.cpp
