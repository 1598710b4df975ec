#!/bin/bash
# Compile the reference's self-contained interpolation-weight sources
# (bfps/cpp/spline_n{1..10}.cpp, bfps/cpp/Lagrange_polys.cpp; they include
# only their own headers and <cmath>) into oracle/_ref/libref_interp.so.
# Test infrastructure only: used to pin the generated weight tables.
# The rest of the reference's hot path needs MPI, FFTW3-MPI and HDF5, which
# are absent here, so it is not built.
set -e
REF=${BFPS_REFERENCE:-/root/reference}/bfps/cpp
HERE=$(cd "$(dirname "$0")" && pwd)
if [ ! -d "$REF" ]; then
    echo "reference tree not present; skipping oracle/_ref build"
    exit 0
fi
mkdir -p "$HERE/_ref"
SRCS=""
for n in 1 2 3 4 5 6 7 8 9 10; do SRCS="$SRCS $REF/spline_n$n.cpp"; done
SRCS="$SRCS $REF/Lagrange_polys.cpp"
g++ -O2 -fPIC -shared -I"$REF" $SRCS -o "$HERE/_ref/libref_interp.so.tmp"
mv "$HERE/_ref/libref_interp.so.tmp" "$HERE/_ref/libref_interp.so"
echo "built $HERE/_ref/libref_interp.so"
