#!/bin/sh
# Converts the reference's golden grid fixture tests/gridone.h5 (HPG v2.0.1,
# Apache-2.0, Associated Universities, Inc.) into a raw little-endian array:
# dataset "grid", 91 x 91 complex<double> indexed [x][y] (x slowest), the
# window [9523, 9614) x [8522, 8613) of the 16384^2 grid of the GridOne test
# (tests/GridderStateUnitTests.cc:1343-1463).  Needs h5dump (HDF5 tools).
set -e
H5DUMP=${H5DUMP:-/opt/conda/bin/h5dump}
SRC=${1:-/root/reference/tests/gridone.h5}
OUT=$(dirname "$0")/gridone.bin
"$H5DUMP" -d grid -b MEMORY -o "$OUT" "$SRC" > /dev/null
ls -l "$OUT"
