#!/bin/sh
# Builds the CPU oracle (test infrastructure only): an FP64 parity oracle and an
# FP32 port used as the CPU baseline.  Outputs stay in oracle/ (git-ignored .so).
set -e
here=$(cd "$(dirname "$0")" && pwd)
gcc -O2 -fopenmp -shared -fPIC -o "$here/liboracle_f64.so" "$here/mbx_oracle.c" -lm
gcc -O3 -fopenmp -shared -fPIC -DREAL=float -o "$here/liboracle_f32.so" "$here/mbx_oracle.c" -lm
