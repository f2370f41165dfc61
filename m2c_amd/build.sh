#!/bin/bash
# Builds libm2c_amd.so (gfx950) in-tree.  No FMA contraction and no fast-math:
# results must match the reference's non-contracted x86-64 arithmetic bit for
# bit.  Packed-FP32 instructions are disabled: some of their forms have given
# wrong results on this hardware.
set -e
cd "$(dirname "$0")"
HIPCC=${HIPCC:-/opt/rocm/bin/hipcc}
$HIPCC --offload-arch=gfx950 -O3 -std=c++17 -fPIC -shared \
  -ffp-contract=off -fno-fast-math \
  -Xclang -target-feature -Xclang -packed-fp32-ops \
  -Wall -Wno-unused-variable -Wno-unused-but-set-variable -Wno-unused-value -Wno-unused-result \
  -I../include -o libm2c_amd.so.tmp csrc/m2c_capi.hip "$@"
# replace the library in one step, so that a reader never sees a partial file
mv -f libm2c_amd.so.tmp libm2c_amd.so
