#!/bin/bash
# Builds libfmh_emu.so: the library's HIP sources compiled for the CPU with
# g++ and AddressSanitizer against tests/emu/hip/hip_runtime.h.
# usage: build_emu.sh OUTDIR
set -e
ROOT=$(cd "$(dirname "$0")/../.." && pwd)
OUT=${1:?outdir}
mkdir -p "$OUT"
sed 's/extern __shared__ u64 lds\[\];/u64* lds = (u64*)emu_dyn_lds();/' \
    "$ROOT/fast_mul_ntt_amd/csrc/fmh_ntt.hip" > "$OUT/fmh_ntt_emu.cpp"
g++ -std=c++20 -O1 -g -fsanitize=address -fno-omit-frame-pointer -fPIC -shared -w -DFMH_SCAN_TILE=16 -DFMH_PIN_CHUNK=4096 -DFMH_OVERLAP_BYTES=16384 -DFMH_D2H_RAMP_MIN=16384 -DFMH_FUSED_CHUNKS=4 \
    -I "$ROOT/tests/emu" -I "$ROOT/fast_mul_ntt_amd/csrc" \
    "$OUT/fmh_ntt_emu.cpp" -x c++ "$ROOT/fast_mul_ntt_amd/csrc/fmh_api.hip" \
    "$ROOT/fast_mul_ntt_amd/csrc/fmh_serve.hip" -x none \
    "$ROOT/tests/emu/emu_runtime.cpp" -o "$OUT/libfmh_emu.so" -lpthread
