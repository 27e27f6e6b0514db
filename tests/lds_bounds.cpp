// Prints the LDS bounds of fast_mul_ntt_amd/csrc/fmh_internal.h (compiled
// with g++ against the CPU emulation of the HIP header):
//   kLdsPerGroup, the largest dynamic LDS of a column tile, of a row tile.
#include <stdio.h>

#include "fmh_internal.h"

int main() {
    using namespace fmh;
    printf("%zu %zu %zu %d %d\n", kLdsPerGroup, col_lds_bytes(kMaxColTileLog, 0), row_lds_bytes(2, kMaxRowTileLog, 0),
           kMaxColTileLog, kMaxRowTileLog);
    return 0;
}
