// Batched launches: one launch runs a stage for many worlds. The grid's blocks are split among the worlds by a
// prefix array (world w owns blocks [start[w], start[w + 1])); a block finds its world by binary search and runs the
// stage's body on that world's descriptor (device memory, read with scalar loads since it is uniform per block).
#pragma once

#include <hip/hip_runtime.h>

#include "rtgi_internal.h"

__device__ __forceinline__ int batch_world(const BatchK& b, uint32_t blk) {
  int lo = 0, hi = b.n;
  while (hi - lo > 1) {
    int m = (lo + hi) >> 1;
    if (b.start[m] <= blk)
      lo = m;
    else
      hi = m;
  }
  return lo;
}
