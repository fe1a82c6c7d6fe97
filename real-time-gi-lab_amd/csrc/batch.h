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

// Workgroups are dispatched to the 8 XCDs round-robin (block i runs on XCD i mod 8), and each XCD has its own L2.
// xcd_block maps hardware block i to a logical block so that every XCD runs one contiguous logical range: in a
// batch, a few whole worlds per XCD instead of a slice of every world, so each L2 holds fewer worlds' clipmaps. It is a
// bijection on [0, nblk), so no work moves between blocks, only where it runs.
__device__ __forceinline__ uint32_t xcd_block(uint32_t blk, uint32_t nblk) {
  const uint32_t X = 8u;
  uint32_t x = blk % X, i = blk / X, q = nblk / X, r = nblk % X;
  return x < r ? x * (q + 1u) + i : r * (q + 1u) + (x - r) * q + i;
}
// The logical block of this workgroup in a batched launch. The remap pays when the worlds of the launch cost about
// the same per block (a uniform batch: +8 % for 16 city worlds); with worlds of unequal cost, a contiguous range per
// XCD leaves the XCD holding the costliest worlds running alone at the end (-8 % for a mixed batch), so those
// launches keep the hardware order.
__device__ __forceinline__ uint32_t batch_block(const BatchK& b) {
  return b.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
}
