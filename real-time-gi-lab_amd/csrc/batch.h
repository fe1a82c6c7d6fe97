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
// The logical block of this workgroup in a launch over a work queue (queue: VQ_LANES zeroed counters VQ_STRIDE words
// apart). XCD x owns the contiguous logical range xcd_block gives it and takes its blocks in order from counter x; a
// workgroup whose XCD's range is used up takes from the next XCD's counter, and so on round the eight. So each L2
// serves a few whole worlds as under the static remap, while an XCD that runs ahead finishes by taking the others'
// remaining blocks instead of idling at the end. Every workgroup takes exactly one logical block: each increment below
// a range's size names a distinct block of it, and there are as many logical blocks as workgroups, so a workgroup
// finds every range used up only if all blocks are taken (then it returns ~0u). For the trace+merge steps of 64 city
// worlds this took the trace stage from 14.50 to 13.6 ms; on a mixed batch's launches it lost to the hardware order
// (32 mixed worlds: 4.09 -> 4.36 ms), so those keep it.
__device__ __forceinline__ uint32_t queue_block(uint32_t* queue) {
  __shared__ uint32_t s_lb;
  if (threadIdx.x == 0) {
    const uint32_t X = 8u, n = gridDim.x, q = n / X, r = n % X, x = blockIdx.x % X;
    uint32_t lb = ~0u;
    for (uint32_t k = 0; k < X; k++) {
      const uint32_t xx = (x + k) % X, size = q + (xx < r ? 1u : 0u);
      if (size == 0u) continue;
      const uint32_t i = atomicAdd(queue + xx * VQ_STRIDE, 1u);
      if (i < size) {
        lb = (xx < r ? xx * (q + 1u) : r * (q + 1u) + (xx - r) * q) + i;
        break;
      }
    }
    s_lb = lb;
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(s_lb);  // uniform: the world's descriptor is then read with scalar loads
}
