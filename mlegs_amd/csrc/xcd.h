// XCD-aware block order for 1D grids (device code only).
//
// Workgroups go to the 8 XCDs of the chip (each with its own L2) round-robin
// by linear id.  A grid padded to a multiple of 8 with its linear id remapped
// by xcd_logical runs consecutive logical ids -- blocks that share cache
// lines or table tiles -- on one XCD, in order.
#pragma once
#include <hip/hip_runtime.h>

namespace mlg {

// Logical block id of this block in a grid of nblk logical blocks launched
// as xcd_grid(nblk) blocks; -1 for a padding block.
// first: the linear id where the remapped range starts (a multiple of 8;
// the blocks before it keep their own order and get -1 here).
__device__ __forceinline__ int xcd_logical(int nblk, int first = 0) {
  const int nblkp = (nblk + 7) & ~7;
  const int bid = (int)blockIdx.x - first;
  if (bid < 0) return -1;
  const int logical = (bid & 7) * (nblkp >> 3) + (bid >> 3);
  return logical < nblk ? logical : -1;
}
__host__ __device__ inline int xcd_grid(int nblk) { return (nblk + 7) & ~7; }

}  // namespace mlg
