// rotated_store.hpp -- write a workgroup's solutions transposed.
//
// The x-like kernels own S = blockDim / L consecutive systems per workgroup
// (L lanes per system, M rows per lane).  The ADI driver wants the solution
// in [row][system] layout, where the next direction's lines are contiguous.
// Each lane drops its M values into an LDS tile [row][S] (one padding
// element per chunk keeps the lanes of a wavefront on distinct banks), then
// the workgroup writes the tile row by row: S consecutive elements per row,
// so every global store instruction covers whole runs instead of one scattered
// element per lane.  Workgroups with adjacent systems are dealt to the same
// XCD (xcd_tile), so the partial 128-byte lines they share meet in one L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pentad_internal.hpp"

namespace pentad {

template <typename T>
__host__ __device__ constexpr size_t rotated_tile_bytes(int block, int M, int64_t len) {
  // S * len elements + one pad per chunk of M rows
  return (size_t)sizeof(T) * ((size_t)(block) / (size_t)(len / M) * (size_t)len + (size_t)(len / M));
}

template <typename T, int M>
__device__ __forceinline__ void stage_rotated_store(T *tile, const T (&v)[M], int p,
                                                    int L, int64_t len, T *out,
                                                    int64_t n_sys, int64_t blk,
                                                    const XMap &map) {
  const int S = (int)blockDim.x / L;        // systems in this workgroup
  const int sl = (int)threadIdx.x / L;      // this lane's system
  __syncthreads();                          // LDS may still hold other data
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int row = p * M + i;
    tile[row * S + sl + p] = v[i];
  }
  __syncthreads();
  const int64_t sys0 = blk * S;
  const int total = (int)len * S;
  for (int e = threadIdx.x; e < total; e += blockDim.x) {
    const int row = e / S, s = e - row * S;
    const int64_t sys = sys0 + s;
    if (sys < n_sys)
      out[map.lo(sys) * map.out_lo + (int64_t)row * map.out_row + map.hi(sys) * map.out_hi] =
          tile[e + row / M];
  }
}

} // namespace pentad
