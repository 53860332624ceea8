// rotated_store.hpp -- write a workgroup's solutions transposed.
//
// The x-like kernels own S = blockDim / L consecutive systems per workgroup
// (L lanes per system, M rows per lane; blockDim and L powers of two).  The
// ADI driver wants the solution in [row][system] layout, where the next
// direction's lines are contiguous.  Each lane drops its M values into an
// LDS tile [row][S] (one padding element per chunk keeps the lanes of a
// wavefront on distinct banks), then the workgroup writes the tile row by
// row: S consecutive elements per row, so every global store instruction
// covers whole runs instead of one scattered element per lane.
//
// In the write-out each thread keeps one system (or V adjacent systems) for
// every row it writes, so the system's output base -- a division by the
// mapped extent -- is computed once per thread.  When the S systems are
// adjacent in the output (unit system stride, no wrap inside the
// workgroup, 16-byte aligned rows), each thread stores V = 16 / sizeof(T)
// systems of a row with one 16-byte store.  Workgroups with adjacent
// systems are dealt to the same XCD (xcd_tile), so the partial 128-byte
// lines they share meet in one L2.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pentad_internal.hpp"
#include "rows_io.hpp"

namespace pentad {

template <typename T>
__host__ __device__ constexpr size_t rotated_tile_bytes(int block, int M, int64_t len) {
  // S * len elements + one pad per chunk of M rows
  return (size_t)sizeof(T) * ((size_t)(block) / (size_t)(len / M) * (size_t)len + (size_t)(len / M));
}

template <typename T, int M>
__device__ __forceinline__ void stage_rotated_store(T *tile, const T (&v)[M], int p,
                                                    int lshift, int len, T *out,
                                                    int64_t n_sys, int64_t blk,
                                                    const XMap &map) {
  const int sshift = __builtin_ctz(blockDim.x) - lshift;
  const int S = 1 << sshift;                      // systems in this workgroup
  const int sl = (int)threadIdx.x >> lshift;      // this lane's system
  __syncthreads();                                // LDS may still hold other data
#pragma unroll
  for (int i = 0; i < M; ++i) tile[(p * M + i) * S + sl + p] = v[i];
  __syncthreads();
  const int64_t sys0 = blk * S;
  constexpr int V = 16 / (int)sizeof(T);
  const bool vec = S >= V && map.out_lo == 1 && map.S1 % S == 0 &&
                   map.out_row % V == 0 && map.out_hi % V == 0 &&
                   ((uintptr_t)out & 15) == 0 && sys0 + S <= n_sys;
  if (vec) {
    const int gshift = sshift - __builtin_ctz(V); // thread groups of V systems
    const int s = ((int)threadIdx.x & ((1 << gshift) - 1)) * V;
    const int64_t sys = sys0 + s;
    T *o = out + map.lo(sys) + map.hi(sys) * map.out_hi;
    const int rstep = (int)blockDim.x >> gshift;
    for (int row = (int)threadIdx.x >> gshift; row < len; row += rstep) {
      const T *t = tile + row * S + s + row / M;
      T w[V];
#pragma unroll
      for (int k = 0; k < V; ++k) w[k] = t[k];
      RowsIO<T, V>::store(o + (int64_t)row * map.out_row, w);
    }
  } else {
    const int s = (int)threadIdx.x & (S - 1);
    const int64_t sys = sys0 + s;
    if (sys >= n_sys) return;
    T *o = out + map.lo(sys) * map.out_lo + map.hi(sys) * map.out_hi;
    const int rstep = (int)blockDim.x >> sshift;
    for (int row = (int)threadIdx.x >> sshift; row < len; row += rstep)
      o[(int64_t)row * map.out_row] = tile[row * S + s + row / M];
  }
}

} // namespace pentad
