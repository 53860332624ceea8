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
// In the write-out each thread keeps one system for every row it writes, so
// the system's output base -- a division by the mapped extent -- is computed
// once per thread; a store instruction covers 64 / S rows of S consecutive
// elements.  (16-byte stores of two or four systems per thread measured no
// faster and wrote more partial lines.)  Workgroups with adjacent systems are
// dealt to the same XCD (xcd_tile), so the partial 128-byte lines they share
// meet in one L2.
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

// column col (0 <= col < S) of a tile of S systems: this lane's M rows
template <typename T, int M>
__device__ __forceinline__ void tile_put(T *tile, const T (&v)[M], int p, int S, int col) {
#pragma unroll
  for (int i = 0; i < M; ++i) tile[(p * M + i) * S + col + p] = v[i];
}

// write a complete tile of S = 1 << sshift systems (sys0 ...) row by row
template <typename T, int M>
__device__ __forceinline__ void tile_flush(const T *tile, int sshift, int len, T *out,
                                           int64_t n_sys, int64_t sys0, const XMap &map) {
  const int S = 1 << sshift;
  const int bshift = __builtin_ctz(blockDim.x);
  if (sshift > bshift) {
    // more systems than threads: each thread writes whole systems, the
    // threads of one instruction covering consecutive systems of a row
    for (int sb = 0; sb < S; sb += (int)blockDim.x) {
      const int s = sb + (int)threadIdx.x;
      const int64_t sys = sys0 + s;
      if (sys >= n_sys) continue;
      T *o = out + map.lo(sys) * map.out_lo + map.hi(sys) * map.out_hi;
      for (int row = 0; row < len; ++row)
        o[(int64_t)row * map.out_row] = tile[row * S + s + row / M];
    }
    return;
  }
  const int s = (int)threadIdx.x & (S - 1);
  const int64_t sys = sys0 + s;
  if (sys >= n_sys) return;
  T *o = out + map.lo(sys) * map.out_lo + map.hi(sys) * map.out_hi;
  const int rstep = (int)blockDim.x >> sshift; // >= 1 here
  if (S * (int)sizeof(T) >= 128) {
    // whole 128-byte lines: streaming stores, so that the solution does not
    // displace lines still being read from L2
    for (int row = (int)threadIdx.x >> sshift; row < len; row += rstep)
      __builtin_nontemporal_store(tile[row * S + s + row / M], o + (int64_t)row * map.out_row);
  } else {
    for (int row = (int)threadIdx.x >> sshift; row < len; row += rstep)
      o[(int64_t)row * map.out_row] = tile[row * S + s + row / M];
  }
}

// Tiles of exactly 16 systems without the per-chunk pad: chunk p's column
// col is stored at col ^ (p & 15).  A wavefront's puts (one column, 64
// chunks) then fall on 16 bank pairs instead of all on one, and the rows the
// flush reads stay whole 128-byte runs; the pad's LDS is saved, which is
// what lets two workgroups' tiles of 640-row lines (80 KB each) share one
// CU's 160 KB.
template <typename T, int M>
__device__ __forceinline__ void tile_put_x16(T *tile, const T (&v)[M], int p, int col) {
#pragma unroll
  for (int i = 0; i < M; ++i) tile[(p * M + i) * 16 + (col ^ (p & 15))] = v[i];
}

template <typename T, int M>
__device__ __forceinline__ void tile_flush_x16(const T *tile, int len, T *out, int64_t n_sys,
                                               int64_t sys0, const XMap &map) {
  const int s = (int)threadIdx.x & 15;
  const int64_t sys = sys0 + s;
  if (sys >= n_sys) return;
  T *o = out + map.lo(sys) * map.out_lo + map.hi(sys) * map.out_hi;
  const int rstep = (int)blockDim.x >> 4;
  for (int row = (int)threadIdx.x >> 4; row < len; row += rstep)
    __builtin_nontemporal_store(tile[row * 16 + (s ^ ((row / M) & 15))],
                                o + (int64_t)row * map.out_row);
}

template <typename T, int M>
__device__ __forceinline__ void stage_rotated_store(T *tile, const T (&v)[M], int p,
                                                    int lshift, int len, T *out,
                                                    int64_t n_sys, int64_t blk,
                                                    const XMap &map) {
  const int sshift = __builtin_ctz(blockDim.x) - lshift;
  const int sl = (int)threadIdx.x >> lshift;      // this lane's system
  __syncthreads();                                // LDS may still hold other data
  tile_put<T, M>(tile, v, p, 1 << sshift, sl);
  __syncthreads();
  tile_flush<T, M>(tile, sshift, len, out, n_sys, blk << sshift, map);
}

} // namespace pentad
