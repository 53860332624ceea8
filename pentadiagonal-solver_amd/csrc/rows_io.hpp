// rows_io.hpp -- load/store of M consecutive elements owned by one lane.
//
// The widest naturally aligned vector that divides the chunk is used:
// 16 bytes when M * sizeof(T) is a multiple of 16, else 8 bytes, else
// scalar.  Chunk starts are multiples of M elements, so the alignment holds
// whenever the array base is 16-byte aligned (checked on the host).
#pragma once

#include <hip/hip_runtime.h>

namespace pentad {

template <typename T, int M> struct RowsIO {
  static constexpr int BYTES = M * (int)sizeof(T);
  static constexpr int VB = (BYTES % 16 == 0) ? 16 : (BYTES % 8 == 0 ? 8 : (int)sizeof(T));
  static constexpr int PER = VB / (int)sizeof(T);

  __device__ __forceinline__ static void load(const T *__restrict__ src, T (&v)[M]) {
    if constexpr (VB == 16 && sizeof(T) == 4) {
      const float4 *q = reinterpret_cast<const float4 *>(src);
#pragma unroll
      for (int k = 0; k < M / 4; ++k) {
        const float4 t = q[k];
        v[4 * k] = t.x; v[4 * k + 1] = t.y; v[4 * k + 2] = t.z; v[4 * k + 3] = t.w;
      }
    } else if constexpr (VB == 8 && sizeof(T) == 4) {
      const float2 *q = reinterpret_cast<const float2 *>(src);
#pragma unroll
      for (int k = 0; k < M / 2; ++k) {
        const float2 t = q[k];
        v[2 * k] = t.x; v[2 * k + 1] = t.y;
      }
    } else if constexpr (VB == 16 && sizeof(T) == 8) {
      const double2 *q = reinterpret_cast<const double2 *>(src);
#pragma unroll
      for (int k = 0; k < M / 2; ++k) {
        const double2 t = q[k];
        v[2 * k] = t.x; v[2 * k + 1] = t.y;
      }
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) v[k] = src[k];
    }
  }

  __device__ __forceinline__ static void store(T *__restrict__ dst, const T (&v)[M]) {
    if constexpr (VB == 16 && sizeof(T) == 4) {
      float4 *q = reinterpret_cast<float4 *>(dst);
#pragma unroll
      for (int k = 0; k < M / 4; ++k)
        q[k] = float4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
    } else if constexpr (VB == 8 && sizeof(T) == 4) {
      float2 *q = reinterpret_cast<float2 *>(dst);
#pragma unroll
      for (int k = 0; k < M / 2; ++k) q[k] = float2{v[2 * k], v[2 * k + 1]};
    } else if constexpr (VB == 16 && sizeof(T) == 8) {
      double2 *q = reinterpret_cast<double2 *>(dst);
#pragma unroll
      for (int k = 0; k < M / 2; ++k) q[k] = double2{v[2 * k], v[2 * k + 1]};
    } else {
#pragma unroll
      for (int k = 0; k < M; ++k) dst[k] = v[k];
    }
  }
};

} // namespace pentad
