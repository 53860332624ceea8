// partition_exchange.hpp -- how the lanes that share one system talk.
//
// swapc<K>(send, from_prev, from_next, h) gives every lane the K values of
// the lane h chunks before it and h chunks after it in the same system.
// Where that chunk does not exist the lane receives its own value (or 0 at
// the edge of the wavefront) -- no zeroing selects: the callers multiply
// those values by exact zeros.  Values never cross from one
// system to another, so a system with non-finite data cannot spoil another.
// all(pred) is true when pred holds on every lane that takes part in the
// exchange (wavefront or workgroup), so loops around swap stay uniform.
// max_f32(v) is the maximum of v over the chunks of the system.
//
// The LDS exchanges need a barrier before their write (every lane has read
// the previous exchange) and one after it.  all() is itself a workgroup
// barrier that every lane reaches after its last read, so the exchange that
// follows it skips its first barrier (the flag synced); the callers write
// LDS only through swapc between an all() and the next exchange.
#pragma once

#include <hip/hip_runtime.h>

namespace pentad {

// x direction: the P lanes of a system are consecutive lanes of a wavefront.
struct WaveEx {
  int lane; // lane id in the wavefront
  int p;    // chunk index = lane position inside the system's lane group
  int P;    // lanes per system
  template <int K, typename T>
  __device__ __forceinline__ void swapc(const T *send, T *from_prev, T *from_next,
                                        int h) const {
    if (h == 1) { // DPP wave shifts: no LDS round trip, edges of the wave get 0
#pragma unroll
      for (int k = 0; k < K; ++k) {
        from_prev[k] = dpp_shift<T, 0x138>(send[k]); // wave_shr:1
        from_next[k] = dpp_shift<T, 0x130>(send[k]); // wave_shl:1
        if (P < 64) { // several systems per wavefront: never read a neighbour system
          from_prev[k] = (p == 0) ? send[k] : from_prev[k];
          from_next[k] = (p == P - 1) ? send[k] : from_next[k];
        }
      }
    } else {
      const int sp = (p - h) >= 0 ? lane - h : lane;
      const int sn = (p + h) < P ? lane + h : lane;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        from_prev[k] = __shfl(send[k], sp, 64);
        from_next[k] = __shfl(send[k], sn, 64);
      }
    }
  }
  __device__ __forceinline__ bool all(bool pred) const { return __all(pred); }
  // maximum over aligned groups of P lanes (P a power of two <= 64)
  __device__ __forceinline__ float max_f32(float v) const {
    if (P >= 2) v = fmaxf(v, dpp_move<0xB1>(v));  // quad_perm [1,0,3,2]
    if (P >= 4) v = fmaxf(v, dpp_move<0x4E>(v));  // quad_perm [2,3,0,1]
    if (P >= 8) v = fmaxf(v, dpp_move<0x141>(v)); // row_half_mirror
    if (P >= 16) v = fmaxf(v, dpp_move<0x140>(v)); // row_mirror
    if (P >= 32)
      v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(
                                                 __builtin_bit_cast(int, v), 0x401F))); // lane ^ 16
    if (P >= 64)
      v = fmaxf(__builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 0)),
                __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 32)));
    return v;
  }

 private:
  // lane i receives lane i -/+ 1 of the whole wavefront (CTRL 0x138 wave_shr:1,
  // 0x130 wave_shl:1); lanes shifted in from outside the wavefront read 0
  template <typename T, int CTRL>
  __device__ __forceinline__ static T dpp_shift(T v) {
    if constexpr (sizeof(T) == 4) {
      return __builtin_bit_cast(
          T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, true));
    } else {
      const int2 w = __builtin_bit_cast(int2, v);
      const int2 r{__builtin_amdgcn_update_dpp(0, w.x, CTRL, 0xF, 0xF, true),
                   __builtin_amdgcn_update_dpp(0, w.y, CTRL, 0xF, 0xF, true)};
      return __builtin_bit_cast(T, r);
    }
  }
  // in-row permutation (quad_perm / mirrors): every source lane exists
  template <int CTRL>
  __device__ __forceinline__ static float dpp_move(float v) {
    return __builtin_bit_cast(
        float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
  }
};

// strided directions: thread t of the workgroup owns system t % S, chunk
// t / S; values travel through LDS laid out [K][P][S] (conflict-free: the S
// threads of one chunk touch S consecutive words).
struct LdsEx {
  unsigned char *raw; // at least K * P * S * sizeof(T) bytes
  int s, p;           // system, chunk of this thread
  int S, P;
  bool synced = false; // the last workgroup-wide step was all()
  template <int K, typename T>
  __device__ __forceinline__ void swapc(const T *send, T *from_prev, T *from_next,
                                        int h) {
    T *lds = reinterpret_cast<T *>(raw);
    const int PS = P * S;
    if (!synced) __syncthreads();
    synced = false;
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * PS + p * S + s] = send[k];
    __syncthreads();
    const int op = ((p - h) >= 0 ? p - h : p) * S + s;
    const int on = ((p + h) < P ? p + h : p) * S + s;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      from_prev[k] = lds[k * PS + op];
      from_next[k] = lds[k * PS + on];
    }
  }
  __device__ __forceinline__ bool all(bool pred) {
    const bool r = __syncthreads_and(pred) != 0;
    synced = true;
    return r;
  }
  __device__ __forceinline__ float max_f32(float v) {
    for (int h = 1; h < P; h <<= 1) {
      float a[1] = {v}, bp[1], bn[1];
      swapc<1>(a, bp, bn, h);
      v = fmaxf(v, fmaxf(bp[0], bn[0]));
    }
    return v;
  }
};

// Block index remap so that consecutive tiles land on the same XCD: the
// hardware deals workgroups to the 8 XCDs round-robin (block b -> XCD b % 8),
// and each XCD has its own L2, so tiles that share cache lines (adjacent
// systems of the strided directions) should share an XCD.  A bijection on
// [0, nblocks) for any nblocks.
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t nblocks) {
  constexpr int NX = 8;
  const int64_t q = nblocks / NX, r = nblocks % NX;
  const int64_t x = b % NX, idx = b / NX;
  return (x < r) ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

} // namespace pentad
