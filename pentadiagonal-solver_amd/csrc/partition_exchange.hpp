// partition_exchange.hpp -- how the lanes that share one system talk.
//
// swap<K>(send, from_prev, from_next, h) gives every lane the K values of
// the lane h chunks before it and h chunks after it in the same system
// (zeros when that chunk does not exist).  all(pred) is true when pred holds
// on every lane that takes part in the exchange (wavefront or workgroup), so
// loops around swap stay uniform.  max_all(v) is the maximum of v over the
// chunks of the system.
#pragma once

#include <hip/hip_runtime.h>

namespace pentad {

// x direction: the P lanes of a system are consecutive lanes of a wavefront.
struct WaveEx {
  int lane; // lane id in the wavefront
  int p;    // chunk index = lane position inside the system's lane group
  int P;    // lanes per system
  template <int K, typename T>
  __device__ __forceinline__ void swap(const T *send, T *from_prev,
                                       T *from_next, int h) const {
    const bool hp = (p - h) >= 0, hn = (p + h) < P;
    const int sp = hp ? lane - h : lane;
    const int sn = hn ? lane + h : lane;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T vp = __shfl(send[k], sp, 64);
      const T vn = __shfl(send[k], sn, 64);
      from_prev[k] = hp ? vp : T(0);
      from_next[k] = hn ? vn : T(0);
    }
  }
  __device__ __forceinline__ bool all(bool pred) const { return __all(pred); }
};

// strided directions: thread t of the workgroup owns system t % S, chunk
// t / S; values travel through LDS laid out [K][P][S] (conflict-free: the S
// threads of one chunk touch S consecutive words).
struct LdsEx {
  unsigned char *raw; // at least K * P * S * sizeof(T) bytes
  int s, p;           // system, chunk of this thread
  int S, P;
  template <int K, typename T>
  __device__ __forceinline__ void swap(const T *send, T *from_prev,
                                       T *from_next, int h) const {
    T *lds = reinterpret_cast<T *>(raw);
    const bool hp = (p - h) >= 0, hn = (p + h) < P;
    const int PS = P * S;
    __syncthreads(); // earlier readers are done with the buffer
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * PS + p * S + s] = send[k];
    __syncthreads();
    const int op = (hp ? p - h : p) * S + s;
    const int on = (hn ? p + h : p) * S + s;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const T vp = lds[k * PS + op];
      const T vn = lds[k * PS + on];
      from_prev[k] = hp ? vp : T(0);
      from_next[k] = hn ? vn : T(0);
    }
  }
  __device__ __forceinline__ bool all(bool pred) const {
    return __syncthreads_and(pred) != 0;
  }
};

template <typename Ex>
__device__ __forceinline__ double max_all(Ex &ex, double v, int P) {
  for (int h = 1; h < P; h <<= 1) {
    double a[1] = {v}, bp[1], bn[1];
    ex.template swap<1>(a, bp, bn, h);
    v = fmax(v, fmax(bp[0], bn[0]));
  }
  return v;
}

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
