// kernels_partition.hip -- register-resident partitioned solves for gfx950.
//
// k_part_x   x direction (contiguous lines).  A wavefront holds 64/L systems,
//            L lanes per system, each lane M consecutive rows loaded with
//            16-byte vector loads.  The reduced block system is exchanged
//            between the L lanes with wavefront shuffles (ds_bpermute).
// k_part_yz  strided directions.  A workgroup holds S adjacent systems and
//            P = len/M chunks of each; thread t owns system t % S, chunk
//            t / S, so at every row the S lanes of a chunk read S consecutive
//            words (coalesced).  The reduced system goes through LDS.
//
// Both keep every row in registers between load and store: HBM traffic is
// the compulsory 5 coefficient reads + 1 RHS read + 1 solution write per row.
#include "partition_core.hpp"
#include "pentad_internal.hpp"

namespace pentad {

// ---------------------------------------------------------------------------
// exchange policies
// ---------------------------------------------------------------------------
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
};

template <typename T> struct LdsEx {
  T *lds;   // [K][P][S]
  int s, p; // system, chunk of this thread
  int S, P;
  template <int K, typename U>
  __device__ __forceinline__ void swap(const U *send, U *from_prev,
                                       U *from_next, int h) const {
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
      const U vp = lds[k * PS + op];
      const U vn = lds[k * PS + on];
      from_prev[k] = hp ? vp : U(0);
      from_next[k] = hn ? vn : U(0);
    }
  }
};

// ---------------------------------------------------------------------------
// vector load/store of M consecutive elements (16-byte aligned)
// ---------------------------------------------------------------------------
template <typename T, int M>
__device__ __forceinline__ void load_rows(const T *__restrict__ src, T (&v)[M]) {
  constexpr int PER = 16 / sizeof(T);
  static_assert(M % PER == 0, "row count must fill 16-byte vectors");
  using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
  const V *q = reinterpret_cast<const V *>(src);
#pragma unroll
  for (int k = 0; k < M / PER; ++k) {
    const V t = q[k];
    if constexpr (sizeof(T) == 8) {
      v[2 * k] = t.x;
      v[2 * k + 1] = t.y;
    } else {
      v[4 * k] = t.x;
      v[4 * k + 1] = t.y;
      v[4 * k + 2] = t.z;
      v[4 * k + 3] = t.w;
    }
  }
}

template <typename T, int M>
__device__ __forceinline__ void store_rows(T *__restrict__ dst, const T (&v)[M]) {
  constexpr int PER = 16 / sizeof(T);
  using V = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
  V *q = reinterpret_cast<V *>(dst);
#pragma unroll
  for (int k = 0; k < M / PER; ++k) {
    V t;
    if constexpr (sizeof(T) == 8) {
      t.x = v[2 * k];
      t.y = v[2 * k + 1];
    } else {
      t.x = v[4 * k];
      t.y = v[4 * k + 1];
      t.z = v[4 * k + 2];
      t.w = v[4 * k + 3];
    }
    q[k] = t;
  }
}

// ---------------------------------------------------------------------------
// x direction
// ---------------------------------------------------------------------------
template <typename T, int M>
__global__ void __launch_bounds__(256)
    k_part_x(const T *__restrict__ ds, const T *__restrict__ dl,
             const T *__restrict__ d, const T *__restrict__ du,
             const T *__restrict__ dw, T *__restrict__ x, int64_t n_sys,
             int64_t len, int lshift /* log2 lanes per system */) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int L = 1 << lshift;
  const int64_t sys = g >> lshift;
  const int p = (int)(g & (L - 1));
  const bool valid = sys < n_sys;
  const int64_t off = (valid ? sys : n_sys - 1) * len + (int64_t)p * M;

  T s[M], l[M], dd[M], u[M], w[M], f[M];
  load_rows<T, M>(ds + off, s);
  load_rows<T, M>(dl + off, l);
  load_rows<T, M>(d + off, dd);
  load_rows<T, M>(du + off, u);
  load_rows<T, M>(dw + off, w);
  load_rows<T, M>(x + off, f);

  WaveEx ex{(int)(threadIdx.x & 63), p, L};
  part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
  if (valid) store_rows<T, M>(x + off, f);
}

// ---------------------------------------------------------------------------
// strided (y/z) directions
// ---------------------------------------------------------------------------
template <typename T, int M, int MAXT>
__global__ void __launch_bounds__(MAXT)
    k_part_yz(const T *__restrict__ ds, const T *__restrict__ dl,
              const T *__restrict__ d, const T *__restrict__ du,
              const T *__restrict__ dw, T *__restrict__ x, Lines ln, int S,
              int P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int t = threadIdx.x;
  const int sl = t % S;
  const int p = t / S;
  const int64_t sys = (int64_t)blockIdx.x * S + sl;
  const bool valid = sys < ln.n_sys;
  const int64_t st = ln.stride;
  const int64_t b = ln.base(valid ? sys : ln.n_sys - 1) + (int64_t)p * M * st;

  T s[M], l[M], dd[M], u[M], w[M], f[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int64_t e = b + i * st;
    s[i] = ds[e];
    l[i] = dl[e];
    dd[i] = d[e];
    u[i] = du[e];
    w[i] = dw[e];
    f[i] = x[e];
  }

  LdsEx<T> ex{lds, sl, p, S, P};
  part_solve<T, M>(s, l, dd, u, w, f, p, P, ex);
  if (valid) {
#pragma unroll
    for (int i = 0; i < M; ++i) x[b + i * st] = f[i];
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {

template <typename T, int M>
hipError_t go_x(const Diags<T> &a, T *x, const Lines &ln, int lanes,
                hipStream_t st) {
  int lshift = 0;
  while ((1 << lshift) < lanes) ++lshift;
  const int block = 256;
  const int64_t threads = ln.n_sys * lanes;
  const int64_t grid = (threads + block - 1) / block;
  hipLaunchKernelGGL((k_part_x<T, M>), dim3((unsigned)grid), dim3(block), 0,
                     st, a.ds, a.dl, a.d, a.du, a.dw, x, ln.n_sys, ln.len,
                     lshift);
  return hipGetLastError();
}

// Largest workgroup each (type, M) instantiation is compiled for.  With W
// waves per SIMD a lane may use 512/W VGPRs; a 1024-thread workgroup means
// 4 waves per SIMD (128 VGPRs), which only the small chunks fit without
// spilling (checked with -Rpass-analysis=kernel-resource-usage).
constexpr int yz_maxt_rt(int elem_bytes, int rows) {
  return (elem_bytes == 8) ? (rows <= 4 ? 1024 : 512)
                           : (rows <= 12 ? 1024 : 512);
}
template <typename T, int M> constexpr int yz_maxt() {
  return yz_maxt_rt((int)sizeof(T), M);
}

template <typename T, int M>
hipError_t go_yz(const Diags<T> &a, T *x, const Lines &ln, int S,
                 hipStream_t st) {
  constexpr int MAXT = yz_maxt<T, M>();
  const int P = (int)(ln.len / M);
  const int block = S * P;
  if (block > MAXT || block % 64 != 0) return hipErrorInvalidConfiguration;
  const size_t smem = (size_t)10 * S * P * sizeof(T);
  const int64_t grid = (ln.n_sys + S - 1) / S;
  if (smem > 65536) {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_part_yz<T, M, MAXT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_part_yz<T, M, MAXT>), dim3((unsigned)grid),
                     dim3(block), smem, st, a.ds, a.dl, a.d, a.du, a.dw, x, ln,
                     S, P);
  return hipGetLastError();
}

} // namespace

bool part_x_supported(int elem_bytes, int rows, int lanes) {
  if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0) return false;
  if (elem_bytes == 8) return rows == 4 || rows == 6 || rows == 8 || rows == 10 || rows == 12 || rows == 16;
  return rows == 4 || rows == 8 || rows == 12 || rows == 16;
}

template <typename T>
hipError_t launch_part_x(const Diags<T> &a, T *x, const Lines &ln, int rows,
                         int lanes, hipStream_t st) {
  if (!part_x_supported(sizeof(T), rows, lanes) || rows * lanes != ln.len)
    return hipErrorInvalidValue;
  if constexpr (sizeof(T) == 8) {
    switch (rows) {
    case 4: return go_x<T, 4>(a, x, ln, lanes, st);
    case 6: return go_x<T, 6>(a, x, ln, lanes, st);
    case 8: return go_x<T, 8>(a, x, ln, lanes, st);
    case 10: return go_x<T, 10>(a, x, ln, lanes, st);
    case 12: return go_x<T, 12>(a, x, ln, lanes, st);
    case 16: return go_x<T, 16>(a, x, ln, lanes, st);
    }
  } else {
    switch (rows) {
    case 4: return go_x<T, 4>(a, x, ln, lanes, st);
    case 8: return go_x<T, 8>(a, x, ln, lanes, st);
    case 12: return go_x<T, 12>(a, x, ln, lanes, st);
    case 16: return go_x<T, 16>(a, x, ln, lanes, st);
    }
  }
  return hipErrorInvalidValue;
}

bool part_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 12 && rows != 16)
    return false;
  if (len % rows != 0) return false;
  const int64_t P = len / rows;
  const int64_t block = P * sys_per_wg;
  return block <= yz_maxt_rt(elem_bytes, rows) && block % 64 == 0;
}

template <typename T>
hipError_t launch_part_yz(const Diags<T> &a, T *x, const Lines &ln, int rows,
                          int sys_per_wg, hipStream_t st) {
  if (!part_yz_supported(sizeof(T), rows, sys_per_wg, ln.len))
    return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_yz<T, 4>(a, x, ln, sys_per_wg, st);
  case 6: return go_yz<T, 6>(a, x, ln, sys_per_wg, st);
  case 8: return go_yz<T, 8>(a, x, ln, sys_per_wg, st);
  case 12: return go_yz<T, 12>(a, x, ln, sys_per_wg, st);
  case 16: return go_yz<T, 16>(a, x, ln, sys_per_wg, st);
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_part_x<float>(const Diags<float> &, float *,
                                         const Lines &, int, int, hipStream_t);
template hipError_t launch_part_x<double>(const Diags<double> &, double *,
                                          const Lines &, int, int,
                                          hipStream_t);
template hipError_t launch_part_yz<float>(const Diags<float> &, float *,
                                          const Lines &, int, int,
                                          hipStream_t);
template hipError_t launch_part_yz<double>(const Diags<double> &, double *,
                                           const Lines &, int, int,
                                           hipStream_t);

} // namespace pentad
