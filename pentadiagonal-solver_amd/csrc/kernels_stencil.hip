// kernels_stencil.hip -- batched solves whose five diagonals follow from one
// scale per line (structured coefficients).
//
// Along every line the system is  id * I + s * [w0, w1, w2, w3, w4]:
//   ds = s w0, dl = s w1, d = id + s w2, du = s w3, dw = s w4,
// with the couplings that reach past either end of the line dropped.  The
// kernels generate the diagonals in registers from s, so the HBM traffic per
// row is the right-hand side in and the solution out (16 B FP64, 8 B FP32)
// plus one scale per line, against 56 / 28 B when five diagonal arrays are
// read.  Implicit fourth-order diffusion I + nu dt/h^4 [1, -4, 6, -4, 1] is
// w = (1, -4, 6, -4, 1), id = 1; second-order diffusion is w = (0, -1, 2,
// -1, 0).  The solve is the partitioned one of partition_core.hpp (with its
// truncated block PCR), so the result is the one the general kernels give
// for the same assembled diagonals.
//
// k_stencil_x   contiguous lines through XMap addressing (the ADI driver's
//               x-like directions and the drop-in x direction), rotated
//               store as in k_part_x
// k_stencil_yz  strided lines (drop-in y / z): S systems x P chunks per
//               workgroup, exchange through LDS
//
// No reference counterpart: the reference assembles and reads five
// diagonal arrays for every solve (as apps/preproc.hpp:248 builds them).
#include "ir_core.hpp"
#include "partition_core.hpp"
#include "partition_exchange.hpp"
#include "rows_io.hpp"
#include "rotated_store.hpp"
#include "pentad_internal.hpp"

namespace pentad {

// Coefficients of one chunk of a structured line for the refinement solve:
// the FP32 operator (the FP64 diagonals rounded) for the bulk solve, the
// FP64 operator (generated from the FP64 scale) for the residual, so the
// refinement converges to the FP64 operator's solution.  Chunks are whole
// (len = M x lanes), so only the first chunk's lower couplings of rows 0, 1
// and the last chunk's upper couplings of rows M-2, M-1 reach past an end
// (zero); every other row has all five values.  Five values per line
// instead of five per row: no coefficient array occupies registers.
template <int M> struct StencilIrCoef {
  float a, b, c, e, f;
  double a64, b64, c64, e64, f64;
  bool first, last;
  template <typename T>
  __device__ __forceinline__ T lo(int i, int from, T v) const {
    return (i >= from || !first) ? v : T(0);
  }
  template <typename T>
  __device__ __forceinline__ T hi(int i, int upto, T v) const {
    return (i + upto < M || !last) ? v : T(0);
  }
  __device__ __forceinline__ float s(int i) const { return lo(i, 2, a); }
  __device__ __forceinline__ float l(int i) const { return lo(i, 1, b); }
  __device__ __forceinline__ float d(int) const { return c; }
  __device__ __forceinline__ float u(int i) const { return hi(i, 1, e); }
  __device__ __forceinline__ float w(int i) const { return hi(i, 2, f); }
  __device__ __forceinline__ double s64(int i) const { return lo(i, 2, a64); }
  __device__ __forceinline__ double l64(int i) const { return lo(i, 1, b64); }
  __device__ __forceinline__ double d64(int) const { return c64; }
  __device__ __forceinline__ double u64(int i) const { return hi(i, 1, e64); }
  __device__ __forceinline__ double w64(int i) const { return hi(i, 2, f64); }
  __device__ __forceinline__ void refresh() const {}
};

template <int M>
__device__ __forceinline__ StencilIrCoef<M> stencil_ir_coef(const Stencil<double> &st,
                                                           double sc, bool first, bool last) {
  StencilIrCoef<M> A;
  A.a64 = sc * st.w[0];
  A.b64 = sc * st.w[1];
  A.c64 = st.id + sc * st.w[2];
  A.e64 = sc * st.w[3];
  A.f64 = sc * st.w[4];
  A.a = (float)A.a64;
  A.b = (float)A.b64;
  A.c = (float)A.c64;
  A.e = (float)A.e64;
  A.f = (float)A.f64;
  A.first = first;
  A.last = last;
  return A;
}

// rows r0 .. r0+M-1 of a line of length len with scale s; rows past the end
// (padded chunks) are identity rows
template <typename T, int M>
__device__ __forceinline__ void stencil_rows(const Stencil<T> &st, T s, int64_t r0,
                                             int64_t len, T (&ds)[M], T (&dl)[M],
                                             T (&d)[M], T (&du)[M], T (&dw)[M]) {
  const T a = s * st.w[0], b = s * st.w[1], c = st.id + s * st.w[2];
  const T e = s * st.w[3], f = s * st.w[4];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int64_t r = r0 + i;
    const bool in = r < len;
    ds[i] = (in && r >= 2) ? a : T(0);
    dl[i] = (in && r >= 1) ? b : T(0);
    d[i] = in ? c : T(1);
    du[i] = (r + 1 < len) ? e : T(0);
    dw[i] = (r + 2 < len) ? f : T(0);
  }
}

// x-like lines: as k_part_x (ROT, Q passes into one tile), the five diagonal
// loads replaced by the stencil of the line's scale, scale index
// lo * sa + hi * sb of the system's XMap split.
template <typename T, int M, bool ROT, int Q>
__global__ void __launch_bounds__(512)
    k_stencil_x(Stencil<T> st, const T *b, T *out, int64_t n_sys, int64_t len,
                int lshift, XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t blk = ROT ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  const int L = 1 << lshift;
  const int s0shift = __builtin_ctz(blockDim.x) - lshift; // systems per pass
  const int sl = (int)threadIdx.x >> lshift;
  const int p = (int)threadIdx.x & (L - 1);
  WaveEx ex{(int)(threadIdx.x & 63), p, L};
  // pass q's right-hand side and scale; the next pass's are loaded before
  // pass q is solved (software pipelining across passes)
  auto load = [&](int q, T (&f)[M], T &scale, int64_t &off, bool &valid) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    valid = sys < n_sys;
    const int64_t sc = valid ? sys : n_sys - 1;
    const int64_t slo = map.lo(sc), shi = map.hi(sc);
    off = slo * map.in_lo + shi * map.in_hi + (int64_t)p * M;
    scale = st.scale[slo * st.sa + shi * st.sb];
    RowsIO<T, M>::load(b + off, f);
  };
  T f[M], scale;
  int64_t off;
  bool valid;
  load(0, f, scale, off, valid);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    T fn[M], scale_n = T(0);
    int64_t off_n = 0;
    bool valid_n = false;
    if (q + 1 < Q) load(q + 1, fn, scale_n, off_n, valid_n);
    T s[M], l[M], dd[M], u[M], w[M];
    stencil_rows<T, M>(st, scale, (int64_t)p * M, len, s, l, dd, u, w);
    part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
    if constexpr (ROT) {
      tile_put<T, M>(reinterpret_cast<T *>(smem_raw), f, p, Q << s0shift, (q << s0shift) + sl);
    } else {
      if (valid) RowsIO<T, M>::store(out + off, f);
    }
    if (q + 1 < Q) {
#pragma unroll
      for (int i = 0; i < M; ++i) f[i] = fn[i];
      scale = scale_n;
      off = off_n;
      valid = valid_n;
    }
  }
  if constexpr (ROT) {
    __syncthreads();
    const int sshift = s0shift + __builtin_ctz(Q);
    tile_flush<T, M>(reinterpret_cast<const T *>(smem_raw), sshift, (int)len, out, n_sys,
                     blk << sshift, map);
  }
}

// strided lines (drop-in y / z): scale index = system index of the Lines
// decomposition (the remaining axes in the standard order)
template <typename T, int M, int MAXT>
__global__ void __launch_bounds__(MAXT)
    k_stencil_yz(Stencil<T> st, T *__restrict__ x, Lines ln, int S, int P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int t = threadIdx.x;
  const int sl = t % S;
  const int p = t / S;
  const int64_t sys = xcd_tile(blockIdx.x, gridDim.x) * S + sl;
  const bool valid = sys < ln.n_sys;
  const int64_t sc = valid ? sys : ln.n_sys - 1;
  const int64_t stride = ln.stride;
  const int64_t b0 = ln.base(sc) + (int64_t)p * M * stride;
  const int len = (int)ln.len;
  T s[M], l[M], dd[M], u[M], w[M], f[M];
#pragma unroll
  for (int i = 0; i < M; ++i) f[i] = (p * M + i < len) ? x[b0 + i * stride] : T(0);
  stencil_rows<T, M>(st, st.scale[sc], (int64_t)p * M, len, s, l, dd, u, w);
  LdsEx ex{smem_raw, sl, p, S, P};
  part_solve<T, M>(s, l, dd, u, w, f, p, P, ex);
  if (valid) {
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (p * M + i < len) x[b0 + i * stride] = f[i];
  }
}

// FP32 solve + FP64 refinement over structured coefficients, x-like lines
// with the rotated store (the ADI driver): Q passes of BLOCK >> R systems
// into one tile whose rows are whole 128-byte lines, as k_ir_xq; HBM traffic
// per row: 8 B in, 8 B out (plus one scale per line).
template <int M, int R, int Q, int BLOCK>
__global__ void __launch_bounds__(BLOCK, 512 / BLOCK)
    k_stencil_ir_x(Stencil<double> st, const double *b, double *out, int64_t n_sys,
                   int64_t len, int max_steps, double tol, int *__restrict__ steps_out,
                   XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int L = 1 << R;
  constexpr int s0shift = (BLOCK == 512 ? 9 : 8) - R; // systems per pass
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int sl = (int)threadIdx.x >> R;
  const int p = (int)(threadIdx.x & (L - 1));
  const int lane = threadIdx.x & 63;
  WaveEx ex{lane, p, L};
  int kmax = 0;
#pragma unroll 1
  for (int q = 0; q < Q; ++q) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    const int64_t sc = sys < n_sys ? sys : n_sys - 1;
    const int64_t slo = map.lo(sc), shi = map.hi(sc);
    const int64_t off = slo * map.in_lo + shi * map.in_hi + (int64_t)p * M;
    const StencilIrCoef<M> A =
        stencil_ir_coef<M>(st, st.scale[slo * st.sa + shi * st.sb], p == 0, p == L - 1);
    const MemRhs bv{b + off, 1};
    double xs[M];
    const int k = ir_solve<M, ir_stored_rounds(R)>(A, bv, xs, p, L, ex, max_steps, tol);
    kmax = k > kmax ? k : kmax;
    tile_put<double, M>(reinterpret_cast<double *>(smem_raw), xs, p, Q << s0shift,
                        (q << s0shift) + sl);
  }
  __syncthreads();
  constexpr int sshift = s0shift + (Q == 4 ? 2 : (Q == 2 ? 1 : 0));
  tile_flush<double, M>(reinterpret_cast<const double *>(smem_raw), sshift, (int)len, out,
                        n_sys, blk << sshift, map);
  // largest step count over the launch (read first: the atomic is rarely needed)
  if (lane == 0 && steps_out != nullptr) {
    const int cur = __hip_atomic_load(steps_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (kmax > cur) atomicMax(steps_out, kmax);
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {

template <int M, int R>
hipError_t go_stencil_ir(const Stencil<double> &stc, const double *b, double *out,
                         int64_t n_sys, int64_t len, const XMap &map, int max_steps, double tol,
                         int *steps, hipStream_t st) {
  // 10-row chunks: a 16-system tile of 640-row lines is 82 KB, one
  // 512-thread workgroup per CU (as k_ir_xq)
  constexpr int B = M == 10 ? 512 : 256;
  constexpr int run = (B >> R) * 8;
  constexpr int Q = run >= 128 ? 1 : (run >= 64 ? 2 : 4);
  const int64_t per_wg = (int64_t)(B >> R) * Q;
  const int64_t grid = (n_sys + per_wg - 1) / per_wg;
  const size_t smem = rotated_tile_bytes<double>(B * Q, M, len);
  const hipError_t e = hipFuncSetAttribute(
      reinterpret_cast<const void *>(&k_stencil_ir_x<M, R, Q, B>),
      hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_stencil_ir_x<M, R, Q, B>), dim3((unsigned)grid), dim3(B), smem, st, stc,
                     b, out, n_sys, len, max_steps, tol, steps, map);
  return hipGetLastError();
}

template <int M>
hipError_t go_stencil_ir_r(const Stencil<double> &stc, const double *b, double *out,
                           int64_t n_sys, int64_t len, const XMap &map, int lanes,
                           int max_steps, double tol, int *steps, hipStream_t st) {
  switch (lanes) {
  case 2: return go_stencil_ir<M, 1>(stc, b, out, n_sys, len, map, max_steps, tol, steps, st);
  case 4: return go_stencil_ir<M, 2>(stc, b, out, n_sys, len, map, max_steps, tol, steps, st);
  case 8: return go_stencil_ir<M, 3>(stc, b, out, n_sys, len, map, max_steps, tol, steps, st);
  case 16: return go_stencil_ir<M, 4>(stc, b, out, n_sys, len, map, max_steps, tol, steps, st);
  case 32: return go_stencil_ir<M, 5>(stc, b, out, n_sys, len, map, max_steps, tol, steps, st);
  case 64: return go_stencil_ir<M, 6>(stc, b, out, n_sys, len, map, max_steps, tol, steps, st);
  }
  return hipErrorInvalidValue;
}

template <typename T, int M>
hipError_t go_sx(const Stencil<T> &stc, const T *b, T *out, bool rot, int64_t n_sys,
                 int64_t len, const XMap &map, int lanes, hipStream_t st) {
  int lshift = 0;
  while ((1 << lshift) < lanes) ++lshift;
  if (!rot) {
    const int block = 256;
    const int64_t grid = (n_sys * lanes + block - 1) / block;
    hipLaunchKernelGGL((k_stencil_x<T, M, false, 1>), dim3((unsigned)grid), dim3(block), 0,
                       st, stc, b, out, n_sys, len, lshift, map);
    return hipGetLastError();
  }
  // rotated stores: as launch_part_x_map, Q passes per workgroup make every
  // row of the tile a 128-byte run
  const int block = sizeof(T) == 4 ? 512 : 256;
  const int run = (block >> lshift) * (int)sizeof(T);
  const int Q = run >= 128 ? 1 : (run >= 64 ? 2 : 4);
  const int64_t per_wg = (int64_t)(block >> lshift) * Q;
  const int64_t grid = (n_sys + per_wg - 1) / per_wg;
  const size_t smem = rotated_tile_bytes<T>(block * Q, M, len);
  auto go = [&](auto kernel) {
    if (smem > 65536) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)smem);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), smem, st, stc, b, out,
                       n_sys, len, lshift, map);
    return hipGetLastError();
  };
  if (Q == 1) return go(k_stencil_x<T, M, true, 1>);
  if (Q == 2) return go(k_stencil_x<T, M, true, 2>);
  return go(k_stencil_x<T, M, true, 4>);
}

// Largest workgroup per (type, rows): 1024 threads (4 waves per SIMD, at
// most 128 VGPRs) for chunks of up to 6 (FP64) / 12 (FP32) rows, which fit
// without spilling.
constexpr int syz_maxt_rt(int elem_bytes, int rows) {
  return (elem_bytes == 8) ? (rows <= 6 ? 1024 : 512) : (rows <= 12 ? 1024 : 512);
}

template <typename T, int M>
hipError_t go_syz(const Stencil<T> &stc, T *x, const Lines &ln, int S, hipStream_t st) {
  constexpr int MAXT = syz_maxt_rt((int)sizeof(T), M);
  const int P = (int)((ln.len + M - 1) / M);
  const int block = S * P;
  if (block > MAXT) return hipErrorInvalidConfiguration;
  const size_t smem = (size_t)10 * S * P * sizeof(T);
  const int64_t grid = (ln.n_sys + S - 1) / S;
  if (smem > 65536) {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_stencil_yz<T, M, MAXT>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_stencil_yz<T, M, MAXT>), dim3((unsigned)grid), dim3(block), smem, st,
                     stc, x, ln, S, P);
  return hipGetLastError();
}

} // namespace

bool stencil_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 12 && rows != 16) return false;
  if (len < 3 || sys_per_wg < 1) return false;
  const int64_t P = (len + rows - 1) / rows;
  return P * sys_per_wg <= syz_maxt_rt(elem_bytes, rows);
}

template <typename T>
hipError_t launch_stencil_x_map(const Stencil<T> &stc, const T *b, T *out, bool rot,
                                int64_t n_sys, int64_t len, const XMap &map, int rows,
                                int lanes, hipStream_t st) {
  if (!part_x_supported((int)sizeof(T), rows, lanes) || (int64_t)rows * lanes != len)
    return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_sx<T, 4>(stc, b, out, rot, n_sys, len, map, lanes, st);
  case 6: return go_sx<T, 6>(stc, b, out, rot, n_sys, len, map, lanes, st);
  case 8: return go_sx<T, 8>(stc, b, out, rot, n_sys, len, map, lanes, st);
  case 10: return go_sx<T, 10>(stc, b, out, rot, n_sys, len, map, lanes, st);
  case 12: return go_sx<T, 12>(stc, b, out, rot, n_sys, len, map, lanes, st);
  case 16: return go_sx<T, 16>(stc, b, out, rot, n_sys, len, map, lanes, st);
  }
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_stencil_yz(const Stencil<T> &stc, T *x, const Lines &ln, int rows,
                             int sys_per_wg, hipStream_t st) {
  if (!stencil_yz_supported((int)sizeof(T), rows, sys_per_wg, ln.len))
    return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_syz<T, 4>(stc, x, ln, sys_per_wg, st);
  case 6: return go_syz<T, 6>(stc, x, ln, sys_per_wg, st);
  case 8: return go_syz<T, 8>(stc, x, ln, sys_per_wg, st);
  case 12: return go_syz<T, 12>(stc, x, ln, sys_per_wg, st);
  case 16: return go_syz<T, 16>(stc, x, ln, sys_per_wg, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_stencil_ir_x_map(const Stencil<double> &stc, const double *b, double *out,
                                   int64_t n_sys, int64_t len, const XMap &map, int rows,
                                   int max_steps, double tol, int *steps, hipStream_t st) {
  if (!ir_x_supported(rows, len)) return hipErrorInvalidValue;
  const int lanes = (int)(len / rows);
  switch (rows) {
  case 4: return go_stencil_ir_r<4>(stc, b, out, n_sys, len, map, lanes, max_steps, tol, steps, st);
  case 6: return go_stencil_ir_r<6>(stc, b, out, n_sys, len, map, lanes, max_steps, tol, steps, st);
  case 8: return go_stencil_ir_r<8>(stc, b, out, n_sys, len, map, lanes, max_steps, tol, steps, st);
  case 10:
    return go_stencil_ir_r<10>(stc, b, out, n_sys, len, map, lanes, max_steps, tol, steps, st);
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_stencil_x_map<float>(const Stencil<float> &, const float *, float *,
                                                bool, int64_t, int64_t, const XMap &, int,
                                                int, hipStream_t);
template hipError_t launch_stencil_x_map<double>(const Stencil<double> &, const double *,
                                                 double *, bool, int64_t, int64_t,
                                                 const XMap &, int, int, hipStream_t);
template hipError_t launch_stencil_yz<float>(const Stencil<float> &, float *, const Lines &,
                                             int, int, hipStream_t);
template hipError_t launch_stencil_yz<double>(const Stencil<double> &, double *,
                                              const Lines &, int, int, hipStream_t);

} // namespace pentad
