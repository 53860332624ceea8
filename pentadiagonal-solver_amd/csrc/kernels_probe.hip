// kernels_probe.hip -- HBM bandwidth probe with the solve's access pattern.
//
// out[i] = f(in0[i], ..., in5[i]) with `fmas` dependent FMAs per element:
// six streams read and one written, the compulsory traffic of one batched
// solve (five diagonals + RHS in, solution out).  fmas = 0 measures the
// memory floor; growing fmas shows where arithmetic starts to cost, the
// method of the reference's apps/roofline_probe_cuda.cu:27-49.
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "partition_exchange.hpp"

namespace pentad {

template <typename T>
__global__ void __launch_bounds__(256)
    k_probe(const T *__restrict__ a, const T *__restrict__ b, const T *__restrict__ c,
            const T *__restrict__ d, const T *__restrict__ e, const T *__restrict__ f,
            T *__restrict__ out, size_t n, int fmas) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    T acc = a[i] + b[i] + c[i] + d[i] + e[i] + f[i];
    for (int k = 0; k < fmas; ++k) acc = acc * T(0.999) + T(1e-3);
    out[i] = acc;
  }
}

hipError_t launch_probe(const void *const in[6], void *out, size_t n,
                        int elem_bytes, int fmas, hipStream_t st) {
  const int block = 256;
  size_t grid = (n + block - 1) / block;
  if (grid > 256 * 64) grid = 256 * 64; // 64 workgroups per CU, grid-stride
  if (elem_bytes == 8) {
    hipLaunchKernelGGL(k_probe<double>, dim3((unsigned)grid), dim3(block), 0, st,
                       (const double *)in[0], (const double *)in[1],
                       (const double *)in[2], (const double *)in[3],
                       (const double *)in[4], (const double *)in[5],
                       (double *)out, n, fmas);
  } else {
    hipLaunchKernelGGL(k_probe<float>, dim3((unsigned)grid), dim3(block), 0, st,
                       (const float *)in[0], (const float *)in[1],
                       (const float *)in[2], (const float *)in[3],
                       (const float *)in[4], (const float *)in[5],
                       (float *)out, n, fmas);
  }
  return hipGetLastError();
}

// The byte mix of the FP32 + FP64 refinement solve: per row 20 B of FP32
// coefficients and 8 B of FP64 right-hand side read, 8 B of FP64 solution
// written, as 7 streams (five coefficient arrays, b, x), 3 streams (the
// coefficients interleaved per 4 rows: 5 x 16 B, then b and x) or 2 streams
// (coefficients and b interleaved per 4 rows: 5 x 16 B + 2 x 16 B, then x).
// One thread moves 4 rows with 16-byte accesses; n must be a multiple of 4.
__device__ __forceinline__ float sum5(float4 a, float4 b, float4 c, float4 d, float4 e, int k) {
  const float4 *v[5] = {&a, &b, &c, &d, &e};
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < 5; ++j) s += k == 0 ? v[j]->x : (k == 1 ? v[j]->y : (k == 2 ? v[j]->z : v[j]->w));
  return s;
}

template <int STREAMS>
__global__ void __launch_bounds__(256)
    k_probe_ir_mix(const float4 *__restrict__ c0, const float4 *__restrict__ c1,
                   const float4 *__restrict__ c2, const float4 *__restrict__ c3,
                   const float4 *__restrict__ c4, const double2 *__restrict__ b,
                   double2 *__restrict__ x, size_t n4) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n4) return;
  float4 a, bb, c, d, e;
  double2 f0, f1;
  if constexpr (STREAMS == 7) {
    a = c0[i]; bb = c1[i]; c = c2[i]; d = c3[i]; e = c4[i];
    f0 = b[2 * i]; f1 = b[2 * i + 1];
  } else if constexpr (STREAMS == 3) {
    const float4 *p = c0 + 5 * i;
    a = p[0]; bb = p[1]; c = p[2]; d = p[3]; e = p[4];
    f0 = b[2 * i]; f1 = b[2 * i + 1];
  } else {
    const float4 *p = c0 + 7 * i;
    a = p[0]; bb = p[1]; c = p[2]; d = p[3]; e = p[4];
    f0 = *reinterpret_cast<const double2 *>(p + 5);
    f1 = *reinterpret_cast<const double2 *>(p + 6);
  }
  x[2 * i] = double2{f0.x + sum5(a, bb, c, d, e, 0), f0.y + sum5(a, bb, c, d, e, 1)};
  x[2 * i + 1] = double2{f1.x + sum5(a, bb, c, d, e, 2), f1.y + sum5(a, bb, c, d, e, 3)};
}

hipError_t launch_probe_ir_mix(int streams, const float *const coef[5], const double *b,
                               double *x, size_t n, hipStream_t st) {
  if (n % 4) return hipErrorInvalidValue;
  const size_t n4 = n / 4;
  const int block = 256;
  const size_t grid = (n4 + block - 1) / block;
  const float4 *c[5];
  for (int k = 0; k < 5; ++k) c[k] = reinterpret_cast<const float4 *>(coef[k]);
  const double2 *b2 = reinterpret_cast<const double2 *>(b);
  double2 *x2 = reinterpret_cast<double2 *>(x);
  if (streams == 7)
    hipLaunchKernelGGL(k_probe_ir_mix<7>, dim3((unsigned)grid), dim3(block), 0, st, c[0], c[1],
                       c[2], c[3], c[4], b2, x2, n4);
  else if (streams == 3)
    hipLaunchKernelGGL(k_probe_ir_mix<3>, dim3((unsigned)grid), dim3(block), 0, st, c[0], c[0],
                       c[0], c[0], c[0], b2, x2, n4);
  else if (streams == 2)
    hipLaunchKernelGGL(k_probe_ir_mix<2>, dim3((unsigned)grid), dim3(block), 0, st, c[0], c[0],
                       c[0], c[0], c[0], b2, x2, n4);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// The refinement kernel's I/O without its solve, in two layouts.  Rows come
// in 8-row chunks, one per thread: 10 float4 of coefficients interleaved per
// 4 rows (lane stride 160 B, the driver's layout) and 4 double2 of b (lane
// stride 64 B), x = b + some of the coefficients, then `rounds` rounds of
// 8 dependent-free FP64 FMAs on the chunk's values stand in for the solve's
// arithmetic (0 = none).
//   LANES: one chunk per thread, every thread of a 256-thread workgroup on
//          its own chunk; x written over the wavefront's 512 rows with
//          coalesced 16-byte stores (1 KB per instruction, the values'
//          order within those rows is immaterial here); as many workgroups
//          as the rows need.
//   TILE:  the headline kernel's workgroup structure for 512-row lines:
//          256 threads, two workgroups per CU (66 KB of LDS each), four
//          passes of four systems (one per wavefront) into one 16-system
//          tile, written rotated after a barrier (a 128-byte run per tile
//          row, rows 4 KB apart, non-temporal stores), the x direction of a
//          512^3 field (system s: lo = s % 512, hi = s / 512; row r of s at
//          lo + 512 r + 512^2 hi).
template <bool TILE>
__global__ void __launch_bounds__(256, 2)
    k_probe_ir_layout(const float4 *__restrict__ c, const double2 *__restrict__ b,
                      double *__restrict__ x, int rounds) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *tile = reinterpret_cast<double *>(smem_raw);
  const int w = (int)threadIdx.x >> 6, l = (int)threadIdx.x & 63;
  const int passes = TILE ? 4 : 1;
  for (int q = 0; q < passes; ++q) {
    const size_t chunk = TILE ? ((size_t)blockIdx.x * 16 + q * 4 + w) * 64 + l
                              : (size_t)blockIdx.x * 256 + threadIdx.x;
    float4 a[10];
    double2 f[4];
#pragma unroll
    for (int k = 0; k < 10; ++k) a[k] = c[chunk * 10 + k];
#pragma unroll
    for (int j = 0; j < 4; ++j) f[j] = b[chunk * 4 + j];
    double v[8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      v[2 * j] = f[j].x + (double)(a[j].x + a[j + 5].y + a[(j + 2) % 10].z);
      v[2 * j + 1] = f[j].y + (double)(a[j + 1].w + a[j + 4].x + a[(j + 7) % 10].y);
    }
    for (int it = 0; it < rounds; ++it) {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = fma(v[i], 0.9990000000001, 1e-3);
    }
    if constexpr (TILE) {
      const int col = q * 4 + w;
#pragma unroll
      for (int i = 0; i < 8; ++i) tile[(l * 8 + i) * 16 + col + l] = v[i];
    } else {
      // the wavefront's 512 rows as four 1 KB stores, lanes side by side
      // (streaming stores, as the solve kernels write whole lines)
      double *o = x + ((chunk - l) * 4 + l) * 2;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        __builtin_nontemporal_store(v[2 * j], o + j * 128);
        __builtin_nontemporal_store(v[2 * j + 1], o + j * 128 + 1);
      }
    }
  }
  if constexpr (TILE) {
    __syncthreads();
    const int s = (int)threadIdx.x & 15;
    const size_t sys = (size_t)blockIdx.x * 16 + s;
    double *o = x + (sys & 511) + (sys >> 9) * (512 * 512);
    for (int row = (int)threadIdx.x >> 4; row < 512; row += 16)
      __builtin_nontemporal_store(tile[row * 16 + s + row / 8], o + (size_t)row * 512);
  }
}

hipError_t launch_probe_ir_layout(int layout, int rounds, const float *coef, const double *b,
                                  double *x, size_t n, hipStream_t st) {
  const float4 *c = reinterpret_cast<const float4 *>(coef);
  const double2 *b2 = reinterpret_cast<const double2 *>(b);
  if (layout == 0) {
    if (n % 2048) return hipErrorInvalidValue; // whole workgroups of 8-row chunks
    hipLaunchKernelGGL(k_probe_ir_layout<false>, dim3((unsigned)(n / 2048)), dim3(256), 0, st,
                       c, b2, x, rounds);
  } else {
    if (n != (size_t)512 * 512 * 512) return hipErrorInvalidValue;
    const size_t smem = sizeof(double) * (16 * 512 + 64);
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_probe_ir_layout<true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_probe_ir_layout<true>, dim3((unsigned)(n / (16 * 512))), dim3(256),
                       smem, st, c, b2, x, rounds);
  }
  return hipGetLastError();
}

// The strided directions' I/O without the solve, lines of 512 FP64 rows:
// a 512-thread workgroup holds S adjacent systems x P chunks of 16 rows,
// loads every row of the five diagonals and x into registers, waits at a
// barrier for the whole workgroup's loads (the solve exchanges through LDS
// before it stores anything), then writes x += ds + dl + d + du + dw in
// place.  S = 16 (K = 1) is the FP64 drop-in kernel's structure, whole lines
// per workgroup; S = 32 / 64 gives each workgroup half / a quarter of every
// line (K = 2 / 4 workgroups per line set), so that one visit of a 2 MB page
// of a z line (stride 512^2 elements) serves 256 / 512 B per array instead
// of 128.  System sys starts at (sys / lo) * lo * 512 + sys % lo with row
// stride lo (lo = the extent below the solve axis: nx for y, nx * ny for z).
// Tiles are dealt to the XCDs as the solve kernel deals them (xcd_tile), the
// K workgroups of one line set and the neighbouring line sets on one XCD.
template <int S>
__global__ void __launch_bounds__(512, 1)
    k_probe_strided_io(const double *__restrict__ a0, const double *__restrict__ a1,
                       const double *__restrict__ a2, const double *__restrict__ a3,
                       const double *__restrict__ a4, double *__restrict__ x, int64_t lo) {
  constexpr int M = 16, P = 512 / S, K = S / 16;
  const int t = (int)threadIdx.x, s = t % S, c = t / S;
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t g = blk / K;
  const int k = (int)(blk % K);
  const int64_t sys = g * S + s;
  const int64_t base = (sys / lo) * lo * 512 + sys % lo + ((int64_t)k * P * M + (int64_t)c * M) * lo;
  const double *src[6] = {a0, a1, a2, a3, a4, x};
  double v[6][M];
#pragma unroll
  for (int q = 0; q < 6; ++q)
#pragma unroll
    for (int i = 0; i < M; ++i) v[q][i] = src[q][base + (int64_t)i * lo];
  __syncthreads();
#pragma unroll
  for (int i = 0; i < M; ++i)
    x[base + (int64_t)i * lo] = v[5][i] + v[0][i] + v[1][i] + v[2][i] + v[3][i] + v[4][i];
}

hipError_t launch_probe_strided_io(int split, const double *const a[5], double *x, int64_t lo,
                                   int64_t n_sys, hipStream_t st) {
  const int S = 16 * split;
  if ((split != 1 && split != 2 && split != 4) || n_sys % S != 0) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(n_sys / S * split)), block(512);
  if (split == 1)
    hipLaunchKernelGGL(k_probe_strided_io<16>, grid, block, 0, st, a[0], a[1], a[2], a[3], a[4], x, lo);
  else if (split == 2)
    hipLaunchKernelGGL(k_probe_strided_io<32>, grid, block, 0, st, a[0], a[1], a[2], a[3], a[4], x, lo);
  else
    hipLaunchKernelGGL(k_probe_strided_io<64>, grid, block, 0, st, a[0], a[1], a[2], a[3], a[4], x, lo);
  return hipGetLastError();
}

} // namespace pentad
