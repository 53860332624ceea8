// kernels_layout.hip -- layout transforms used by the ADI driver.
//
// rotate: view the input as a [rows][cols] matrix (cols fastest) and write
// its transpose [cols][rows], converting the element type on the way
// (double -> double, double -> float, float -> float, float -> double).
// With rows = number of lines and cols = line length this turns the
// [system][row] layout of one ADI direction into the [row][system] layout in
// which the next direction's lines are contiguous.  64x64 tiles staged
// through LDS (padded against bank conflicts); reads and writes are 256-byte
// (FP32) / 512-byte (FP64) coalesced runs.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pentad {

constexpr int TILE = 64;
constexpr int TROWS = 16; // threads per tile column: 64 x 16 threads

template <typename Ti, typename To>
__global__ void __launch_bounds__(TILE * TROWS)
    k_rotate(const Ti *__restrict__ in, To *__restrict__ out, int64_t rows,
             int64_t cols) {
  __shared__ To tile[TILE][TILE + 1];
  const int64_t r0 = (int64_t)blockIdx.x * TILE, c0 = (int64_t)blockIdx.y * TILE;
  const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
#pragma unroll
  for (int k = 0; k < TILE; k += TROWS) {
    const int64_t r = r0 + ty + k, c = c0 + tx;
    if (r < rows && c < cols) tile[ty + k][tx] = (To)in[r * cols + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < TILE; k += TROWS) {
    const int64_t c = c0 + ty + k, r = r0 + tx;
    if (r < rows && c < cols) out[c * rows + r] = tile[tx][ty + k];
  }
}

template <typename Ti, typename To>
__global__ void __launch_bounds__(256)
    k_convert(const Ti *__restrict__ in, To *__restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (To)in[i];
}

// in_bytes / out_bytes: 4 or 8.  rows * cols elements.
hipError_t launch_rotate(const void *in, int in_bytes, void *out, int out_bytes,
                         int64_t rows, int64_t cols, hipStream_t st) {
  const dim3 grid((unsigned)((rows + TILE - 1) / TILE), (unsigned)((cols + TILE - 1) / TILE));
  const dim3 block(TILE * TROWS);
  if (grid.y > 65535u) return hipErrorInvalidValue; // lines longer than 4M rows
  if (in_bytes == 8 && out_bytes == 8)
    hipLaunchKernelGGL((k_rotate<double, double>), grid, block, 0, st,
                       (const double *)in, (double *)out, rows, cols);
  else if (in_bytes == 8 && out_bytes == 4)
    hipLaunchKernelGGL((k_rotate<double, float>), grid, block, 0, st,
                       (const double *)in, (float *)out, rows, cols);
  else if (in_bytes == 4 && out_bytes == 4)
    hipLaunchKernelGGL((k_rotate<float, float>), grid, block, 0, st,
                       (const float *)in, (float *)out, rows, cols);
  else if (in_bytes == 4 && out_bytes == 8)
    hipLaunchKernelGGL((k_rotate<float, double>), grid, block, 0, st,
                       (const float *)in, (double *)out, rows, cols);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_convert(const void *in, int in_bytes, void *out, int out_bytes,
                          int64_t n, hipStream_t st) {
  const int block = 256;
  int64_t grid = (n + block - 1) / block;
  if (grid > 256 * 32) grid = 256 * 32;
  if (in_bytes == 8 && out_bytes == 4)
    hipLaunchKernelGGL((k_convert<double, float>), dim3((unsigned)grid), dim3(block), 0,
                       st, (const double *)in, (float *)out, n);
  else if (in_bytes == 4 && out_bytes == 8)
    hipLaunchKernelGGL((k_convert<float, double>), dim3((unsigned)grid), dim3(block), 0,
                       st, (const float *)in, (double *)out, n);
  else if (in_bytes == out_bytes)
    return hipMemcpyAsync(out, in, (size_t)n * in_bytes, hipMemcpyDeviceToDevice, st);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

// out[i*os0 + j*os1 + k*os2] = in[i + d0*(j + d1*k)]: any axis permutation
// of a 3-D array (reads coalesced, writes scattered; used once at set-up).
template <typename Ti, typename To>
__global__ void __launch_bounds__(256)
    k_permute3(const Ti *__restrict__ in, To *__restrict__ out, int64_t d0, int64_t d1,
               int64_t n, int64_t os0, int64_t os1, int64_t os2) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
    const int64_t i = e % d0, jk = e / d0;
    const int64_t j = jk % d1, k = jk / d1;
    out[i * os0 + j * os1 + k * os2] = (To)in[e];
  }
}

hipError_t launch_permute3(const void *in, int in_bytes, void *out, int out_bytes,
                           const int64_t dims[3], const int64_t ostride[3],
                           hipStream_t st) {
  const int64_t n = dims[0] * dims[1] * dims[2];
  const int block = 256;
  int64_t grid = (n + block - 1) / block;
  if (grid > 256 * 64) grid = 256 * 64;
#define PERM(TI, TO)                                                                   \
  hipLaunchKernelGGL((k_permute3<TI, TO>), dim3((unsigned)grid), dim3(block), 0, st,  \
                     (const TI *)in, (TO *)out, dims[0], dims[1], n, ostride[0],      \
                     ostride[1], ostride[2])
  if (in_bytes == 8 && out_bytes == 8) PERM(double, double);
  else if (in_bytes == 8 && out_bytes == 4) PERM(double, float);
  else if (in_bytes == 4 && out_bytes == 4) PERM(float, float);
  else if (in_bytes == 4 && out_bytes == 8) PERM(float, double);
  else return hipErrorInvalidValue;
#undef PERM
  return hipGetLastError();
}

// out[20 * g + 4 * k + r] = src[k][4 * g + r]: five arrays interleaved per
// group of 4 rows, one thread per group (set-up only).
__global__ void __launch_bounds__(256)
    k_interleave4(const float4 *__restrict__ s0, const float4 *__restrict__ s1,
                  const float4 *__restrict__ s2, const float4 *__restrict__ s3,
                  const float4 *__restrict__ s4, float4 *__restrict__ out, int64_t groups) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    float4 *o = out + 5 * g;
    o[0] = s0[g];
    o[1] = s1[g];
    o[2] = s2[g];
    o[3] = s3[g];
    o[4] = s4[g];
  }
}

// out[10 * g + 2 * k + r] = src[k][2 * g + r]: the same per group of 2 rows
// (8-byte pieces), for chunks of 6 or 10 rows.
__global__ void __launch_bounds__(256)
    k_interleave2(const float2 *__restrict__ s0, const float2 *__restrict__ s1,
                  const float2 *__restrict__ s2, const float2 *__restrict__ s3,
                  const float2 *__restrict__ s4, float2 *__restrict__ out, int64_t groups) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += stride) {
    float2 *o = out + 5 * g;
    o[0] = s0[g];
    o[1] = s1[g];
    o[2] = s2[g];
    o[3] = s3[g];
    o[4] = s4[g];
  }
}

hipError_t launch_interleave(const float *const src[5], float *out, int64_t n, int group,
                             hipStream_t st) {
  if ((group != 2 && group != 4) || n % group) return hipErrorInvalidValue;
  const uintptr_t align = (uintptr_t)group * sizeof(float) - 1;
  for (int k = 0; k < 5; ++k)
    if (reinterpret_cast<uintptr_t>(src[k]) & align) return hipErrorInvalidValue;
  if (reinterpret_cast<uintptr_t>(out) & align) return hipErrorInvalidValue;
  const int64_t groups = n / group;
  const int block = 256;
  int64_t grid = (groups + block - 1) / block;
  if (grid > 256 * 64) grid = 256 * 64;
  if (group == 4)
    hipLaunchKernelGGL(k_interleave4, dim3((unsigned)grid), dim3(block), 0, st,
                       reinterpret_cast<const float4 *>(src[0]),
                       reinterpret_cast<const float4 *>(src[1]),
                       reinterpret_cast<const float4 *>(src[2]),
                       reinterpret_cast<const float4 *>(src[3]),
                       reinterpret_cast<const float4 *>(src[4]), reinterpret_cast<float4 *>(out),
                       groups);
  else
    hipLaunchKernelGGL(k_interleave2, dim3((unsigned)grid), dim3(block), 0, st,
                       reinterpret_cast<const float2 *>(src[0]),
                       reinterpret_cast<const float2 *>(src[1]),
                       reinterpret_cast<const float2 *>(src[2]),
                       reinterpret_cast<const float2 *>(src[3]),
                       reinterpret_cast<const float2 *>(src[4]), reinterpret_cast<float2 *>(out),
                       groups);
  return hipGetLastError();
}

__global__ void k_fill_int(int *p, int v) { *p = v; }

// Writes one int with a kernel (usable inside captured graphs, where a
// memset node is avoided).
hipError_t launch_fill_int(int *p, int v, hipStream_t st) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(1), 0, st, p, v);
  return hipGetLastError();
}

} // namespace pentad
