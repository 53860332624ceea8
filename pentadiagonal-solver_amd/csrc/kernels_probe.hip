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

} // namespace pentad
