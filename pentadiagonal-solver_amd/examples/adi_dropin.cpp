// adi_dropin.cpp -- a caller written against the reference's C++ interface.
//
// It uses only what the reference's apps use (include/pentadsolver.hpp:36-131
// of SidhantKrAgrawal/pentadiagonal-solver): pentadsolver_create, the
// templated-by-overload pentadsolver_gpsv_batch_buffer_size_ext /
// pentadsolver_gpsv_batch, pentadsolver_kernel_that_ran and
// pentadsolver_destroy, in the pattern of the reference's ADI step
// (apps/adi_cuda.cu:167-176: solvedim 0, 1, 2 in place on one field with the
// same five coefficient arrays), with hipMalloc for cudaMalloc.
//
// usage: adi_dropin <in.bin> <out.bin> <n0> <n1> <n2> <iterations> <f64|f32>
//   in.bin : ds, dl, d, du, dw, x (n0*n1*n2 values each, the chosen type)
//   out.bin: x after the iterations
#include "pentadsolver.hpp"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#define CHECK(call)                                                                   \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));          \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

template <typename F>
static int run(const char *in, const char *out, const int dims[3], int iters) {
  const size_t n = (size_t)dims[0] * dims[1] * dims[2];
  std::vector<F> host(6 * n);
  FILE *f = std::fopen(in, "rb");
  if (!f || std::fread(host.data(), sizeof(F), 6 * n, f) != 6 * n) {
    std::fprintf(stderr, "cannot read %s\n", in);
    return 1;
  }
  std::fclose(f);
  F *dev[6];
  for (int k = 0; k < 6; ++k) {
    CHECK(hipMalloc(&dev[k], n * sizeof(F)));
    CHECK(hipMemcpy(dev[k], host.data() + k * n, n * sizeof(F), hipMemcpyHostToDevice));
  }
  pentadsolver_handle_t handle;
  pentadsolver_create(&handle, nullptr, 3, nullptr);
  void *buffer[3] = {nullptr, nullptr, nullptr};
  for (int sd = 0; sd < 3; ++sd) {
    const size_t bytes = pentadsolver_gpsv_batch_buffer_size_ext(
        handle, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dims, 3, sd);
    if (bytes) CHECK(hipMalloc(&buffer[sd], bytes));
  }
  for (int it = 0; it < iters; ++it)
    for (int sd = 0; sd < 3; ++sd)
      pentadsolver_gpsv_batch(handle, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dims, 3,
                              sd, buffer[sd]);
  CHECK(hipDeviceSynchronize());
  std::printf("kernels: %s %s %s\n", pentadsolver_kernel_that_ran(0),
              pentadsolver_kernel_that_ran(1), pentadsolver_kernel_that_ran(2));
  CHECK(hipMemcpy(host.data(), dev[5], n * sizeof(F), hipMemcpyDeviceToHost));
  f = std::fopen(out, "wb");
  if (!f || std::fwrite(host.data(), sizeof(F), n, f) != n) {
    std::fprintf(stderr, "cannot write %s\n", out);
    return 1;
  }
  std::fclose(f);
  for (int sd = 0; sd < 3; ++sd)
    if (buffer[sd]) CHECK(hipFree(buffer[sd]));
  for (int k = 0; k < 6; ++k) CHECK(hipFree(dev[k]));
  pentadsolver_destroy(&handle);
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s in.bin out.bin n0 n1 n2 iterations f64|f32\n", argv[0]);
    return 2;
  }
  const int dims[3] = {std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5])};
  const int iters = std::atoi(argv[6]);
  if (std::strcmp(argv[7], "f64") == 0) return run<double>(argv[1], argv[2], dims, iters);
  return run<float>(argv[1], argv[2], dims, iters);
}
