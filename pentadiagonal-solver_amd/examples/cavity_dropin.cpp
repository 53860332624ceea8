// cavity_dropin.cpp -- the pressure-correction step of the reference's 2-D
// cavity consumer (apps/2D_cavity_cpu.cpp:195-219), written as that consumer
// would call this library from C++: the assembly into padded arrays
// (pentadsolver_D_cavity_pressure_coef, in place of pressure_coef_matrix,
// apps/preproc.hpp:248), then the x solve and the y solve of the padded
// field through pentadsolver_D_gpsv_batch_padded with the reference's
// padding (nx_pad = nx rounded up to 4, 2D_cavity_cpu.cpp:70), and the same
// step as one fused call (pentadsolver_D_cavity_pressure_solve).
//
// usage: cavity_dropin <in.bin> <out.bin> <nx> <ny> <nx_pad>
//   in.bin : u_prime (ny+2)x(nx+1), v_prime (ny+1)x(nx+2), uv (ny+2)x(nx+2),
//            row-major doubles
//   out.bin: the field after the assembled-array step, then after the fused
//            step (ny * nx_pad doubles each, padding included)
#include "pentadsolver.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#define CHECK(call)                                                                   \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess) {                                                           \
      std::fprintf(stderr, "%s failed: %s\n", #call, hipGetErrorString(e_));          \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)
#define STATUS(call)                                                                  \
  do {                                                                                \
    pentadStatus_t s_ = (call);                                                       \
    if (s_ != PENTAD_STATUS_SUCCESS) {                                                \
      std::fprintf(stderr, "%s returned status %d\n", #call, (int)s_);               \
      std::exit(1);                                                                   \
    }                                                                                 \
  } while (0)

int main(int argc, char **argv) {
  if (argc != 6) {
    std::fprintf(stderr, "usage: %s in.bin out.bin nx ny nx_pad\n", argv[0]);
    return 2;
  }
  const int nx = std::atoi(argv[3]), ny = std::atoi(argv[4]), nx_pad = std::atoi(argv[5]);
  const size_t nu = (size_t)(ny + 2) * (nx + 1), nv = (size_t)(ny + 1) * (nx + 2),
               nuv = (size_t)(ny + 2) * (nx + 2), n = (size_t)nx_pad * ny;
  std::vector<double> host(nu + nv + nuv);
  FILE *f = std::fopen(argv[1], "rb");
  if (!f || std::fread(host.data(), sizeof(double), host.size(), f) != host.size()) {
    std::fprintf(stderr, "cannot read %s\n", argv[1]);
    return 1;
  }
  std::fclose(f);
  const double alpha_u = 0.7, dx = 1.0 / nx, dy = 1.0 / ny; // 2D_cavity_cpu.cpp:119-123

  double *in = nullptr, *arr = nullptr;
  CHECK(hipMalloc(&in, host.size() * sizeof(double)));
  CHECK(hipMalloc(&arr, 12 * n * sizeof(double)));
  CHECK(hipMemcpy(in, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice));
  CHECK(hipMemset(arr, 0, 12 * n * sizeof(double)));
  const double *up = in, *vp = in + nu, *uv = in + nu + nv;
  double *x = arr, *c[10];
  for (int k = 0; k < 10; ++k) c[k] = arr + (size_t)(k + 1) * n;
  double *xf = arr + 11 * n;

  pentadsolver_handle_t h = nullptr;
  pentadsolver_create(&h, nullptr, 2, nullptr);
  // the reference's sequence: assembly, x solve, y solve on the padded field
  STATUS(pentadsolver_D_cavity_pressure_coef(h, nx, ny, nx_pad, alpha_u, dx, dy, up, vp, uv, x,
                                             c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7],
                                             c[8], c[9]));
  const int dims[2] = {nx, ny}, pads[2] = {nx_pad, ny};
  void *buf = nullptr;
  size_t need = 0;
  for (int sd = 0; sd < 2; ++sd) {
    const size_t b = pentadsolver_D_gpsv_batch_padded_buffer_size(
        h, c[5 * sd], c[5 * sd + 1], c[5 * sd + 2], c[5 * sd + 3], c[5 * sd + 4], x, dims, pads,
        2, sd);
    need = b > need ? b : need;
  }
  if (need) CHECK(hipMalloc(&buf, need));
  for (int sd = 0; sd < 2; ++sd)
    STATUS(pentadsolver_D_gpsv_batch_padded(h, c[5 * sd], c[5 * sd + 1], c[5 * sd + 2],
                                            c[5 * sd + 3], c[5 * sd + 4], x, dims, pads, 2, sd,
                                            buf));
  const char *kx = pentadsolver_kernel_that_ran(0), *ky = pentadsolver_kernel_that_ran(2);
  std::printf("assembled: x %s, y %s\n", kx, ky);
  // the same step in one call
  STATUS(pentadsolver_D_cavity_pressure_solve(h, nx, ny, nx_pad, alpha_u, dx, dy, up, vp, uv,
                                              xf));
  CHECK(hipDeviceSynchronize());
  std::printf("fused: x %s, y %s\n", pentadsolver_kernel_that_ran(0),
              pentadsolver_kernel_that_ran(2));

  std::vector<double> out(2 * n);
  CHECK(hipMemcpy(out.data(), x, n * sizeof(double), hipMemcpyDeviceToHost));
  CHECK(hipMemcpy(out.data() + n, xf, n * sizeof(double), hipMemcpyDeviceToHost));
  f = std::fopen(argv[2], "wb");
  if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size()) {
    std::fprintf(stderr, "cannot write %s\n", argv[2]);
    return 1;
  }
  std::fclose(f);
  if (buf) CHECK(hipFree(buf));
  CHECK(hipFree(arr));
  CHECK(hipFree(in));
  pentadsolver_destroy(&h);
  return 0;
}
