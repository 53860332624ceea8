// microbench_dropin.cpp -- the reference's GPU micro-benchmark matrix
// (benchmarks/cuda/my_bm.cu:55-71: N = 32..max_n by powers of two; ndims 1,
// 2, 3 with every solve dimension; FP32 and FP64) run natively against this
// library, one pentadsolver_gpsv_batch call per iteration as the reference's
// BM_PentadSolver does (:6-44), through its C++ overloads.
//
// Per case it reports
//   launch_us  host time of one asynchronous call (what the reference's
//              Google Benchmark loop times: the call returns before the
//              solve ends);
//   gpu_us     steady-state GPU time per call: `reps` calls back to back on
//              the handle's stream between two HIP events;
//   resid      max-norm relative residual |A x - b| / |b| of the first
//              solve, computed on the host in FP64 (an independent check,
//              no oracle).
// Inputs: off-diagonals U(-1, 1) and d = U(2, 3) * sum|off| from a seeded
// 64-bit LCG, right-hand side U(-1, 1).
//
// usage: microbench_dropin [max_n] [reps]      (defaults 256, 50)
// One line per case, then nothing else on stdout.
#include "pentadsolver.hpp"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
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

namespace {

struct Lcg {
  unsigned long long s;
  double uniform(double lo, double hi) {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return lo + (hi - lo) * (double)(s >> 11) * (1.0 / 9007199254740992.0);
  }
};

// max |A x - b| / max |b| along solve dimension sd of a dense grid
template <typename F>
double residual(const std::vector<F> (&c)[5], const std::vector<F> &b, const std::vector<F> &x,
                const std::vector<int> &dims, int sd) {
  size_t stride = 1;
  for (int k = 0; k < sd; ++k) stride *= (size_t)dims[k];
  const size_t len = (size_t)dims[sd], n = b.size();
  double rmax = 0.0, bmax = 0.0;
  for (size_t e = 0; e < n; ++e) {
    const size_t r = (e / stride) % len;
    double ax = (double)c[2][e] * (double)x[e];
    if (r >= 2) ax += (double)c[0][e] * (double)x[e - 2 * stride];
    if (r >= 1) ax += (double)c[1][e] * (double)x[e - stride];
    if (r + 1 < len) ax += (double)c[3][e] * (double)x[e + stride];
    if (r + 2 < len) ax += (double)c[4][e] * (double)x[e + 2 * stride];
    rmax = std::fmax(rmax, std::fabs(ax - (double)b[e]));
    bmax = std::fmax(bmax, std::fabs((double)b[e]));
  }
  return rmax / bmax;
}

template <typename F>
void run_case(int n, int ndims, int sd, int reps, pentadsolver_handle_t h, hipStream_t st) {
  const std::vector<int> dims(ndims, n);
  size_t total = 1;
  for (int v : dims) total *= (size_t)v;
  Lcg g{0x9E3779B97F4A7C15ULL ^ (unsigned long long)(n * 131 + ndims * 17 + sd)};
  std::vector<F> c[5], b(total);
  for (auto &v : c) v.resize(total);
  for (size_t e = 0; e < total; ++e) {
    double off = 0.0;
    for (int k : {0, 1, 3, 4}) {
      c[k][e] = (F)g.uniform(-1.0, 1.0);
      off += std::fabs((double)c[k][e]);
    }
    c[2][e] = (F)(g.uniform(2.0, 3.0) * off);
    b[e] = (F)g.uniform(-1.0, 1.0);
  }
  F *dev[7];
  for (int k = 0; k < 7; ++k) CHECK(hipMalloc(&dev[k], total * sizeof(F)));
  for (int k = 0; k < 5; ++k)
    CHECK(hipMemcpy(dev[k], c[k].data(), total * sizeof(F), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dev[5], b.data(), total * sizeof(F), hipMemcpyHostToDevice));
  CHECK(hipMemcpy(dev[6], dev[5], total * sizeof(F), hipMemcpyDeviceToDevice));
  // scratch allocated outside the timed loop, as the reference does
  const size_t bytes = pentadsolver_gpsv_batch_buffer_size_ext(
      h, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dims.data(), ndims, sd);
  void *buf = nullptr;
  if (bytes) CHECK(hipMalloc(&buf, bytes));
  auto call = [&]() {
    pentadsolver_gpsv_batch(h, dev[0], dev[1], dev[2], dev[3], dev[4], dev[5], dims.data(),
                            (size_t)ndims, sd, buf);
  };
  // first solve: the residual check
  call();
  CHECK(hipStreamSynchronize(st));
  std::vector<F> x(total);
  CHECK(hipMemcpy(x.data(), dev[5], total * sizeof(F), hipMemcpyDeviceToHost));
  const double res = residual<F>(c, b, x, dims, sd);
  // host time of one asynchronous call, the right-hand side restored
  // outside the timed region (the reference's loop)
  double launch_sum = 0.0;
  for (int r = 0; r < reps; ++r) {
    CHECK(hipMemcpyAsync(dev[5], dev[6], total * sizeof(F), hipMemcpyDeviceToDevice, st));
    CHECK(hipStreamSynchronize(st));
    const auto t0 = std::chrono::steady_clock::now();
    call();
    const auto t1 = std::chrono::steady_clock::now();
    launch_sum += std::chrono::duration<double, std::micro>(t1 - t0).count();
    CHECK(hipStreamSynchronize(st));
  }
  // GPU time per call, calls back to back (each solves the previous
  // solution: the same work)
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  call();
  CHECK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r) call();
  CHECK(hipEventRecord(e1, st));
  CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  CHECK(hipEventElapsedTime(&ms, e0, e1));
  const double gpu_us = 1e3 * ms / reps;
  const double rows = (double)total;
  std::printf("%s n=%d ndims=%d solvedim=%d launch_us=%.2f gpu_us=%.2f grows=%.4f gbs=%.1f "
              "resid=%.2e kernel=%s config=\"%s\"\n",
              sizeof(F) == 8 ? "f64" : "f32", n, ndims, sd, launch_sum / reps, gpu_us,
              rows / (gpu_us * 1e-6) / 1e9, 7.0 * sizeof(F) * rows / (gpu_us * 1e-6) / 1e9, res,
              pentadsolver_kernel_that_ran(sd == 0 ? 0 : (sd == ndims - 1 ? 2 : 1)),
              pentadsolver_kernel_config(h, sd == 0 ? 0 : (sd == ndims - 1 ? 2 : 1)));
  std::fflush(stdout);
  CHECK(hipEventDestroy(e0));
  CHECK(hipEventDestroy(e1));
  if (buf) CHECK(hipFree(buf));
  for (int k = 0; k < 7; ++k) CHECK(hipFree(dev[k]));
}

} // namespace

int main(int argc, char **argv) {
  const int max_n = argc > 1 ? std::atoi(argv[1]) : 256;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 50;
  if (max_n < 32 || reps < 1) {
    std::fprintf(stderr, "usage: %s [max_n >= 32] [reps >= 1]\n", argv[0]);
    return 2;
  }
  hipStream_t st;
  CHECK(hipStreamCreate(&st));
  pentadsolver_handle_t h;
  pentadsolver_create(&h, nullptr, 0, nullptr);
  pentadsolver_set_stream(h, st);
  for (int prec = 0; prec < 2; ++prec)
    for (int ndims = 1; ndims <= 3; ++ndims)
      for (int sd = 0; sd < ndims; ++sd)
        for (int n = 32; n <= max_n; n *= 2) {
          if (prec == 0)
            run_case<float>(n, ndims, sd, reps, h, st);
          else
            run_case<double>(n, ndims, sd, reps, h, st);
        }
  pentadsolver_destroy(&h);
  CHECK(hipStreamDestroy(st));
  return 0;
}
