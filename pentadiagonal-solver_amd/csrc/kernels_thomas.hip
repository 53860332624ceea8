// kernels_thomas.hip -- one lane per system, global-memory intermediates.
//
// The general-purpose gfx950 kernel: any system length >= 3, any direction.
// Each lane runs the pentadiagonal Thomas recurrence of
// src/cpu/singlenode/pentad_orig_cpu.cpp:73-122 over one line; the forward
// intermediates du2, dw2, x2 go to the caller's scratch buffer laid out
// [3][len][n_sys] so that, at a fixed row, the 64 lanes of a wavefront touch
// 64 consecutive scratch words.  For the strided (y/z) directions the
// coefficient loads are coalesced too (adjacent lanes = adjacent systems);
// for x they are not (the reference's "naive" kernel class,
// src/cuda/singlenode/pentad_cuda.cu:45-117).
//
// HBM traffic: 7 compulsory words + 6 scratch words per row, so this kernel
// is bounded at 7/13 of the bandwidth roofline.  It is the fallback for
// shapes the register-resident kernels do not cover.
#include "pentad_device.hpp"
#include "pentad_internal.hpp"

namespace pentad {

template <typename T>
__global__ void __launch_bounds__(256)
    k_thomas_lane(const T *__restrict__ ds, const T *__restrict__ dl,
                  const T *__restrict__ d, const T *__restrict__ du,
                  const T *__restrict__ dw, T *__restrict__ x, Lines ln,
                  T *__restrict__ scratch) {
  const int64_t sys = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= ln.n_sys) return;
  const int64_t n = ln.len, st = ln.stride, ns = ln.n_sys;
  const int64_t b = ln.base(sys);
  T *du2 = scratch + sys;
  T *dw2 = du2 + n * ns;
  T *x2 = dw2 + n * ns;

  // row 0
  T inv = recip(d[b]);
  T u1 = du[b] * inv, w1 = dw[b] * inv, z1 = x[b] * inv;
  du2[0] = u1;
  dw2[0] = w1;
  x2[0] = z1;
  // row 1
  T ddl = dl[b + st];
  inv = recip(d[b + st] - ddl * u1);
  T u0 = (du[b + st] - ddl * w1) * inv;
  T w0 = dw[b + st] * inv;
  T z0 = (x[b + st] - ddl * z1) * inv;
  du2[ns] = u0;
  dw2[ns] = w0;
  x2[ns] = z0;
  // rows 2..n-1: (u1,w1,z1) hold row i-2, (u0,w0,z0) row i-1
  for (int64_t i = 2; i < n; ++i) {
    const int64_t e = b + i * st;
    const T s = ds[e];
    ddl = dl[e] - s * u1;
    inv = recip(d[e] - s * w1 - ddl * u0);
    const T un = (du[e] - ddl * w0) * inv;
    const T wn = dw[e] * inv;
    const T zn = (x[e] - s * z1 - ddl * z0) * inv;
    du2[i * ns] = un;
    dw2[i * ns] = wn;
    x2[i * ns] = zn;
    u1 = u0;
    w1 = w0;
    z1 = z0;
    u0 = un;
    w0 = wn;
    z0 = zn;
  }
  // backward substitution: xa = x[i+1], xb = x[i+2]
  T xb = z0;
  x[b + (n - 1) * st] = xb;
  T xa = z1 - u1 * xb;
  x[b + (n - 2) * st] = xa;
  for (int64_t i = n - 3; i >= 0; --i) {
    const T xi = x2[i * ns] - dw2[i * ns] * xb - du2[i * ns] * xa;
    x[b + i * st] = xi;
    xb = xa;
    xa = xi;
  }
}

template <typename T>
hipError_t launch_thomas_lane(const Diags<T> &a, T *x, const Lines &ln,
                              T *scratch, hipStream_t st) {
  const int block = 256;
  const int64_t grid = (ln.n_sys + block - 1) / block;
  hipLaunchKernelGGL(k_thomas_lane<T>, dim3((unsigned)grid), dim3(block), 0,
                     st, a.ds, a.dl, a.d, a.du, a.dw, x, ln, scratch);
  return hipGetLastError();
}

template hipError_t launch_thomas_lane<float>(const Diags<float> &, float *,
                                              const Lines &, float *,
                                              hipStream_t);
template hipError_t launch_thomas_lane<double>(const Diags<double> &, double *,
                                               const Lines &, double *,
                                               hipStream_t);

} // namespace pentad
