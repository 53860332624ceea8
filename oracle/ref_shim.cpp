// ref_shim.cpp -- C entry points over the reference CPU solver, built from the
// reference's own sources by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ to pin the oracle against the
// real reference, and by bench.py's cpu_baseline leg as the timed CPU path.
// The product library never links or loads it.
//
// The reference exposes C++ overloads pentadsolver_gpsv_batch(handle, ...)
// (include/pentadsolver.hpp:121-130).  Two reference quirks are handled here:
//  * the AVX build reads t_dims[2] even for 2-D input
//    (src/cpu/singlenode/pentad_cpu.cpp:613), so dims are copied into a
//    3-entry array padded with 1;
//  * the AVX x-path assumes the leading dimension is padded to the SIMD width
//    (pentad_cpu.cpp:611-613); callers must pass dims[0] % 8 == 0 for 2-D/3-D
//    x-solves, which the tests do.
#include "pentadsolver.hpp"

namespace {
void pad_dims(const int *dims, int ndim, int out[3]) {
  out[0] = out[1] = out[2] = 1;
  for (int i = 0; i < ndim && i < 3; ++i) out[i] = dims[i];
}
} // namespace

extern "C" int ref_gpsv_batch_d(const double *ds, const double *dl,
                                const double *d, const double *du,
                                const double *dw, double *x, const int *dims,
                                int ndim, int solvedim) {
  int pd[3];
  pad_dims(dims, ndim, pd);
  pentadsolver_handle_t h{};
  pentadsolver_gpsv_batch(h, ds, dl, d, du, dw, x, pd, ndim, solvedim,
                          nullptr);
  return 0;
}

extern "C" int ref_gpsv_batch_s(const float *ds, const float *dl,
                                const float *d, const float *du,
                                const float *dw, float *x, const int *dims,
                                int ndim, int solvedim) {
  int pd[3];
  pad_dims(dims, ndim, pd);
  pentadsolver_handle_t h{};
  pentadsolver_gpsv_batch(h, ds, dl, d, du, dw, x, pd, ndim, solvedim,
                          nullptr);
  return 0;
}

// OpenMP thread count of the reference solver's parallel loops (the CPU
// baseline's thread sweep) and the binding policy in force
// (OMP_PROC_BIND; 0 false, 1 true, 2 primary, 3 close, 4 spread).
#include <omp.h>
extern "C" void ref_set_threads(int n) { omp_set_num_threads(n); }
extern "C" int ref_max_threads(void) { return omp_get_max_threads(); }
extern "C" int ref_proc_bind(void) { return (int)omp_get_proc_bind(); }
