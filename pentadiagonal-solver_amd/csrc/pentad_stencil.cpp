// pentad_stencil.cpp -- host side of the structured-coefficient solves
// (pentadsolver_{D,S}_gpsv_stencil_batch): the five diagonals of every line
// are id * I + scale[line] * [w0 .. w4], generated inside the kernels
// (kernels_stencil.hip), so a solve streams only the right-hand side and the
// solution.  Same shape contract as pentadsolver_*_gpsv_batch (t_dims[0]
// fastest, in place on x, asynchronous on the handle's stream); no scratch.
// No reference counterpart.
#include "pentad_api_internal.hpp"

#include <cstdint>
#include <cstdio>
#include <string>

namespace pentad {

namespace {

bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

} // namespace

// x lines: a partition of the line into lanes * rows with lanes a power of
// two <= 64 (rows per lane preferred 8 for FP64, 16 for FP32)
bool stencil_plan_x(int elem_bytes, int64_t len, Plan &pl) {
  static const int pref64[] = {8, 6, 10, 12, 4, 16};
  static const int pref32[] = {16, 12, 8, 10, 6, 4};
  const int *pref = elem_bytes == 8 ? pref64 : pref32;
  for (int k = 0; k < 6; ++k) {
    const int r = pref[k];
    if (len % r) continue;
    const int64_t lanes = len / r;
    if (lanes < 2 || lanes > 64 || (lanes & (lanes - 1))) continue;
    if (!part_x_supported(elem_bytes, r, (int)lanes)) continue;
    pl.kind = Plan::PartX;
    pl.rows = r;
    pl.lanes = (int)lanes;
    return true;
  }
  return false;
}

namespace {

// strided lines (or x lines no partition above fits): rows per lane and
// systems per workgroup, whole chunks first, then a padded last chunk
bool plan_yz(int elem_bytes, int64_t len, Plan &pl) {
  static const int pref64[] = {8, 6, 12, 4, 16};
  static const int pref32[] = {16, 12, 8, 6, 4};
  const int *pref = elem_bytes == 8 ? pref64 : pref32;
  for (int k = 0; k < 10; ++k) {
    const int r = pref[k % 5];
    if ((k < 5) != (len % r == 0)) continue;
    for (int S = 64; S >= 1; S >>= 1) {
      if (stencil_yz_supported(elem_bytes, r, S, len)) {
        pl.kind = Plan::PartYZ;
        pl.rows = r;
        pl.sys_per_wg = S;
        return true;
      }
    }
  }
  return false;
}

template <typename T>
pentadStatus_t stencil_batch(Context *c, const T *scale, const double *weights,
                             double id_weight, T *x, const int *dims, int ndim,
                             int solvedim) {
  Lines ln;
  if (decompose(dims, ndim, solvedim, &ln) != 0 || ln.len < 3) {
    std::fprintf(stderr, "[penta] stencil: invalid shape (ndim %d, solvedim %d)\n", ndim,
                 solvedim);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (!scale || !weights || !x) {
    std::fprintf(stderr, "[penta] stencil: null pointer\n");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  Stencil<T> stc;
  stc.scale = scale;
  stc.sa = 1; // x lines in place: the XMap split is lo = system, hi = 0
  stc.sb = 0;
  for (int k = 0; k < 5; ++k) stc.w[k] = (T)weights[k];
  stc.id = (T)id_weight;
  Plan pl;
  hipError_t err;
  if (ln.dir == 0 && aligned16(x) && stencil_plan_x((int)sizeof(T), ln.len, pl)) {
    err = launch_stencil_x_map<T>(stc, x, x, false, ln.n_sys, ln.len, xmap_inplace(ln),
                                  pl.rows, pl.lanes, c->stream);
  } else if (plan_yz((int)sizeof(T), ln.len, pl)) {
    err = launch_stencil_yz<T>(stc, x, ln, pl.rows, pl.sys_per_wg, c->stream);
  } else {
    std::fprintf(stderr, "[penta] stencil: no kernel for length %lld\n", (long long)ln.len);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (err != hipSuccess) {
    std::fprintf(stderr, "[penta] stencil solvedim %d launch FAILED: %s -- output is invalid\n",
                 solvedim, hipGetErrorString(err));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  note_kernel(c, ln.dir, "stencil", "stencil " + plan_config(pl));
  return PENTAD_STATUS_SUCCESS;
}

} // namespace

} // namespace pentad

using namespace pentad;

extern "C" {

pentadStatus_t pentadsolver_D_gpsv_stencil_batch(pentadsolver_handle_t handle,
                                                 const double *scale, const double *weights,
                                                 double id_weight, double *x,
                                                 const int *t_dims, int t_ndim,
                                                 int t_solvedim) {
  Context *c = ctx_of(handle);
  const pentadStatus_t s =
      stencil_batch<double>(c, scale, weights, id_weight, x, t_dims, t_ndim, t_solvedim);
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_S_gpsv_stencil_batch(pentadsolver_handle_t handle,
                                                 const float *scale, const double *weights,
                                                 double id_weight, float *x,
                                                 const int *t_dims, int t_ndim,
                                                 int t_solvedim) {
  Context *c = ctx_of(handle);
  const pentadStatus_t s =
      stencil_batch<float>(c, scale, weights, id_weight, x, t_dims, t_ndim, t_solvedim);
  set_status(c, s);
  return s;
}

} // extern "C"
