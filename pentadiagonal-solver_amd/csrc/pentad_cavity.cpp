// pentad_cavity.cpp -- host side of the 2-D cavity consumer's pressure-
// correction step (apps/2D_cavity_cpu.cpp:195-219 with the assembly of
// apps/preproc.hpp:248-525) on the GPU:
//
//   pentadsolver_D_cavity_pressure_coef   the assembly, writing the padded
//                                         arrays (drop-in for
//                                         pressure_coef_matrix)
//   pentadsolver_D_cavity_pressure_solve  assembly + x solve + y solve with
//                                         the coefficients generated inside
//                                         the solve kernels (no arrays)
//
// Both are asynchronous on the handle's stream and return the status.
#include "pentad_api_internal.hpp"

#include <cstdio>

namespace pentad {
namespace {

pentadStatus_t check_shape(int nx, int ny, int nx_pad, const double *up, const double *vp,
                           const double *uv, const void *x) {
  if (nx < 3 || ny < 3 || nx_pad < nx) {
    std::fprintf(stderr, "[penta] cavity: need nx, ny >= 3 and nx_pad >= nx (got %d, %d, %d)\n",
                 nx, ny, nx_pad);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (!up || !vp || !uv || !x) {
    std::fprintf(stderr, "[penta] cavity: null pointer\n");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  return PENTAD_STATUS_SUCCESS;
}

Cavity make(int nx, int ny, int nx_pad, double alpha_u, double dx, double dy,
            const double *up, const double *vp, const double *uv) {
  Cavity c;
  c.nx = nx;
  c.ny = ny;
  c.nx_pad = nx_pad;
  c.alpha_u = alpha_u;
  c.dx = dx;
  c.dy = dy;
  c.up = up;
  c.vp = vp;
  c.uv = uv;
  return c;
}

pentadStatus_t finish(Context *c, hipError_t e, const char *what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] cavity %s launch FAILED: %s -- output is invalid\n", what,
                 hipGetErrorString(e));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  (void)c;
  return PENTAD_STATUS_SUCCESS;
}

} // namespace
} // namespace pentad

using namespace pentad;

extern "C" {

pentadStatus_t pentadsolver_D_cavity_pressure_coef(
    pentadsolver_handle_t handle, int nx, int ny, int nx_pad, double alphaU, double dx,
    double dy, const double *u_prime, const double *v_prime, const double *uv, double *x,
    double *s_x, double *l_x, double *d_x, double *u_x, double *w_x, double *s_y,
    double *l_y, double *d_y, double *u_y, double *w_y) {
  Context *c = ctx_of(handle);
  pentadStatus_t s = check_shape(nx, ny, nx_pad, u_prime, v_prime, uv, x);
  const bool yset = s_y || l_y || d_y || u_y || w_y;
  if (s == PENTAD_STATUS_SUCCESS &&
      (!s_x || !l_x || !d_x || !u_x || !w_x || (yset && (!s_y || !l_y || !d_y || !u_y || !w_y)))) {
    std::fprintf(stderr, "[penta] cavity: the x set needs five arrays, the y set five or none\n");
    s = PENTAD_STATUS_INVALID_VALUE;
  }
  if (s == PENTAD_STATUS_SUCCESS) {
    CavityOut o;
    double *const p[11] = {x, s_x, l_x, d_x, u_x, w_x, s_y, l_y, d_y, u_y, w_y};
    for (int k = 0; k < 11; ++k) o.p[k] = p[k];
    if (!yset) o.p[6] = nullptr;
    s = finish(c, launch_cavity_assemble(make(nx, ny, nx_pad, alphaU, dx, dy, u_prime, v_prime,
                                              uv),
                                         o, c->stream),
               "assembly");
  }
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_D_cavity_pressure_solve(pentadsolver_handle_t handle, int nx,
                                                    int ny, int nx_pad, double alphaU,
                                                    double dx, double dy,
                                                    const double *u_prime,
                                                    const double *v_prime, const double *uv,
                                                    double *x) {
  Context *c = ctx_of(handle);
  pentadStatus_t s = check_shape(nx, ny, nx_pad, u_prime, v_prime, uv, x);
  int mx = 0, lanes = 0, my = 0, spw = 0;
  if (s == PENTAD_STATUS_SUCCESS &&
      (!cavity_plan_x(nx, &mx, &lanes) || !cavity_plan_y(ny, &my, &spw))) {
    std::fprintf(stderr,
                 "[penta] cavity solve: nx <= 2048 and ny <= 8192 (got %d x %d); assemble with "
                 "pentadsolver_D_cavity_pressure_coef and solve with the padded batch solve\n",
                 nx, ny);
    s = PENTAD_STATUS_INVALID_VALUE;
  }
  if (s == PENTAD_STATUS_SUCCESS) {
    // pentadsolver_set_kernel_params: y (direction 2 of a 2-D grid) takes
    // rows per lane and columns per workgroup when they fit
    int fr, fw;
    {
      std::lock_guard<std::mutex> lk(c->mu);
      fr = c->force_rows[2];
      fw = c->force_width[2];
    }
    if (fr > 0 && fw > 0 && cavity_y_supported(ny, fr, fw)) {
      my = fr;
      spw = fw;
    }
    s = finish(c, launch_cavity_solve(make(nx, ny, nx_pad, alphaU, dx, dy, u_prime, v_prime,
                                           uv),
                                      x, mx, lanes, my, spw, c->stream),
               "solve");
    if (s == PENTAD_STATUS_SUCCESS) {
      char b0[64], b1[64];
      std::snprintf(b0, sizeof b0, "cavity rows=%d lanes=%d", mx, lanes);
      std::snprintf(b1, sizeof b1, "cavity rows=%d sys_per_wg=%d", my, spw);
      note_kernel(c, 0, "cavity", b0);
      note_kernel(c, 2, "cavity", b1);
    }
  }
  set_status(c, s);
  return s;
}

} // extern "C"
