// cavity_shim.cpp -- C entry point over the reference 2-D cavity consumer's
// pressure-correction assembly (pressure_coef_matrix, apps/preproc.hpp:248),
// compiled from the reference's own header by oracle/Makefile into
// oracle/_ref/libref_cavity.so.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ to pin the oracle's
// restatement (po_cavity_pressure_coef_d) bit for bit.  The product library
// never links or loads it.
//
// The reference takes its inputs as std::vector<std::vector<double>> with
// extents u_prime (ny+2) x (nx+1), v_prime (ny+1) x (nx+2), uv (ny+2) x
// (nx+2) (apps/2D_cavity_cpu.cpp:141-160); this shim takes them row-major.
#include "preproc.hpp"

#include <vector>

namespace {
std::vector<std::vector<double>> rows(const double *a, int nr, int nc) {
  std::vector<std::vector<double>> v(nr, std::vector<double>(nc));
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) v[i][j] = a[(size_t)i * nc + j];
  return v;
}
} // namespace

extern "C" int ref_cavity_pressure_coef(int nx, int ny, int nx_pad, double alphaU,
                                        double dx, double dy, const double *u_prime,
                                        const double *v_prime, const double *uv, double *x,
                                        double *s_x, double *l_x, double *d_x, double *u_x,
                                        double *w_x, double *s_y, double *l_y, double *d_y,
                                        double *u_y, double *w_y) {
  auto up = rows(u_prime, ny + 2, nx + 1);
  auto vp = rows(v_prime, ny + 1, nx + 2);
  auto a = rows(uv, ny + 2, nx + 2);
  pressure_coef_matrix<double>(nx, ny, nx_pad, alphaU, dx, dy, up, vp, a, x, s_x, l_x, d_x,
                               u_x, w_x, s_y, l_y, d_y, u_y, w_y);
  return 0;
}
