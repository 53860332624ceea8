// kernels_cavity.hip -- the 2-D cavity consumer's pressure-correction solve.
//
// The reference's 2-D consumer (apps/2D_cavity_cpu.cpp:195-219) assembles,
// every iteration, a pentadiagonal system per grid row and per grid column
// from the momentum diagonal uv = A_p and the face velocities u', v'
// (apps/preproc.hpp:248-525, pressure_coef_matrix), writes eleven arrays (the
// right-hand side and two identical sets of five diagonals), then solves
// along x and along y in place on the padded field (rows of nx_pad).
//
// cavity_point() is that assembly for one grid point.  Three kernels use it:
//   k_cavity_assemble  writes the arrays (the drop-in for the assembly);
//   k_cavity_x         the x solve with the coefficients and right-hand side
//                      generated in registers (partitioned solve, lanes of a
//                      row exchange through the wavefront);
//   k_cavity_y         the y solve with the coefficients generated in
//                      registers and the right-hand side the x solve left
//                      (S columns x P chunks per workgroup, LDS exchange).
// The fused pair reads uv, u' and v' (24 B per point) and the field once
// more for y: 56 B per point for assembly + both solves, against 224 B when
// the eleven arrays are written and the two solves read them back.
//
// Arithmetic follows the reference's expressions in their order, with
// contraction off, except the right-hand side, which the reference's build
// (-O3 -march=core-avx2) fuses into one multiply-add: the values are the
// reference's bit for bit.
#include "partition_core.hpp"
#include "partition_exchange.hpp"
#include "pentad_internal.hpp"

namespace pentad {

// The five entries (ds, dl, d, du, dw) and right-hand side of point (i, j),
// 1 <= i <= ny, 1 <= j <= nx (nx, ny >= 3).  With
// c(a, b) = 0.5 * alphaU * (1/a + 1/b) the couplings are E = c(uv_ij,
// uv_i,j+1) dy^2, W = c(uv_ij, uv_i,j-1) dy^2, N = c(uv_ij, uv_i-1,j) dx^2,
// S = c(uv_ij, uv_i+1,j) dx^2, assigned per position as the reference does
// (1 where it assigns nothing, d = du + dl + ds + dw):
//   interior j = 1: ds 1, dl N, du E, dw S        (preproc.hpp:273-295)
//   interior j = nx: ds N, dl W, du 1, dw S       (:297-318)
//   interior: ds N, dl W, du E, dw S              (:319-341)
//   top row: ds 1, dl W, du E, dw S               (:347-376)
//   bottom row: ds N, dl W, du E, dw = S of (1, nx-1), the value its loop
//     inherits from the top row's last point     (:378-401)
//   corners (1,1): E, S; (1,nx): W, S; (ny,1): E, N; (ny,nx): W, N  (:403-520)
__device__ __forceinline__ double cav_c(double ia, double ib, double half_au, double h2) {
#pragma clang fp contract(off)
  return half_au * (ia + ib) * h2;
}

// The same with the five reciprocals given: 1/uv of the point (ic) and of
// its east, west, north and south neighbours.  Callers walking along a line
// pass each reciprocal to the next point (it is that point's west or north
// one), three divisions per point instead of five; division is correctly
// rounded, so the values are the same bit for bit.
__device__ __forceinline__ void cavity_coef_inv(const Cavity &c, int i, int j, double ic,
                                                double ie, double iw, double in_, double is_,
                                                double &ds, double &dl, double &d, double &du,
                                                double &dw) {
#pragma clang fp contract(off)
  const int lv = c.nx + 2;
  const double half_au = 0.5 * c.alpha_u;
  const double dx2 = c.dx * c.dx, dy2 = c.dy * c.dy;
  const double E = cav_c(ic, ie, half_au, dy2);
  const double W = cav_c(ic, iw, half_au, dy2);
  const double N = cav_c(ic, in_, half_au, dx2);
  const double S = cav_c(ic, is_, half_au, dx2);
  const bool top = i == 1, bot = i == c.ny, left = j == 1, right = j == c.nx;
  double s = N, l = W, u = E, w = S;
  if (top) {
    s = 1.0;
    if (left) l = 1.0;
    if (right) u = 1.0;
  } else if (bot) {
    if (left) { l = 1.0; w = 1.0; }
    else if (right) { u = 1.0; w = 1.0; }
    else {
      const double *u1 = c.uv + lv + (c.nx - 1); // (1, nx - 1)
      w = cav_c(1.0 / u1[0], 1.0 / u1[lv], half_au, dx2);
    }
  } else if (left) {
    s = 1.0;
    l = N;
  } else if (right) {
    u = 1.0;
  }
  ds = s;
  dl = l;
  du = u;
  dw = w;
  d = u + l + s + w;
}

// The right-hand side of point (i, j) from its four face values: u'(i, j),
// u'(i, j-1), v'(i-1, j), v'(i, j).
__device__ __forceinline__ double cavity_rhs(const Cavity &c, double ue, double uw, double vn,
                                             double vs) {
#pragma clang fp contract(off)
  const double p = -(ue - uw);
  const double q = vn - vs;
  return __builtin_fma(p, c.dy, -(q * c.dx));
}

__device__ __forceinline__ void cavity_point_inv(const Cavity &c, int i, int j, double ic,
                                                 double ie, double iw, double in_, double is_,
                                                 double &ds, double &dl, double &d, double &du,
                                                 double &dw, double &rhs) {
  const int lu = c.nx + 1, lv = c.nx + 2;
  cavity_coef_inv(c, i, j, ic, ie, iw, in_, is_, ds, dl, d, du, dw);
  rhs = cavity_rhs(c, c.up[(int64_t)i * lu + j], c.up[(int64_t)i * lu + j - 1],
                   c.vp[(int64_t)(i - 1) * lv + j], c.vp[(int64_t)i * lv + j]);
}

// v[k] = p[k] for k < N with 16-byte loads.  p is 8-byte aligned; the pairs
// read are exactly the 16-byte blocks holding p[0] .. p[N-1] (the last load
// repeats an earlier block when the run starts on a block boundary), so
// nothing outside the run's blocks is touched.
template <int N>
__device__ __forceinline__ void load_run(const double *p, double (&v)[N]) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  const bool odd = (a & 8u) != 0;
  const double2 *q = reinterpret_cast<const double2 *>(a & ~uintptr_t(15));
  constexpr int NP = N / 2 + 1;
  const int last = odd ? N / 2 : (N + 1) / 2 - 1;
  double2 t[NP];
#pragma unroll
  for (int k = 0; k < NP; ++k) t[k] = q[k < last ? k : last];
  // a select between two loaded values, not between two indices (an index
  // select would leave the array in scratch memory)
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double lo = (k & 1) ? t[k >> 1].y : t[k >> 1].x;
    const double hi = (k & 1) ? t[(k + 1) >> 1].x : t[k >> 1].y;
    v[k] = odd ? hi : lo;
  }
}

__device__ __forceinline__ void cavity_point(const Cavity &c, int i, int j, double &ds,
                                             double &dl, double &d, double &du, double &dw,
                                             double &rhs) {
#pragma clang fp contract(off)
  const int lv = c.nx + 2;
  const double *uv = c.uv + (int64_t)i * lv + j;
  cavity_point_inv(c, i, j, 1.0 / uv[0], 1.0 / uv[1], 1.0 / uv[-1], 1.0 / uv[-lv],
                   1.0 / uv[lv], ds, dl, d, du, dw, rhs);
}

// one thread per grid point; point (i, j) to (i-1) * nx_pad + (j-1)
__global__ void __launch_bounds__(256) k_cavity_assemble(Cavity c, CavityOut o) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)c.nx * c.ny) return;
  const int i = (int)(t / c.nx) + 1, j = (int)(t % c.nx) + 1;
  double v[6];
  cavity_point(c, i, j, v[1], v[2], v[3], v[4], v[5], v[0]);
  const int64_t k = (int64_t)(i - 1) * c.nx_pad + (j - 1);
#pragma unroll
  for (int a = 0; a < 6; ++a) o.p[a][k] = v[a];
  if (o.p[6] != nullptr) {
#pragma unroll
    for (int a = 1; a < 6; ++a) o.p[5 + a][k] = v[a];
  }
}

// rows r0 .. r0+M-1 of the line through (i, j) along x (LINE_X) or y: the
// reference's entries, couplings past either end dropped, rows past the end
// identity rows with right-hand side 0
template <int M, bool LINE_X>
__device__ __forceinline__ void cavity_rows(const Cavity &c, int fixed, int r0, double (&s)[M],
                                            double (&l)[M], double (&d)[M], double (&u)[M],
                                            double (&w)[M], double (&f)[M]) {
#pragma clang fp contract(off)
  const int len = LINE_X ? c.nx : c.ny;
  const int lv = c.nx + 2;
  // 1/uv along the line: point r (grid index r + 1) needs line indices r,
  // r + 1, r + 2, carried from point to point; indices past the far halo
  // (rows past the end, whose values are discarded) are clamped to it
  auto along = [&](int q) {
    const int qc = q <= len + 1 ? q : len + 1;
    return LINE_X ? c.uv[(int64_t)fixed * lv + qc] : c.uv[(int64_t)qc * lv + fixed];
  };
  if constexpr (LINE_X) {
    // a whole chunk inside the line: its runs of uv, u' and v' come in with
    // 16-byte loads (consecutive lanes are M elements apart, so one scalar
    // load per element would touch a new 128-byte line per lane each time)
    if (r0 + M <= len) {
      const int lu = c.nx + 1;
      const int64_t row = (int64_t)fixed * lv + r0;
      double a[M + 2], un[M], us[M], ur[M + 1], vn[M], vs[M];
      load_run(c.uv + row, a);
      load_run(c.uv + row - lv + 1, un);
      load_run(c.uv + row + lv + 1, us);
      load_run(c.up + (int64_t)fixed * lu + r0, ur);
      load_run(c.vp + row - lv + 1, vn);
      load_run(c.vp + row + 1, vs);
      double iw = 1.0 / a[0], ic = 1.0 / a[1];
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int r = r0 + k;
        const double ie = 1.0 / a[k + 2];
        cavity_coef_inv(c, fixed, r + 1, ic, ie, iw, 1.0 / un[k], 1.0 / us[k], s[k], l[k],
                        d[k], u[k], w[k]);
        iw = ic;
        ic = ie;
        f[k] = cavity_rhs(c, ur[k + 1], ur[k], vn[k], vs[k]);
        if (r < 2) s[k] = 0.0;
        if (r < 1) l[k] = 0.0;
        if (r + 1 >= len) u[k] = 0.0;
        if (r + 2 >= len) w[k] = 0.0;
      }
      return;
    }
  }
  double iw = 1.0 / along(r0), ic = 1.0 / along(r0 + 1);
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int r = r0 + k;
    const double ie = 1.0 / along(r + 2);
    // rows past the end evaluate a real point (the line's last) and are
    // then overwritten by identity rows: no divergent branch around the
    // generator
    const int rc = r < len ? r : len - 1;
    double rhs;
    if constexpr (LINE_X) {
      const double *uvp = c.uv + (int64_t)fixed * lv + (rc + 1);
      cavity_point_inv(c, fixed, rc + 1, ic, ie, iw, 1.0 / uvp[-lv], 1.0 / uvp[lv], s[k], l[k],
                       d[k], u[k], w[k], rhs);
    } else {
      const double *uvp = c.uv + (int64_t)(rc + 1) * lv + fixed;
      // along a column the line's neighbours are north (before) and south
      cavity_point_inv(c, rc + 1, fixed, ic, 1.0 / uvp[1], 1.0 / uvp[-1], iw, ie, s[k], l[k],
                       d[k], u[k], w[k], rhs);
    }
    iw = ic;
    ic = ie;
    if constexpr (LINE_X) f[k] = rhs;
    if (r < 2) s[k] = 0.0;
    if (r < 1) l[k] = 0.0;
    if (r + 1 >= len) u[k] = 0.0;
    if (r + 2 >= len) w[k] = 0.0;
    if (r >= len) {
      d[k] = 1.0;
      if constexpr (LINE_X) f[k] = 0.0;
    }
  }
}

// x solve of every grid row i (system i - 1), L = 2^lshift lanes of M rows
// each, 256 / L rows per workgroup; the solution goes to x[(i-1)*nx_pad + j-1].
// WIDE: more than 64 lanes per row (lines above 64 x M points), the lanes of
// a row span wavefronts and exchange through LDS; 8-row chunks that way keep
// three wavefronts per SIMD where 16-row chunks in one wavefront hold one.
template <int M, bool WIDE>
__global__ void __launch_bounds__(256) k_cavity_x(Cavity c, double *__restrict__ x,
                                                  int lshift) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = 1 << lshift;
  const int sl = (int)threadIdx.x >> lshift;
  const int p = (int)threadIdx.x & (L - 1);
  const int64_t sys = ((int64_t)blockIdx.x << (8 - lshift)) + sl;
  const int i = (int)(sys < c.ny ? sys : c.ny - 1) + 1;
  double s[M], l[M], d[M], u[M], w[M], f[M];
  cavity_rows<M, true>(c, i, p * M, s, l, d, u, w, f);
  if constexpr (WIDE) {
    LdsEx ex{smem_raw, sl, p, 256 >> lshift, L};
    part_solve<double, M>(s, l, d, u, w, f, p, L, ex);
  } else {
    WaveEx ex{(int)(threadIdx.x & 63), p, L};
    part_solve<double, M>(s, l, d, u, w, f, p, L, ex);
  }
  if (sys < c.ny) {
    double *row = x + (int64_t)(i - 1) * c.nx_pad;
#pragma unroll
    for (int k = 0; k < M; ++k)
      if (p * M + k < c.nx) row[p * M + k] = f[k];
  }
}

// y solve of every grid column j (system j - 1): S columns x P chunks of M
// rows per workgroup, thread t -> column t % S, chunk t / S
template <int M>
__global__ void __launch_bounds__(512) k_cavity_y(Cavity c, double *__restrict__ x, int S,
                                                  int P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int t = threadIdx.x;
  const int sl = t % S, p = t / S;
  const int64_t sys = xcd_tile(blockIdx.x, gridDim.x) * S + sl;
  const bool valid = sys < c.nx;
  const int j = (int)(valid ? sys : c.nx - 1) + 1;
  double *col = x + (j - 1);
  double s[M], l[M], d[M], u[M], w[M], f[M];
#pragma unroll
  for (int k = 0; k < M; ++k) {
    const int r = p * M + k;
    f[k] = r < c.ny ? col[(int64_t)r * c.nx_pad] : 0.0;
  }
  cavity_rows<M, false>(c, j, p * M, s, l, d, u, w, f);
  LdsEx ex{smem_raw, sl, p, S, P};
  part_solve<double, M>(s, l, d, u, w, f, p, P, ex);
  if (valid) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const int r = p * M + k;
      if (r < c.ny) col[(int64_t)r * c.nx_pad] = f[k];
    }
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// 8-row chunks, a power-of-two number of lanes per row: up to 64 in one
// wavefront, up to 256 (2048 points) across the workgroup
bool cavity_plan_x(int nx, int *rows, int *lanes) {
  int l = 1;
  while ((int64_t)l * 8 < nx) l <<= 1;
  if (l > 256) return false;
  *rows = 8;
  *lanes = l;
  return true;
}

// The widest workgroup of columns 8-row chunks allow; 16-row chunks (two
// wavefronts per SIMD instead of three) when 8-row chunks leave fewer than 4
// columns per workgroup and 16-row ones leave more.  Measured, fused step:
// 1024 x 2048 0.071 -> 0.054 ms (4 columns of 16-row chunks against 2 of
// 8-row ones); at 1024 x 1024 4 columns of either are equal (0.036) and 8
// columns of 16-row chunks slower (0.041).
bool cavity_plan_y(int ny, int *rows, int *sys_per_wg) {
  auto widest = [&](int m) {
    const int64_t P = (ny + m - 1) / m;
    int S = 64;
    while (S >= 1 && S * P > 512) S >>= 1;
    return S; // 0 when not even one column fits
  };
  const int s8 = widest(8), s16 = widest(16);
  if (s8 >= 4 || (s8 >= 1 && s16 <= s8)) {
    *rows = 8;
    *sys_per_wg = s8;
    return true;
  }
  if (s16 >= 1) {
    *rows = 16;
    *sys_per_wg = s16 < 4 ? s16 : 4;
    return true;
  }
  return false;
}

hipError_t launch_cavity_assemble(const Cavity &c, const CavityOut &o, hipStream_t st) {
  const int64_t n = (int64_t)c.nx * c.ny;
  const int64_t grid = (n + 255) / 256;
  hipLaunchKernelGGL(k_cavity_assemble, dim3((unsigned)grid), dim3(256), 0, st, c, o);
  return hipGetLastError();
}

bool cavity_y_supported(int ny, int rows, int sys_per_wg) {
  if (rows != 8 && rows != 16) return false;
  if (sys_per_wg < 1 || sys_per_wg > 64 || (sys_per_wg & (sys_per_wg - 1))) return false;
  return (int64_t)sys_per_wg * ((ny + rows - 1) / rows) <= 512;
}

hipError_t launch_cavity_solve(const Cavity &c, double *x, int mx, int lanes, int my, int S,
                               hipStream_t st) {
  if (mx != 8 || lanes < 1 || lanes > 256 || (lanes & (lanes - 1)) ||
      (int64_t)lanes * mx < c.nx || !cavity_y_supported(c.ny, my, S))
    return hipErrorInvalidValue;
  int lshift = 0;
  while ((1 << lshift) < lanes) ++lshift;
  const int64_t rows_per_wg = 256 >> lshift;
  const int64_t gx = (c.ny + rows_per_wg - 1) / rows_per_wg;
  if (lanes <= 64)
    hipLaunchKernelGGL((k_cavity_x<8, false>), dim3((unsigned)gx), dim3(256), 0, st, c, x,
                       lshift);
  else
    hipLaunchKernelGGL((k_cavity_x<8, true>), dim3((unsigned)gx), dim3(256),
                       (size_t)10 * 256 * sizeof(double), st, c, x, lshift);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int P = (c.ny + my - 1) / my;
  const size_t smem = (size_t)10 * S * P * sizeof(double);
  const int64_t gy = (c.nx + S - 1) / S;
  auto go = [&](auto kernel) {
    if (smem > 65536) {
      const hipError_t a = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)smem);
      if (a != hipSuccess) return a;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)gy), dim3(S * P), smem, st, c, x, S, P);
    return hipGetLastError();
  };
  return my == 8 ? go(k_cavity_y<8>) : go(k_cavity_y<16>);
}

} // namespace pentad
