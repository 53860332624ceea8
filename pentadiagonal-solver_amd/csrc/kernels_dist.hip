// kernels_dist.hip -- one rank's share of a line solve split across ranks.
//
// Every line of length P*m along the solve axis is cut into P slabs of m
// consecutive rows, slab k held by rank k (the layout of the reference's
// distributed solver, src/cpu/mpi/pentad_cpu.cpp:834-856).  The last two
// rows of slab k are its separator T_k; rows 0..m-3 are its interior.  With
// the separators known the interiors decouple, which gives three phases:
//
//  1. forward (per rank, k_dist_forward): the interior's first two and last
//     two rows as affine functions of (T_{k-1}, T_k), from one top-down
//     elimination sweep over the interior (banded LU without pivoting, the
//     recurrence of src/cpu/singlenode/pentad_orig_cpu.cpp:94-104) that
//     carries the first two rows along.  Substituted into the separator rows they
//     give the rank's partial 2x2 block row; the interior's first two rows
//     are what rank k-1 needs to complete its own.  Both go to the rank's
//     record (DIST_REC words per line).
//  2. reduced (all ranks' records, k_dist_reduced): each block row is
//     completed with the next rank's first rows, normalised, and the block
//     tridiagonal system in the T_k (one block row per rank and line) is
//     solved by block PCR in ceil(log2 P) rounds, one lane per rank.  On
//     several GPUs this phase is the exchange of records between neighbours
//     and the same rounds across ranks; on one GPU every rank's record sits
//     in one buffer.
//  3. backward (per rank, k_dist_halo + the batched solve): the separator
//     couplings are moved to the right-hand side of the interior's first
//     two and last two rows, the separator rows are written, and the
//     interior is solved as an ordinary batch of lines of m-2 rows by the
//     fastest kernel the planner has for it.
//
// The reference's distributed solver eliminates toward each slab's FIRST two
// rows, sends its last two rows down the line and then runs PCR or Jacobi
// on 10-word records (pentad_cpu.cpp:105-416, :544-787).  Here the
// separators are each slab's LAST two rows -- the formulation of the
// partitioned single-GPU solve (partition_core.hpp) with a slab as a chunk
// -- and the interior is re-solved in phase 3 instead of carrying four
// spike arrays from phase 1 to phase 3: phase 1 reads 6 words per row and
// writes none, phase 3 reads 6 and writes 1.
#include "partition_core.hpp"
#include "partition_exchange.hpp"
#include "pentad_internal.hpp"

namespace pentad {

// Record of one line of one rank, DIST_REC words, stored [word][line]:
//   0..4   XF0: interior row 0 as [f, s0, s1, t0, t1] (x = f + s0 T_{k-1}[0]
//          + s1 T_{k-1}[1] + t0 T_k[0] + t1 T_k[1])
//   5..9   XF1: interior row 1, the same
//   10..13 A (2x2, row-major): coupling of the separator rows to T_{k-1}
//   14..17 B (2x2): coupling to T_k, without the next slab's part
//   18..19 r: right-hand side, without the next slab's part
//   20..22 wa, ub, wb: couplings of the separator rows to the next slab's
//          rows 0 (wa, ub) and 1 (wb); zero on the last rank
// DIST_REC is declared in pentad_internal.hpp.

// One top-down elimination sweep over the interior of one line (banded LU
// without pivoting) with 5 right-hand sides [f, s0, s1, t0, t1].  Row i
// leaves x_i = z_i - u_i x_{i+1} - w_i x_{i+2}.  Interior rows 0 and 1 are
// carried along as x_r = a_r + b_r x_{i+1} + g_r x_{i+2}: substituting row
// i's relation moves them one row on (a += b z_i, b' = g - b u_i,
// g' = -b w_i), and after the last row (whose u, w vanish) they are the
// constants a_r.  So one pass gives the interior's last two rows (XL2, XL1)
// and its first two (XF0, XF1), each row read once.
template <typename T>
__device__ __forceinline__ void dist_sweep(const Diags<T> &a, const T *__restrict__ x,
                                           int64_t b, int64_t st, int n, bool first,
                                           T (&XL2)[5], T (&XL1)[5], T (&XF0)[5],
                                           T (&XF1)[5]) {
  // state of the two previous rows (1 = i-1, 2 = i-2)
  T u1 = 0, w1 = 0, u2 = 0, w2 = 0;
  T z1[5] = {0, 0, 0, 0, 0}, z2[5] = {0, 0, 0, 0, 0};
  T b0 = 0, g0 = 0, b1 = 0, g1 = 0;
#pragma unroll 2
  for (int r = 0; r < n; ++r) {
    const int64_t e = b + (int64_t)r * st;
    // couplings inside the interior
    const T back2 = r >= 2 ? a.ds[e] : T(0), back1 = r >= 1 ? a.dl[e] : T(0);
    const T fwd1 = r + 1 < n ? a.du[e] : T(0), fwd2 = r + 2 < n ? a.dw[e] : T(0);
    // right-hand sides: couplings leaving the interior move across
    T rhs[5] = {x[e], T(0), T(0), T(0), T(0)};
    if (!first && r == 0) {
      rhs[1] = -a.ds[e];
      rhs[2] = -a.dl[e];
    }
    if (!first && r == 1) rhs[2] = -a.ds[e];
    if (r == n - 2) rhs[3] = -a.dw[e];
    if (r == n - 1) {
      rhs[3] = -a.du[e];
      rhs[4] = -a.dw[e];
    }
    const T ddl = back1 - back2 * u2;
    const T iv = recip(a.d[e] - back2 * w2 - ddl * u1);
    const T un = (fwd1 - ddl * w1) * iv;
    const T wn = fwd2 * iv;
    T zn[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) zn[c] = (rhs[c] - back2 * z2[c] - ddl * z1[c]) * iv;
    // rows 0 and 1 moved on past row r (uniform branches: r is the same on
    // every lane)
    if (r == 0) {
#pragma unroll
      for (int c = 0; c < 5; ++c) XF0[c] = zn[c];
      b0 = -un;
      g0 = -wn;
    } else {
#pragma unroll
      for (int c = 0; c < 5; ++c) XF0[c] += b0 * zn[c];
      const T nb = g0 - b0 * un;
      g0 = -b0 * wn;
      b0 = nb;
    }
    if (r == 1) {
#pragma unroll
      for (int c = 0; c < 5; ++c) XF1[c] = zn[c];
      b1 = -un;
      g1 = -wn;
    } else if (r > 1) {
#pragma unroll
      for (int c = 0; c < 5; ++c) XF1[c] += b1 * zn[c];
      const T nb = g1 - b1 * un;
      g1 = -b1 * wn;
      b1 = nb;
    }
    u2 = u1;
    w2 = w1;
    u1 = un;
    w1 = wn;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      z2[c] = z1[c];
      z1[c] = zn[c];
    }
  }
  // the last row is solved; the one before it needs it (its w is 0)
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    XL1[c] = z1[c];
    XL2[c] = z2[c] - u2 * z1[c];
  }
}

// One lane per line (adjacent lines are adjacent in memory along y / z, so
// every row load is coalesced across the wavefront).
template <typename T>
__global__ void __launch_bounds__(256)
    k_dist_forward(Diags<T> a, const T *__restrict__ x, Lines ln, bool first, bool last,
                   T *__restrict__ rec) {
  const int64_t sys = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ns = ln.n_sys, st = ln.stride;
  if (sys >= ns) return;
  const int64_t b = ln.base(sys);
  const int n = (int)ln.len - 2; // interior rows

  T XL2[5], XL1[5], XF0[5], XF1[5];
  dist_sweep<T>(a, x, b, st, n, first, XL2, XL1, XF0, XF1);
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    rec[c * ns + sys] = XF0[c];
    rec[(5 + c) * ns + sys] = XF1[c];
  }
  // separator rows n (eq a) and n+1 (eq b):
  //   sa x_{n-2} + la x_{n-1} + da t0 + ua t1 + wa y0           = f_n
  //   sb x_{n-1} + lb t0 + db t1 + ub y0 + wb y1                = f_{n+1}
  // (y0, y1: the next slab's interior rows 0, 1)
  const int64_t ea = b + (int64_t)n * st, eb = ea + st;
  const T sa = a.ds[ea], la = a.dl[ea], da = a.d[ea], ua = a.du[ea];
  const T sb = a.ds[eb], lb = a.dl[eb], db = a.d[eb];
  const T wa = last ? T(0) : a.dw[ea];
  const T ub = last ? T(0) : a.du[eb];
  const T wb = last ? T(0) : a.dw[eb];
  T out[DIST_REC - 10];
  out[0] = sa * XL2[1] + la * XL1[1];
  out[1] = sa * XL2[2] + la * XL1[2];
  out[2] = sb * XL1[1];
  out[3] = sb * XL1[2];
  out[4] = sa * XL2[3] + la * XL1[3] + da;
  out[5] = sa * XL2[4] + la * XL1[4] + ua;
  out[6] = sb * XL1[3] + lb;
  out[7] = sb * XL1[4] + db;
  out[8] = x[ea] - sa * XL2[0] - la * XL1[0];
  out[9] = x[eb] - sb * XL1[0];
  out[10] = wa;
  out[11] = ub;
  out[12] = wb;
#pragma unroll
  for (int k = 0; k < DIST_REC - 10; ++k) rec[(10 + k) * ns + sys] = out[k];
}

// Phase 2 on one device: recs holds P records of n_sys lines, record k at
// recs + k * DIST_REC * n_sys.  One lane per (line, rank), L lanes per line
// (L = P rounded up to a power of two <= 64; lanes past P carry identity
// rows).  seps receives per rank [4][n_sys]: T_{k-1} (zero for k = 0) and
// T_k.
template <typename T>
__global__ void __launch_bounds__(256)
    k_dist_reduced(const T *__restrict__ recs, int64_t n_sys, int P, int lshift,
                   T *__restrict__ seps) {
  const int L = 1 << lshift;
  const int lane = (int)(threadIdx.x & 63);
  const int k = lane & (L - 1);
  const int64_t line = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> lshift;
  // every lane of a wavefront takes part in the exchanges: lines past the
  // end repeat the last line (their results are not written)
  const bool line_ok = line < n_sys;
  const int64_t ln = line_ok ? line : n_sys - 1;
  const bool real = k < P;
  const int64_t R = (int64_t)DIST_REC * n_sys;
  M2<T> A{0, 0, 0, 0}, C{0, 0, 0, 0};
  T r0 = 0, r1 = 0;
  if (real) {
    const T *own = recs + k * R + ln;
    auto g = [&](int w) { return own[(int64_t)w * n_sys]; };
    A = {g(10), g(11), g(12), g(13)};
    M2<T> B{g(14), g(15), g(16), g(17)};
    r0 = g(18);
    r1 = g(19);
    if (k + 1 < P) {
      const T *nx = recs + (k + 1) * R + ln;
      auto y = [&](int w) { return nx[(int64_t)w * n_sys]; };
      const T wa = g(20), ub = g(21), wb = g(22);
      // next slab's interior rows 0 (Y0) and 1 (Y1): [f, s0, s1, t0, t1]
      // with (s0, s1) = T_k and (t0, t1) = T_{k+1}
      const T Y0[5] = {y(0), y(1), y(2), y(3), y(4)};
      const T Y1[5] = {y(5), y(6), y(7), y(8), y(9)};
      B.a00 += wa * Y0[1];
      B.a01 += wa * Y0[2];
      B.a10 += ub * Y0[1] + wb * Y1[1];
      B.a11 += ub * Y0[2] + wb * Y1[2];
      C = {wa * Y0[3], wa * Y0[4], ub * Y0[3] + wb * Y1[3], ub * Y0[4] + wb * Y1[4]};
      r0 -= wa * Y0[0];
      r1 -= ub * Y0[0] + wb * Y1[0];
    }
    // normalise: T_k + A T_{k-1} + C T_{k+1} = r
    const M2<T> Bi = inv(B);
    A = mul(Bi, A);
    C = mul(Bi, C);
    const T t0 = Bi.a00 * r0 + Bi.a01 * r1;
    const T t1 = Bi.a10 * r0 + Bi.a11 * r1;
    r0 = t0;
    r1 = t1;
  }
  WaveEx ex{lane, k, L};
  block_pcr<T>(A, C, r0, r1, L, ex);
  // T_{k-1} from the lane before (rank 0 has none)
  T prev0, prev1;
  {
    T send[2] = {r0, r1}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    prev0 = k > 0 ? rp[0] : T(0);
    prev1 = k > 0 ? rp[1] : T(0);
  }
  if (real && line_ok) {
    T *o = seps + (int64_t)k * 4 * n_sys + line;
    o[0] = prev0;
    o[n_sys] = prev1;
    o[2 * n_sys] = r0;
    o[3 * n_sys] = r1;
  }
}

// Phase 3, first half: move the separator couplings of the interior's end
// rows to the right-hand side and write the separator rows.  sep: this
// rank's [4][n_sys] from phase 2.
template <typename T>
__global__ void __launch_bounds__(256)
    k_dist_halo(Diags<T> a, T *__restrict__ x, Lines ln, bool first,
                const T *__restrict__ sep) {
  const int64_t sys = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= ln.n_sys) return;
  const int64_t ns = ln.n_sys, st = ln.stride;
  const int64_t b = ln.base(sys);
  const int n = (int)ln.len - 2;
  const T p0 = sep[sys], p1 = sep[ns + sys], t0 = sep[2 * ns + sys], t1 = sep[3 * ns + sys];
  // interior rows n-2, n-1 couple to T_k; rows 0, 1 to T_{k-1} (n >= 3, so
  // rows 0 and n-1 differ; when n == 3 row 1 is row n-2 and gets both)
  T f[4];
  const int64_t e0 = b, e1 = b + st, ea = b + (int64_t)(n - 2) * st, eb = ea + st;
  f[0] = x[e0];
  f[1] = x[e1];
  if (!first) {
    f[0] -= a.ds[e0] * p0 + a.dl[e0] * p1;
    f[1] -= a.ds[e1] * p1;
  }
  x[e0] = f[0];
  x[e1] = f[1];
  f[2] = x[ea] - a.dw[ea] * t0; // re-read: row n-2 may be row 1
  f[3] = x[eb] - a.du[eb] * t0 - a.dw[eb] * t1;
  x[ea] = f[2];
  x[eb] = f[3];
  x[eb + st] = t0;
  x[eb + 2 * st] = t1;
}

template <typename T>
hipError_t launch_dist_forward(const Diags<T> &a, const T *x, const Lines &ln, bool first,
                               bool last, T *rec, hipStream_t st) {
  if (ln.len < 5) return hipErrorInvalidValue;
  const int block = 256;
  const int64_t grid = (ln.n_sys + block - 1) / block;
  hipLaunchKernelGGL(k_dist_forward<T>, dim3((unsigned)grid), dim3(block), 0, st, a, x, ln,
                     first, last, rec);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_dist_reduced(const T *recs, int64_t n_sys, int P, T *seps, hipStream_t st) {
  if (P < 1 || P > 64 || n_sys < 1) return hipErrorInvalidValue;
  int lshift = 0;
  while ((1 << lshift) < P) ++lshift;
  const int block = 256;
  const int64_t lanes = n_sys << lshift;
  const int64_t grid = (lanes + block - 1) / block;
  hipLaunchKernelGGL(k_dist_reduced<T>, dim3((unsigned)grid), dim3(block), 0, st, recs, n_sys,
                     P, lshift, seps);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_dist_halo(const Diags<T> &a, T *x, const Lines &ln, bool first, const T *sep,
                            hipStream_t st) {
  if (ln.len < 5) return hipErrorInvalidValue;
  const int block = 256;
  const int64_t grid = (ln.n_sys + block - 1) / block;
  hipLaunchKernelGGL(k_dist_halo<T>, dim3((unsigned)grid), dim3(block), 0, st, a, x, ln, first,
                     sep);
  return hipGetLastError();
}

#define DIST_INST(T)                                                                        \
  template hipError_t launch_dist_forward<T>(const Diags<T> &, const T *, const Lines &, bool, \
                                             bool, T *, hipStream_t);                       \
  template hipError_t launch_dist_reduced<T>(const T *, int64_t, int, T *, hipStream_t);    \
  template hipError_t launch_dist_halo<T>(const Diags<T> &, T *, const Lines &, bool,        \
                                          const T *, hipStream_t);
DIST_INST(float)
DIST_INST(double)

} // namespace pentad
