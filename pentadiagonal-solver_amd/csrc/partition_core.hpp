// partition_core.hpp -- register-resident partitioned pentadiagonal solve.
//
// One system of length len is cut into P chunks of M consecutive rows; one
// lane owns one chunk and keeps every row of it in registers.  The last two
// rows of each chunk are its "separator" S_p; the first M-2 rows are its
// interior.  Because the matrix has bandwidth 2, removing the separators
// decouples the interiors, so:
//
//  1. each lane eliminates its interior (banded LU without pivoting, the
//     recurrence of src/cpu/singlenode/pentad_orig_cpu.cpp:94-104) carrying
//     three right-hand sides: the RHS f and the couplings to the previous
//     separator S_{p-1} = (x[r0-2], x[r0-1]);
//  2. a backward sweep expresses the interior's first two and last two rows
//     as affine functions of (S_{p-1}, S_p);
//  3. substituting them into the two separator equations gives a block
//     tridiagonal system in the S_p with 2x2 blocks, one block row per lane,
//     solved by block parallel cyclic reduction in ceil(log2 P) rounds;
//  4. with S_{p-1} and S_p known, a final backward sweep over the interior
//     gives the solution.
//
// This is the capability of the reference's hybrid Thomas-PCR
// (src/cuda/singlenode/pentad_cuda.cu:236-611), formulated differently:
// separators and 2x2 blocks instead of per-chunk 4x4 blocks, so the reduced
// system is half the size and needs no 4x4 Gauss-Jordan.  Out-of-range
// couplings (ds, dl of the first rows, du, dw of the last rows) are never
// read into the arithmetic, matching the reference's don't-care convention.
//
// The exchange between the lanes of one system is abstracted by Ex:
//   Ex::template swapc<K>(const T *send, T *from_prev, T *from_next, int h)
// delivers the K values of lane p-h and of lane p+h; where that chunk does
// not exist the value is a finite stand-in, and every use multiplies it by
// an exact zero: the first chunk's block A and spikes vanish (it has no
// previous separator), the last chunk's block C and separator couplings
// past the end vanish, and block PCR keeps A = 0 for p < h and C = 0 for
// p + h >= P in every round.  Wavefront DPP/shuffles implement it for x,
// LDS for y/z.
#pragma once

#include "pentad_device.hpp"

namespace pentad {

// 2x2 helpers (row-major a00 a01 a10 a11)
template <typename T> struct M2 {
  T a00, a01, a10, a11;
};
template <typename T> __device__ __forceinline__ M2<T> mul(const M2<T> &x, const M2<T> &y) {
  return {x.a00 * y.a00 + x.a01 * y.a10, x.a00 * y.a01 + x.a01 * y.a11,
          x.a10 * y.a00 + x.a11 * y.a10, x.a10 * y.a01 + x.a11 * y.a11};
}
template <typename T> __device__ __forceinline__ M2<T> inv(const M2<T> &b) {
  const T r = recip(b.a00 * b.a11 - b.a01 * b.a10);
  return {b.a11 * r, -b.a01 * r, -b.a10 * r, b.a00 * r};
}

// Truncated block PCR.  Before each round every lane holds a normalised
// block row S_p + A S_{p-h} + C S_{p+h} = r.  Once every entry of A and C
// (over all lanes taking part in the exchange) is at most tau, the remaining
// rounds change S_p = r - A S_{p-h} - C S_{p+h} by at most 4 tau max|S|:
// tau = 2^-28 (FP32) / 2^-57 (FP64) keeps that below a quarter of the unit
// roundoff, so the rounds are skipped.  In a diagonally dominant system the
// couplings between separators shrink geometrically with their distance,
// and every round doubles the distance (squares the couplings): a system
// of 64 chunks of 8 rows with dominance 2..3 stops after 2 of its 6 rounds
// in FP32 and after 3 in FP64.  The test is uniform over the exchange group,
// so the loops around the exchanges stay uniform.
template <typename T> __device__ __forceinline__ T pcr_drop_tol();
template <> __device__ __forceinline__ float pcr_drop_tol<float>() { return 0x1p-28f; }
template <> __device__ __forceinline__ double pcr_drop_tol<double>() { return 0x1p-57; }

template <typename T>
__device__ __forceinline__ bool pcr_negligible(const M2<T> &A, const M2<T> &C) {
  const T a = fmax(fmax(fabs(A.a00), fabs(A.a01)), fmax(fabs(A.a10), fabs(A.a11)));
  const T c = fmax(fmax(fabs(C.a00), fabs(C.a01)), fmax(fabs(C.a10), fabs(C.a11)));
  return fmax(a, c) <= pcr_drop_tol<T>();
}

// Block parallel cyclic reduction of a block tridiagonal system with 2x2
// blocks, one normalised block row S_p + A S_{p-1} + C S_{p+1} = r per lane
// (A = 0 on the first lane, C = 0 on the last), over the P lanes that Ex
// groups; on return (r0, r1) = S_p.  Round h = 1 is peeled so that its
// exchange is a compile-time lane shift; later rounds stop once the
// couplings are negligible on every lane (pcr_negligible).
// EXT: every block row also carries a 2x2 right-hand side E for an unknown
// pair X from outside the group (S_p + A S_{p-1} + C S_{p+1} = r + E X);
// the rounds reduce E as they reduce r, so on return S_p = r + E X.
template <typename T, bool EXT, typename Ex>
__device__ __forceinline__ void block_pcr_impl(M2<T> &A, M2<T> &C, T &r0, T &r1, M2<T> &E,
                                               int P, Ex &ex) {
  auto round = [&](int h) {
    // EXT: E - A Ep - C En first, through an exchange of its own, so that
    // the neighbours' E and A, C, r are not in registers together
    M2<T> Eq{};
    if constexpr (EXT) {
      T se[4] = {E.a00, E.a01, E.a10, E.a11}, ep[4], en[4];
      ex.template swapc<4>(se, ep, en, h);
      const M2<T> AEp = mul(A, M2<T>{ep[0], ep[1], ep[2], ep[3]});
      const M2<T> CEn = mul(C, M2<T>{en[0], en[1], en[2], en[3]});
      Eq = M2<T>{E.a00 - AEp.a00 - CEn.a00, E.a01 - AEp.a01 - CEn.a01,
                 E.a10 - AEp.a10 - CEn.a10, E.a11 - AEp.a11 - CEn.a11};
    }
    T send[10], rp[10], rn[10];
    send[0] = A.a00; send[1] = A.a01; send[2] = A.a10; send[3] = A.a11;
    send[4] = C.a00; send[5] = C.a01; send[6] = C.a10; send[7] = C.a11;
    send[8] = r0;    send[9] = r1;
    ex.template swapc<10>(send, rp, rn, h);
    const M2<T> Ap{rp[0], rp[1], rp[2], rp[3]}, Cp{rp[4], rp[5], rp[6], rp[7]};
    const M2<T> An{rn[0], rn[1], rn[2], rn[3]}, Cn{rn[4], rn[5], rn[6], rn[7]};
    // B' = I - A Cp - C An ; A' = -A Ap ; C' = -C Cn ; r' = r - A rp - C rn
    const M2<T> ACp = mul(A, Cp), CAn = mul(C, An);
    M2<T> Bn{T(1) - ACp.a00 - CAn.a00, -ACp.a01 - CAn.a01,
             -ACp.a10 - CAn.a10, T(1) - ACp.a11 - CAn.a11};
    const T q0 = r0 - (A.a00 * rp[8] + A.a01 * rp[9]) -
                 (C.a00 * rn[8] + C.a01 * rn[9]);
    const T q1 = r1 - (A.a10 * rp[8] + A.a11 * rp[9]) -
                 (C.a10 * rn[8] + C.a11 * rn[9]);
    M2<T> An2 = mul(A, Ap), Cn2 = mul(C, Cn);
    const M2<T> Bi = inv(Bn);
    if constexpr (EXT) E = mul(Bi, Eq); // normalised like r
    const M2<T> nA{-An2.a00, -An2.a01, -An2.a10, -An2.a11};
    const M2<T> nC{-Cn2.a00, -Cn2.a01, -Cn2.a10, -Cn2.a11};
    A = mul(Bi, nA);
    C = mul(Bi, nC);
    r0 = Bi.a00 * q0 + Bi.a01 * q1;
    r1 = Bi.a10 * q0 + Bi.a11 * q1;
  };
  if (P > 1) round(1);
  for (int h = 2; h < P; h <<= 1) {
    if (ex.all(pcr_negligible(A, C))) break;
    round(h);
  }
}

template <typename T, typename Ex>
__device__ __forceinline__ void block_pcr(M2<T> &A, M2<T> &C, T &r0, T &r1, int P, Ex &ex) {
  M2<T> none{};
  block_pcr_impl<T, false>(A, C, r0, r1, none, P, ex);
}

// Steps 1 and 2 of the solve below for one chunk: the interior eliminated
// with three right-hand sides (du2, dw2, zf, za, zb: row i reads
// x_i + du2_i x_{i+1} + dw2_i x_{i+2} = zf_i + za_i s0 + zb_i s1), then its
// last two rows (XL2, XL1) and first two rows (XF0, XF1) as affine functions
// [f, s0, s1, t0, t1] of the previous separator (s0, s1) and the chunk's own
// separator (t0, t1).  first: the chunk starts the system (no previous
// separator).
template <typename T, int M>
__device__ __forceinline__ void chunk_eliminate(const T (&s)[M], const T (&l)[M],
                                                const T (&d)[M], const T (&u)[M],
                                                const T (&w)[M], const T (&f)[M], bool first,
                                                T (&du2)[M - 2], T (&dw2)[M - 2],
                                                T (&zf)[M - 2], T (&za)[M - 2],
                                                T (&zb)[M - 2], T (&XL1)[5], T (&XL2)[5],
                                                T (&XF0)[5], T (&XF1)[5]) {
  constexpr int m = M - 2;
  // ---- 1. forward elimination of the interior, 3 right-hand sides --------
  // After it, x_i + du2_i x_{i+1} + dw2_i x_{i+2} = zf_i + za_i s0 + zb_i s1
  // where (s0, s1) = S_{p-1}.
  {
    const T iv = recip(d[0]);
    du2[0] = u[0] * iv;
    dw2[0] = w[0] * iv;
    zf[0] = f[0] * iv;
    za[0] = first ? T(0) : -s[0] * iv;
    zb[0] = first ? T(0) : -l[0] * iv;
  }
  {
    const T ddl = l[1];
    const T iv = recip(d[1] - ddl * du2[0]);
    du2[1] = (u[1] - ddl * dw2[0]) * iv;
    dw2[1] = w[1] * iv;
    zf[1] = (f[1] - ddl * zf[0]) * iv;
    za[1] = (-ddl * za[0]) * iv;
    zb[1] = ((first ? T(0) : -s[1]) - ddl * zb[0]) * iv;
  }
#pragma unroll
  for (int i = 2; i < m; ++i) {
    const T ddl = l[i] - s[i] * du2[i - 2];
    const T iv = recip(d[i] - s[i] * dw2[i - 2] - ddl * du2[i - 1]);
    du2[i] = (u[i] - ddl * dw2[i - 1]) * iv;
    dw2[i] = w[i] * iv;
    zf[i] = (f[i] - s[i] * zf[i - 2] - ddl * zf[i - 1]) * iv;
    za[i] = (-s[i] * za[i - 2] - ddl * za[i - 1]) * iv;
    zb[i] = (-s[i] * zb[i - 2] - ddl * zb[i - 1]) * iv;
  }

  // ---- 2. backward sweep with 5 components [f, s0, s1, t0, t1] -----------
  // (t0, t1) = S_p = (x_m, x_{m+1}).  Keep rows m-2, m-1 (XL2, XL1) and
  // rows 0, 1 (XF0, XF1).
  {
    const int i = m - 1;
    XL1[0] = zf[i];
    XL1[1] = za[i];
    XL1[2] = zb[i];
    XL1[3] = -du2[i];
    XL1[4] = -dw2[i];
  }
  {
    const int i = m - 2;
    XL2[0] = zf[i] - du2[i] * XL1[0];
    XL2[1] = za[i] - du2[i] * XL1[1];
    XL2[2] = zb[i] - du2[i] * XL1[2];
    XL2[3] = -du2[i] * XL1[3] - dw2[i];
    XL2[4] = -du2[i] * XL1[4];
  }
  {
    T xa[5], xb[5]; // rows i+1, i+2
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      xa[c] = XL2[c];
      xb[c] = XL1[c];
    }
#pragma unroll
    for (int i = m - 3; i >= 0; --i) {
      const T z3[3] = {zf[i], za[i], zb[i]};
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        const T zc = (c < 3) ? z3[c] : T(0);
        const T xi = zc - du2[i] * xa[c] - dw2[i] * xb[c];
        xb[c] = xa[c];
        xa[c] = xi;
      }
    }
    // now xa = row 0, xb = row 1 (for m == 2 these are XL2, XL1)
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      XF0[c] = xa[c];
      XF1[c] = xb[c];
    }
  }
}

// Step 4 for one chunk: with the previous separator (s0, s1) and the
// chunk's own (t0, t1) known, the backward sweep over the interior; f
// receives the solution of all M rows.
template <typename T, int M>
__device__ __forceinline__ void chunk_backsub(const T (&du2)[M - 2], const T (&dw2)[M - 2],
                                              const T (&zf)[M - 2], const T (&za)[M - 2],
                                              const T (&zb)[M - 2], T s0, T s1, T t0, T t1,
                                              T (&f)[M]) {
  constexpr int m = M - 2;
  T xa = t0, xb = t1; // x_{i+1}, x_{i+2}
#pragma unroll
  for (int i = m - 1; i >= 0; --i) {
    const T hi = zf[i] + za[i] * s0 + zb[i] * s1;
    const T xi = hi - du2[i] * xa - dw2[i] * xb;
    f[i] = xi;
    xb = xa;
    xa = xi;
  }
  f[m] = t0;
  f[m + 1] = t1;
}

// Step 3 for one chunk: its separator equations (rows m and m+1) as the
// block row A S_{p-1} + B S_p + C S_{p+1} = r, with the next chunk's first
// two rows Y0, Y1 (affine functions [f, s0, s1, t0, t1] of S_p and S_{p+1})
// substituted where rows m, m+1 reach past the chunk: wa = w_m, ub =
// u_{m+1}, wb = w_{m+1}, or zeros where those couplings do not go to the
// next chunk of the group.
template <typename T, int M>
__device__ __forceinline__ void separator_block_row(const T (&s)[M], const T (&l)[M],
                                                    const T (&d)[M], const T (&u)[M],
                                                    const T (&f)[M], const T (&XL1)[5],
                                                    const T (&XL2)[5], const T (&Y0)[5],
                                                    const T (&Y1)[5], T wa, T ub, T wb,
                                                    M2<T> &A, M2<T> &B, M2<T> &C, T &r0,
                                                    T &r1) {
  constexpr int m = M - 2;
  const T sa = s[m], la = l[m], da = d[m], ua = u[m];
  const T sb = s[m + 1], lb = l[m + 1], db = d[m + 1];
  // eq a: sa x_{m-2} + la x_{m-1} + da t0 + ua t1 + wa y0 = f_m
  A.a00 = sa * XL2[1] + la * XL1[1];
  A.a01 = sa * XL2[2] + la * XL1[2];
  B.a00 = sa * XL2[3] + la * XL1[3] + da + wa * Y0[1];
  B.a01 = sa * XL2[4] + la * XL1[4] + ua + wa * Y0[2];
  C.a00 = wa * Y0[3];
  C.a01 = wa * Y0[4];
  r0 = f[m] - sa * XL2[0] - la * XL1[0] - wa * Y0[0];
  // eq b: sb x_{m-1} + lb t0 + db t1 + ub y0 + wb y1 = f_{m+1}
  A.a10 = sb * XL1[1];
  A.a11 = sb * XL1[2];
  B.a10 = sb * XL1[3] + lb + ub * Y0[1] + wb * Y1[1];
  B.a11 = sb * XL1[4] + db + ub * Y0[2] + wb * Y1[2];
  C.a10 = ub * Y0[3] + wb * Y1[3];
  C.a11 = ub * Y0[4] + wb * Y1[4];
  r1 = f[m + 1] - sb * XL1[0] - ub * Y0[0] - wb * Y1[0];
}

// Solve the chunk owned by this lane.  On entry s,l,d,u,w,f hold rows
// r0..r0+M-1 of the five diagonals and the RHS; on return f holds the
// solution of those rows.  p is the chunk index, P the chunk count.
template <typename T, int M, typename Ex>
__device__ __forceinline__ void part_solve(const T (&s)[M], const T (&l)[M],
                                           const T (&d)[M], const T (&u)[M],
                                           const T (&w)[M], T (&f)[M], int p,
                                           int P, Ex &ex) {
  static_assert(M >= 4, "a chunk needs at least two interior rows");
  constexpr int m = M - 2; // interior rows 0..m-1, separator rows m, m+1
  const bool first = (p == 0);
  const bool last = (p == P - 1);

  T du2[m], dw2[m], zf[m], za[m], zb[m];
  T XL1[5], XL2[5], XF0[5], XF1[5];
  chunk_eliminate<T, M>(s, l, d, u, w, f, first, du2, dw2, zf, za, zb, XL1, XL2, XF0, XF1);

  // ---- 3. separator equations -> one 2x2 block row per lane --------------
  T Y0[5], Y1[5]; // next chunk's interior rows 0, 1 (zeros for the last)
  {
    T send[10], rp[10], rn[10];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      send[c] = XF0[c];
      send[5 + c] = XF1[c];
    }
    ex.template swapc<10>(send, rp, rn, 1);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      Y0[c] = rn[c];
      Y1[c] = rn[5 + c];
    }
  }
  // separator rows m (eq a) and m+1 (eq b); couplings past the end of the
  // system are out of range for the last chunk.
  M2<T> A, B, C;
  T r0, r1;
  separator_block_row<T, M>(s, l, d, u, f, XL1, XL2, Y0, Y1, last ? T(0) : w[m],
                            last ? T(0) : u[m + 1], last ? T(0) : w[m + 1], A, B, C, r0, r1);

  // normalise: S_p + A S_{p-1} + C S_{p+1} = r
  {
    const M2<T> Bi = inv(B);
    A = mul(Bi, A);
    C = mul(Bi, C);
    const T t0 = Bi.a00 * r0 + Bi.a01 * r1;
    const T t1 = Bi.a10 * r0 + Bi.a11 * r1;
    r0 = t0;
    r1 = t1;
  }

  // ---- block parallel cyclic reduction -----------------------------------
  block_pcr<T>(A, C, r0, r1, P, ex);
  const T t0 = r0, t1 = r1; // S_p

  // ---- 4. previous separator, final backward sweep -----------------------
  T s0, s1;
  {
    T send[2] = {t0, t1}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    s0 = rp[0];
    s1 = rp[1];
  }
  chunk_backsub<T, M>(du2, dw2, zf, za, zb, s0, s1, t0, t1, f);
}

} // namespace pentad
