// shared_core.hpp -- shared factorisation: factor one system, apply to many.
//
// When every system of a direction has the same coefficients (constant-
// coefficient operators, the reference's "shared-fact" mode,
// src/cuda/singlenode/pentad_cuda.cu:908-1195), the partitioned
// factorisation of partition_core.hpp is computed once, from system 0, into
// a table; the solve then streams only the right-hand side and the
// solution: 2 words of HBM traffic per row instead of 7.
//
// Table layout: value k of chunk p at tab[k * P + p] (lanes of one system
// read consecutive words).  Per chunk, with m = M - 2 interior rows:
//   IV  1/pivot            SIV  s * 1/pivot       LIV  (eliminated l) * 1/pivot
//   DU2, DW2               the U factor (unit diagonal)
//   ZA, ZB                 spikes of the previous separator
//   SEP                    separator couplings sa la wa sb ub wb
//   BI0                    inverse of the chunk's 2x2 separator block
//   PCR + 12 j             round j of block PCR: A (4), C (4), B^-1 (4)
// As in the reference, nothing checks that the systems really share
// coefficients: system 0's are used for all (pentad_cuda.cu:962-963).
#pragma once

#include "partition_core.hpp"
#include "partition_exchange.hpp"

namespace pentad {

template <int M> struct SfLayout {
  static constexpr int m = M - 2;
  static constexpr int IV = 0, SIV = m, LIV = 2 * m, DU2 = 3 * m, DW2 = 4 * m;
  static constexpr int ZA = 5 * m, ZB = 6 * m, SEP = 7 * m, BI0 = 7 * m + 6;
  static constexpr int PCR = 7 * m + 10;
  static constexpr int K(int rounds) { return PCR + 12 * rounds; }
};

inline int sf_rounds(int P) {
  int r = 0;
  while ((1 << r) < P) ++r;
  return r;
}

// Factor chunk p of system 0 (coefficient rows in registers, out-of-range
// couplings already zero) and write its table column.
template <typename T, int M, typename Ex>
__device__ __forceinline__ void sf_factor(const T (&s)[M], const T (&l)[M],
                                          const T (&d)[M], const T (&u)[M],
                                          const T (&w)[M], int p, int P, Ex &ex,
                                          T *__restrict__ tab) {
  using L = SfLayout<M>;
  constexpr int m = M - 2;
  auto put = [&](int k, T v) { tab[(int64_t)k * P + p] = v; };
  T du2[m], dw2[m], za[m], zb[m];
  {
    const T iv = recip(d[0]);
    du2[0] = u[0] * iv;
    dw2[0] = w[0] * iv;
    za[0] = -s[0] * iv;
    zb[0] = -l[0] * iv;
    put(L::IV, iv);
    put(L::SIV, T(0));
    put(L::LIV, T(0));
  }
  {
    const T ddl = l[1];
    const T iv = recip(d[1] - ddl * du2[0]);
    du2[1] = (u[1] - ddl * dw2[0]) * iv;
    dw2[1] = w[1] * iv;
    za[1] = (-ddl * za[0]) * iv;
    zb[1] = (-s[1] - ddl * zb[0]) * iv;
    put(L::IV + 1, iv);
    put(L::SIV + 1, T(0));
    put(L::LIV + 1, ddl * iv);
  }
#pragma unroll
  for (int i = 2; i < m; ++i) {
    const T ddl = l[i] - s[i] * du2[i - 2];
    const T iv = recip(d[i] - s[i] * dw2[i - 2] - ddl * du2[i - 1]);
    du2[i] = (u[i] - ddl * dw2[i - 1]) * iv;
    dw2[i] = w[i] * iv;
    za[i] = (-s[i] * za[i - 2] - ddl * za[i - 1]) * iv;
    zb[i] = (-s[i] * zb[i - 2] - ddl * zb[i - 1]) * iv;
    put(L::IV + i, iv);
    put(L::SIV + i, s[i] * iv);
    put(L::LIV + i, ddl * iv);
  }
#pragma unroll
  for (int i = 0; i < m; ++i) {
    put(L::DU2 + i, du2[i]);
    put(L::DW2 + i, dw2[i]);
    put(L::ZA + i, za[i]);
    put(L::ZB + i, zb[i]);
  }
  // backward sweep of the spike components [s0, s1, t0, t1]
  T XL1[4], XL2[4], xa[4], xb[4];
  XL1[0] = za[m - 1];
  XL1[1] = zb[m - 1];
  XL1[2] = -du2[m - 1];
  XL1[3] = -dw2[m - 1];
  XL2[0] = za[m - 2] - du2[m - 2] * XL1[0];
  XL2[1] = zb[m - 2] - du2[m - 2] * XL1[1];
  XL2[2] = -du2[m - 2] * XL1[2] - dw2[m - 2];
  XL2[3] = -du2[m - 2] * XL1[3];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    xa[c] = XL2[c];
    xb[c] = XL1[c];
  }
#pragma unroll
  for (int i = m - 3; i >= 0; --i) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const T zc = (c == 0) ? za[i] : (c == 1 ? zb[i] : T(0));
      const T xi = zc - du2[i] * xa[c] - dw2[i] * xb[c];
      xb[c] = xa[c];
      xa[c] = xi;
    }
  }
  T Y0[4], Y1[4];
  {
    T send[8], rp[8], rn[8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      send[c] = xa[c];
      send[4 + c] = xb[c];
    }
    ex.template swapc<8>(send, rp, rn, 1);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      Y0[c] = rn[c];
      Y1[c] = rn[4 + c];
    }
  }
  // separator couplings past the end of the system are zero in w, u
  const T sa = s[m], la = l[m], wa = w[m], sb = s[m + 1], ub = u[m + 1], wb = w[m + 1];
  put(L::SEP + 0, sa);
  put(L::SEP + 1, la);
  put(L::SEP + 2, wa);
  put(L::SEP + 3, sb);
  put(L::SEP + 4, ub);
  put(L::SEP + 5, wb);
  M2<T> Am, B, C;
  Am.a00 = sa * XL2[0] + la * XL1[0];
  Am.a01 = sa * XL2[1] + la * XL1[1];
  B.a00 = sa * XL2[2] + la * XL1[2] + d[m] + wa * Y0[0];
  B.a01 = sa * XL2[3] + la * XL1[3] + u[m] + wa * Y0[1];
  C.a00 = wa * Y0[2];
  C.a01 = wa * Y0[3];
  Am.a10 = sb * XL1[0];
  Am.a11 = sb * XL1[1];
  B.a10 = sb * XL1[2] + l[m + 1] + ub * Y0[0] + wb * Y1[0];
  B.a11 = sb * XL1[3] + d[m + 1] + ub * Y0[1] + wb * Y1[1];
  C.a10 = ub * Y0[2] + wb * Y1[2];
  C.a11 = ub * Y0[3] + wb * Y1[3];
  const M2<T> Bi0 = inv(B);
  put(L::BI0 + 0, Bi0.a00);
  put(L::BI0 + 1, Bi0.a01);
  put(L::BI0 + 2, Bi0.a10);
  put(L::BI0 + 3, Bi0.a11);
  Am = mul(Bi0, Am);
  C = mul(Bi0, C);
  int j = 0;
  for (int h = 1; h < P; h <<= 1, ++j) {
    T send[8] = {Am.a00, Am.a01, Am.a10, Am.a11, C.a00, C.a01, C.a10, C.a11};
    T rp[8], rn[8];
    ex.template swapc<8>(send, rp, rn, h);
    const M2<T> Ap{rp[0], rp[1], rp[2], rp[3]}, Cp{rp[4], rp[5], rp[6], rp[7]};
    const M2<T> An{rn[0], rn[1], rn[2], rn[3]}, Cn{rn[4], rn[5], rn[6], rn[7]};
    const int o = L::PCR + 12 * j;
    put(o + 0, Am.a00); put(o + 1, Am.a01); put(o + 2, Am.a10); put(o + 3, Am.a11);
    put(o + 4, C.a00);  put(o + 5, C.a01);  put(o + 6, C.a10);  put(o + 7, C.a11);
    const M2<T> ACp = mul(Am, Cp), CAn = mul(C, An);
    const M2<T> Bn{T(1) - ACp.a00 - CAn.a00, -ACp.a01 - CAn.a01,
                   -ACp.a10 - CAn.a10, T(1) - ACp.a11 - CAn.a11};
    const M2<T> Bi = inv(Bn);
    put(o + 8, Bi.a00); put(o + 9, Bi.a01); put(o + 10, Bi.a10); put(o + 11, Bi.a11);
    const M2<T> AA = mul(Am, Ap), CC = mul(C, Cn);
    Am = mul(Bi, M2<T>{-AA.a00, -AA.a01, -AA.a10, -AA.a11});
    C = mul(Bi, M2<T>{-CC.a00, -CC.a01, -CC.a10, -CC.a11});
  }
}

// Solve with the table: f (rows of chunk p) is overwritten by the solution.
// tab points at this chunk's column (value k at tab[k * P]).
template <typename T, int M, typename Ex>
__device__ __forceinline__ void sf_apply(const T *tab, int P, T (&f)[M], Ex &ex) {
  using L = SfLayout<M>;
  constexpr int m = M - 2;
  auto at = [&](int k) { return tab[k * P]; };
  T zf[m];
  zf[0] = f[0] * at(L::IV);
  zf[1] = f[1] * at(L::IV + 1) - at(L::LIV + 1) * zf[0];
#pragma unroll
  for (int i = 2; i < m; ++i)
    zf[i] = f[i] * at(L::IV + i) - at(L::SIV + i) * zf[i - 2] - at(L::LIV + i) * zf[i - 1];
  const T xl1 = zf[m - 1];
  const T xl2 = zf[m - 2] - at(L::DU2 + m - 2) * xl1;
  T xa = xl2, xb = xl1;
#pragma unroll
  for (int i = m - 3; i >= 0; --i) {
    const T xi = zf[i] - at(L::DU2 + i) * xa - at(L::DW2 + i) * xb;
    xb = xa;
    xa = xi;
  }
  T y0, y1;
  {
    T send[2] = {xa, xb}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    y0 = rn[0];
    y1 = rn[1];
  }
  const T r0 = f[m] - at(L::SEP + 0) * xl2 - at(L::SEP + 1) * xl1 - at(L::SEP + 2) * y0;
  const T r1 = f[m + 1] - at(L::SEP + 3) * xl1 - at(L::SEP + 4) * y0 - at(L::SEP + 5) * y1;
  T q0 = at(L::BI0 + 0) * r0 + at(L::BI0 + 1) * r1;
  T q1 = at(L::BI0 + 2) * r0 + at(L::BI0 + 3) * r1;
  auto round = [&](int h, int j) {
    T send[2] = {q0, q1}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, h);
    const int o = L::PCR + 12 * j;
    const T t0 = q0 - (at(o + 0) * rp[0] + at(o + 1) * rp[1]) - (at(o + 4) * rn[0] + at(o + 5) * rn[1]);
    const T t1 = q1 - (at(o + 2) * rp[0] + at(o + 3) * rp[1]) - (at(o + 6) * rn[0] + at(o + 7) * rn[1]);
    q0 = at(o + 8) * t0 + at(o + 9) * t1;
    q1 = at(o + 10) * t0 + at(o + 11) * t1;
  };
  if (P > 1) round(1, 0);
  {
    int j = 1;
    for (int h = 2; h < P; h <<= 1, ++j) round(h, j);
  }
  T s0, s1;
  {
    T send[2] = {q0, q1}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    s0 = rp[0];
    s1 = rp[1];
  }
  T xn1 = q0, xn2 = q1;
#pragma unroll
  for (int i = m - 1; i >= 0; --i) {
    const T hi = zf[i] + at(L::ZA + i) * s0 + at(L::ZB + i) * s1;
    const T xi = hi - at(L::DU2 + i) * xn1 - at(L::DW2 + i) * xn2;
    f[i] = xi;
    xn2 = xn1;
    xn1 = xi;
  }
  f[m] = q0;
  f[m + 1] = q1;
}

} // namespace pentad
