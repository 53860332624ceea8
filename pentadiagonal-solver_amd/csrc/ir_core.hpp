// ir_core.hpp -- FP32 solve + FP64 iterative refinement, fused per line.
//
// Mixed-precision iterative refinement on the partitioned layout of
// partition_core.hpp (same chunks, separators and 2x2 block reduced system):
//
//   factor A (FP32) once: interior LU, separator spikes, and the block
//                         PCR multipliers of every round that is not
//                         negligible (truncated PCR, partition_core.hpp)
//   x0  = solve32(A, fp32(b))
//   repeat:  r  = b - A x      (FP64, A widened exactly from FP32)
//            dx = solve32(A, fp32(r))
//            x += dx           (FP64)
//   until the estimated relative forward error of x,
//   max|dx_k|^2 / (max|dx_{k-1}| max|x|) with dx_0 = x0, is <= tol
//   (maxima over the system), or max_steps corrections were made.
//
// The residual and the update are the only FP64 arithmetic.  Coefficients
// are read once from HBM in FP32 (20 B/row) and the RHS/solution in FP64
// (8 + 8 B/row); every refinement step runs on chip.  The FP64 result solves
// widen(A32) x = b, the system the FP64 oracle solves.
//
// The residual's operator comes from the accessor's FP64 values (s64 ..
// w64): the FP32 coefficients widened for stored coefficients, or an FP64
// operator of its own (structured coefficients generated from an FP64 scale,
// whose solution the refinement then converges to).
//
// Register budget: the coefficients are reached through an accessor (Cf),
// which the x-direction kernel backs with LDS so that the per-lane state is
// the FP64 iterate, the LU factors (du2, 1/pivot, two spikes) and the PCR
// multipliers; the other two factors (the eliminated lower coupling and
// du2's partner dw2) are recomputed from the coefficients when used.  The
// right-hand side is reached through Bv (registers, or re-read from memory).
//
// No reference counterpart (the reference solves in one precision only).
#pragma once

#include "partition_core.hpp"
#include "partition_exchange.hpp"

namespace pentad {

// coefficient accessors: s(i), l(i), d(i), u(i), w(i) for chunk row i
template <int M> struct RegCoef {
  const float (&s_)[M], (&l_)[M], (&d_)[M], (&u_)[M], (&w_)[M];
  __device__ __forceinline__ float s(int i) const { return s_[i]; }
  __device__ __forceinline__ float l(int i) const { return l_[i]; }
  __device__ __forceinline__ float d(int i) const { return d_[i]; }
  __device__ __forceinline__ float u(int i) const { return u_[i]; }
  __device__ __forceinline__ float w(int i) const { return w_[i]; }
  // the operator of the FP64 residual: the FP32 coefficients widened
  __device__ __forceinline__ double s64(int i) const { return (double)s_[i]; }
  __device__ __forceinline__ double l64(int i) const { return (double)l_[i]; }
  __device__ __forceinline__ double d64(int i) const { return (double)d_[i]; }
  __device__ __forceinline__ double u64(int i) const { return (double)u_[i]; }
  __device__ __forceinline__ double w64(int i) const { return (double)w_[i]; }
  __device__ __forceinline__ void refresh() const {}
};

// LDS layout [5][M][64] per wavefront: lane-minor, conflict-free.
template <int M> struct LdsCoef {
  const float *base; // this wavefront's region + lane
  __device__ __forceinline__ float at(int k, int i) const { return base[(k * M + i) * 64]; }
  __device__ __forceinline__ float s(int i) const { return at(0, i); }
  __device__ __forceinline__ float l(int i) const { return at(1, i); }
  __device__ __forceinline__ float d(int i) const { return at(2, i); }
  __device__ __forceinline__ float u(int i) const { return at(3, i); }
  __device__ __forceinline__ float w(int i) const { return at(4, i); }
  __device__ __forceinline__ double s64(int i) const { return (double)s(i); }
  __device__ __forceinline__ double l64(int i) const { return (double)l(i); }
  __device__ __forceinline__ double d64(int i) const { return (double)d(i); }
  __device__ __forceinline__ double u64(int i) const { return (double)u(i); }
  __device__ __forceinline__ double w64(int i) const { return (double)w(i); }
  // Forget earlier loads (compiler-level memory clobber, no instruction):
  // without it the compiler keeps every coefficient it has read in
  // registers, which is what keeping them in LDS is meant to avoid.
  __device__ __forceinline__ void refresh() const { asm volatile("" ::: "memory"); }
};

// right-hand side accessors
template <int M> struct RegRhs {
  const double (&b_)[M];
  __device__ __forceinline__ double operator()(int i) const { return b_[i]; }
};
struct MemRhs { // rows i of this lane's chunk: p[i * stride]
  const double *p;
  int64_t stride;
  __device__ __forceinline__ double operator()(int i) const { return p[i * stride]; }
};

// Zero the couplings of the chunk's rows that reach past the ends of the
// system (diagonal k: 0 ds, 1 dl, 3 du, 4 dw), so that products with halo
// values of chunks that do not exist vanish without selects.
template <int M>
__device__ __forceinline__ void ir_zero_out_of_range(int k, float (&t)[M], bool first,
                                                     bool last) {
  if (k == 0 && first) t[0] = t[1] = 0.0f;
  if (k == 1 && first) t[0] = 0.0f;
  if (k == 3 && last) t[M - 1] = 0.0f;
  if (k == 4 && last) t[M - 2] = t[M - 1] = 0.0f;
}

// Factorisation of one chunk: du2, 1/pivot, the previous-separator spikes
// (za, zb) and the block PCR multipliers (Ah, Ch, Bi) of up to MAXR rounds.
// Rounds are truncated once their couplings are negligible (pcr_negligible);
// a system that still needs rounds after the MAXR stored ones keeps the
// normalised block row reached there (A0, C0) and every apply redoes those
// rounds on the fly (more = true).  12 registers per stored round: MAXR = 3
// covers diagonally dominant systems of up to 64 chunks of 8 rows (they
// stop after 2 or 3 rounds) without pushing the kernels off 3 waves per
// SIMD; MAXR = 0 (long systems, hundreds of chunks) redoes every round.
// The separator rows' couplings are read from the coefficients when used
// (the last chunk's out-of-range ones are zero there).
// SPIKES = false drops za, zb (2 (M-2) registers): the final backward sweep
// of every apply then redoes the forward sweep with the previous separator's
// couplings moved into the first two rows' right-hand side, which by
// linearity is zf + za s0 + zb s1.
template <int M, int MAXR, bool SPIKES = true> struct IrFactors {
  static constexpr int m = M - 2;
  static constexpr int NR = MAXR > 0 ? MAXR : 1;
  static constexpr int NS = SPIKES ? m : 1;
  float du2[m], iv[m], za[NS], zb[NS];
  int nr;    // stored rounds in use
  bool more; // rounds after the stored ones are needed (uniform over the system)
  M2<float> Bi0;
  M2<float> Ah[NR], Ch[NR], Bi[NR];
  M2<float> A0, C0; // block row after the stored rounds (more == true)
};

// Largest number of stored PCR rounds the kernels use (see IrFactors).
constexpr int ir_stored_rounds(int R) { return R < 3 ? R : 3; }

// The FP32 factorisation takes the hardware reciprocal as it is (within an
// ulp or two): the refinement measures and removes the error of the FP32
// solve however it arose, so a Newton step here only costs latency.
__device__ __forceinline__ float rcp32(float d) { return __builtin_amdgcn_rcpf(d); }
__device__ __forceinline__ M2<float> inv32(const M2<float> &b) {
  const float r = rcp32(b.a00 * b.a11 - b.a01 * b.a10);
  return {b.a11 * r, -b.a01 * r, -b.a10 * r, b.a00 * r};
}

// eliminated lower coupling of interior row i (ddl) and dw2 = w / pivot
template <int M, int MAXR, bool SP, typename Cf>
__device__ __forceinline__ float ir_ddl(const IrFactors<M, MAXR, SP> &F, const Cf &A, int i) {
  return (i == 1) ? A.l(1) : A.l(i) - A.s(i) * F.du2[i - 2];
}
template <int M, int MAXR, bool SP, typename Cf>
__device__ __forceinline__ float ir_dw2(const IrFactors<M, MAXR, SP> &F, const Cf &A, int i) {
  return A.w(i) * F.iv[i];
}

template <int M, int MAXR, bool SP, typename Cf, typename Ex>
__device__ __forceinline__ void ir_factor(IrFactors<M, MAXR, SP> &F, const Cf &A, int p,
                                          int P, Ex &ex) {
  constexpr int m = M - 2;
  const bool first = (p == 0), last = (p == P - 1);
  float dw2[m]; // kept only during the factorisation
  float za[m], zb[m]; // the spikes, copied into F when it keeps them
  F.iv[0] = rcp32(A.d(0));
  F.du2[0] = A.u(0) * F.iv[0];
  dw2[0] = A.w(0) * F.iv[0];
  za[0] = first ? 0.0f : -A.s(0) * F.iv[0];
  zb[0] = first ? 0.0f : -A.l(0) * F.iv[0];
  {
    const float l1 = A.l(1);
    F.iv[1] = rcp32(A.d(1) - l1 * F.du2[0]);
    F.du2[1] = (A.u(1) - l1 * dw2[0]) * F.iv[1];
    dw2[1] = A.w(1) * F.iv[1];
    za[1] = (-l1 * za[0]) * F.iv[1];
    zb[1] = ((first ? 0.0f : -A.s(1)) - l1 * zb[0]) * F.iv[1];
  }
#pragma unroll
  for (int i = 2; i < m; ++i) {
    const float si = A.s(i);
    const float dl_ = A.l(i) - si * F.du2[i - 2];
    F.iv[i] = rcp32(A.d(i) - si * dw2[i - 2] - dl_ * F.du2[i - 1]);
    F.du2[i] = (A.u(i) - dl_ * dw2[i - 1]) * F.iv[i];
    dw2[i] = A.w(i) * F.iv[i];
    za[i] = (-si * za[i - 2] - dl_ * za[i - 1]) * F.iv[i];
    zb[i] = (-si * zb[i - 2] - dl_ * zb[i - 1]) * F.iv[i];
  }
  // backward sweep of the spike components [s0, s1, t0, t1]
  float XL1[4], XL2[4], XF0[4], XF1[4];
  XL1[0] = za[m - 1];
  XL1[1] = zb[m - 1];
  XL1[2] = -F.du2[m - 1];
  XL1[3] = -dw2[m - 1];
  XL2[0] = za[m - 2] - F.du2[m - 2] * XL1[0];
  XL2[1] = zb[m - 2] - F.du2[m - 2] * XL1[1];
  XL2[2] = -F.du2[m - 2] * XL1[2] - dw2[m - 2];
  XL2[3] = -F.du2[m - 2] * XL1[3];
  {
    float xa[4], xb[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      xa[c] = XL2[c];
      xb[c] = XL1[c];
    }
#pragma unroll
    for (int i = m - 3; i >= 0; --i) {
      const float z2[2] = {za[i], zb[i]};
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const float zc = (c < 2) ? z2[c] : 0.0f;
        const float xi = zc - F.du2[i] * xa[c] - dw2[i] * xb[c];
        xb[c] = xa[c];
        xa[c] = xi;
      }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      XF0[c] = xa[c];
      XF1[c] = xb[c];
    }
  }
  float Y0[4], Y1[4];
  {
    float send[8], rp[8], rn[8];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      send[c] = XF0[c];
      send[4 + c] = XF1[c];
    }
    ex.template swapc<8>(send, rp, rn, 1);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      Y0[c] = rn[c];
      Y1[c] = rn[4 + c];
    }
  }
  const float sa = A.s(m), la = A.l(m), wa = last ? 0.0f : A.w(m);
  const float sb = A.s(m + 1), ub = last ? 0.0f : A.u(m + 1), wb = last ? 0.0f : A.w(m + 1);
  const float da = A.d(m), ua = A.u(m), lb = A.l(m + 1), db = A.d(m + 1);
  M2<float> Am, B, C;
  Am.a00 = sa * XL2[0] + la * XL1[0];
  Am.a01 = sa * XL2[1] + la * XL1[1];
  B.a00 = sa * XL2[2] + la * XL1[2] + da + wa * Y0[0];
  B.a01 = sa * XL2[3] + la * XL1[3] + ua + wa * Y0[1];
  C.a00 = wa * Y0[2];
  C.a01 = wa * Y0[3];
  Am.a10 = sb * XL1[0];
  Am.a11 = sb * XL1[1];
  B.a10 = sb * XL1[2] + lb + ub * Y0[0] + wb * Y1[0];
  B.a11 = sb * XL1[3] + db + ub * Y0[1] + wb * Y1[1];
  C.a10 = ub * Y0[2] + wb * Y1[2];
  C.a11 = ub * Y0[3] + wb * Y1[3];
  if constexpr (SP) {
#pragma unroll
    for (int i = 0; i < m; ++i) {
      F.za[i] = za[i];
      F.zb[i] = zb[i];
    }
  }
  F.Bi0 = inv32(B);
  Am = mul(F.Bi0, Am);
  C = mul(F.Bi0, C);
  F.nr = 0;
  bool done = (P <= 1);
#pragma unroll
  for (int j = 0; j < MAXR; ++j) {
    const int h = 1 << j;
    if (!done) {
      if (h >= P || (j > 0 && ex.all(pcr_negligible(Am, C)))) {
        done = true;
      } else {
        F.nr = j + 1;
        float send[8] = {Am.a00, Am.a01, Am.a10, Am.a11, C.a00, C.a01, C.a10, C.a11};
        float rp[8], rn[8];
        ex.template swapc<8>(send, rp, rn, h);
        const M2<float> Ap{rp[0], rp[1], rp[2], rp[3]}, Cp{rp[4], rp[5], rp[6], rp[7]};
        const M2<float> An{rn[0], rn[1], rn[2], rn[3]}, Cn{rn[4], rn[5], rn[6], rn[7]};
        F.Ah[j] = Am;
        F.Ch[j] = C;
        const M2<float> ACp = mul(Am, Cp), CAn = mul(C, An);
        const M2<float> Bn{1.0f - ACp.a00 - CAn.a00, -ACp.a01 - CAn.a01,
                           -ACp.a10 - CAn.a10, 1.0f - ACp.a11 - CAn.a11};
        F.Bi[j] = inv32(Bn);
        const M2<float> AA = mul(Am, Ap), CC = mul(C, Cn);
        Am = mul(F.Bi[j], M2<float>{-AA.a00, -AA.a01, -AA.a10, -AA.a11});
        C = mul(F.Bi[j], M2<float>{-CC.a00, -CC.a01, -CC.a10, -CC.a11});
      }
    }
  }
  if (!done) done = ((1 << MAXR) >= P) || ex.all(pcr_negligible(Am, C));
  F.more = !done;
  F.A0 = Am;
  F.C0 = C;
}

// Solve A y = f with the factors; f (M rows of this chunk) is overwritten by y.
template <int M, int MAXR, bool SP, typename Cf, typename Ex>
__device__ __forceinline__ void ir_apply(const IrFactors<M, MAXR, SP> &F, const Cf &A,
                                         float (&f)[M], int p, int P, Ex &ex) {
  constexpr int m = M - 2;
  A.refresh();
  float zf[m];
  zf[0] = f[0] * F.iv[0];
  zf[1] = (f[1] - A.l(1) * zf[0]) * F.iv[1];
#pragma unroll
  for (int i = 2; i < m; ++i)
    zf[i] = (f[i] - A.s(i) * zf[i - 2] - ir_ddl(F, A, i) * zf[i - 1]) * F.iv[i];
  // f-component of the interior's first and last two rows (separators = 0)
  const float xl1 = zf[m - 1];
  const float xl2 = zf[m - 2] - F.du2[m - 2] * xl1;
  float xa = xl2, xb = xl1;
#pragma unroll
  for (int i = m - 3; i >= 0; --i) {
    const float xi = zf[i] - F.du2[i] * xa - ir_dw2(F, A, i) * xb;
    xb = xa;
    xa = xi;
  }
  float y0, y1;
  {
    float send[2] = {xa, xb}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    y0 = rn[0];
    y1 = rn[1];
  }
  // separator couplings (zero past the end of the system)
  const float r0 = f[m] - A.s(m) * xl2 - A.l(m) * xl1 - A.w(m) * y0;
  const float r1 = f[m + 1] - A.s(m + 1) * xl1 - A.u(m + 1) * y0 - A.w(m + 1) * y1;
  float q0 = F.Bi0.a00 * r0 + F.Bi0.a01 * r1;
  float q1 = F.Bi0.a10 * r0 + F.Bi0.a11 * r1;
#pragma unroll
  for (int j = 0; j < MAXR; ++j) {
    const int h = 1 << j;
    if (j < F.nr) {
      float send[2] = {q0, q1}, rp[2], rn[2];
      ex.template swapc<2>(send, rp, rn, h);
      const M2<float> &Ah = F.Ah[j], &Ch = F.Ch[j], &Bi = F.Bi[j];
      const float t0 = q0 - (Ah.a00 * rp[0] + Ah.a01 * rp[1]) - (Ch.a00 * rn[0] + Ch.a01 * rn[1]);
      const float t1 = q1 - (Ah.a10 * rp[0] + Ah.a11 * rp[1]) - (Ch.a10 * rn[0] + Ch.a11 * rn[1]);
      q0 = Bi.a00 * t0 + Bi.a01 * t1;
      q1 = Bi.a10 * t0 + Bi.a11 * t1;
    }
  }
  if (F.more) { // rounds past the stored ones: block PCR on (A, C, q) together
    M2<float> Am = F.A0, C = F.C0;
    auto round = [&](int h) {
      float send[10] = {Am.a00, Am.a01, Am.a10, Am.a11, C.a00, C.a01, C.a10, C.a11, q0, q1};
      float rp[10], rn[10];
      ex.template swapc<10>(send, rp, rn, h);
      const M2<float> Ap{rp[0], rp[1], rp[2], rp[3]}, Cp{rp[4], rp[5], rp[6], rp[7]};
      const M2<float> An{rn[0], rn[1], rn[2], rn[3]}, Cn{rn[4], rn[5], rn[6], rn[7]};
      const M2<float> ACp = mul(Am, Cp), CAn = mul(C, An);
      const M2<float> Bn{1.0f - ACp.a00 - CAn.a00, -ACp.a01 - CAn.a01,
                         -ACp.a10 - CAn.a10, 1.0f - ACp.a11 - CAn.a11};
      const M2<float> Bi = inv32(Bn);
      const float t0 = q0 - (Am.a00 * rp[8] + Am.a01 * rp[9]) - (C.a00 * rn[8] + C.a01 * rn[9]);
      const float t1 = q1 - (Am.a10 * rp[8] + Am.a11 * rp[9]) - (C.a10 * rn[8] + C.a11 * rn[9]);
      const M2<float> AA = mul(Am, Ap), CC = mul(C, Cn);
      Am = mul(Bi, M2<float>{-AA.a00, -AA.a01, -AA.a10, -AA.a11});
      C = mul(Bi, M2<float>{-CC.a00, -CC.a01, -CC.a10, -CC.a11});
      q0 = Bi.a00 * t0 + Bi.a01 * t1;
      q1 = Bi.a10 * t0 + Bi.a11 * t1;
    };
    round(1 << MAXR); // not negligible: that is why more is set
    for (int h = 2 << MAXR; h < P; h <<= 1) {
      if (ex.all(pcr_negligible(Am, C))) break;
      round(h);
    }
  }
  float s0, s1;
  {
    float send[2] = {q0, q1}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    s0 = rp[0];
    s1 = rp[1];
  }
  float xn1 = q0, xn2 = q1;
  A.refresh();
  if constexpr (!SP) {
    // the forward sweep again with the previous separator moved across
    // (zero couplings on the first chunk, whose s(0), s(1), l(0) are zero)
    zf[0] = (f[0] - A.s(0) * s0 - A.l(0) * s1) * F.iv[0];
    zf[1] = (f[1] - A.s(1) * s1 - A.l(1) * zf[0]) * F.iv[1];
#pragma unroll
    for (int i = 2; i < m; ++i)
      zf[i] = (f[i] - A.s(i) * zf[i - 2] - ir_ddl(F, A, i) * zf[i - 1]) * F.iv[i];
  }
#pragma unroll
  for (int i = m - 1; i >= 0; --i) {
    float hi;
    if constexpr (SP)
      hi = zf[i] + F.za[i] * s0 + F.zb[i] * s1;
    else
      hi = zf[i];
    const float xi = hi - F.du2[i] * xn1 - ir_dw2(F, A, i) * xn2;
    f[i] = xi;
    xn2 = xn1;
    xn1 = xi;
  }
  f[m] = q0;
  f[m + 1] = q1;
}

// Full refinement on one chunk.  b: FP64 RHS rows (accessor); x: FP64
// solution rows (output).  Returns the number of refinement steps taken
// (uniform over the lanes that exchange with each other).
// x: the FP64 iterate, a register array double[M] or any X with
// double &operator[](int) (an LDS-backed one keeps M doubles out of the
// registers).
template <int M, int MAXR, bool SP = true, typename Cf, typename Bv, typename X, typename Ex>
__device__ __forceinline__ int ir_solve(const Cf &A, const Bv &b, X &x,
                                        int p, int P, Ex &ex, int max_steps,
                                        double tol) {
  IrFactors<M, MAXR, SP> F;
  ir_factor<M, MAXR>(F, A, p, P, ex);
  float f[M];
#pragma unroll
  for (int i = 0; i < M; ++i) f[i] = (float)b(i);
  ir_apply<M, MAXR>(F, A, f, p, P, ex);
  float xloc = 0.0f;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    x[i] = (double)f[i];
    xloc = fmaxf(xloc, fabsf(f[i]));
  }
  // the norms of the stopping rule only steer it: FP32 maxima of FP32 values
  const double xnorm = (double)ex.max_f32(xloc);
  double dprev = xnorm; // dx_0 := x_0
  int k = 0;
  while (k < max_steps) {
    // halo: the previous chunk's last two rows, the next chunk's first two.
    // Where a neighbour chunk does not exist the halo is some finite value
    // and the coupling multiplying it is zero (the caller zeroes the
    // out-of-range couplings of the first and last chunk).
    double send[4] = {x[0], x[1], x[M - 2], x[M - 1]}, hp[4], hn[4];
    ex.template swapc<4>(send, hp, hn, 1);
    float dxm = 0.0f;
    A.refresh();
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const double xm2 = (i >= 2) ? x[i - 2] : hp[2 + i];
      const double xm1 = (i >= 1) ? x[i - 1] : hp[3];
      const double xp1 = (i + 1 < M) ? x[i + 1] : hn[0];
      const double xp2 = (i + 2 < M) ? x[i + 2] : hn[i + 2 - M];
      double acc = A.d64(i) * x[i];
      acc += A.s64(i) * xm2;
      acc += A.l64(i) * xm1;
      acc += A.u64(i) * xp1;
      acc += A.w64(i) * xp2;
      f[i] = (float)(b(i) - acc);
    }
    ir_apply<M, MAXR>(F, A, f, p, P, ex);
#pragma unroll
    for (int i = 0; i < M; ++i) {
      x[i] += (double)f[i];
      dxm = fmaxf(dxm, fabsf(f[i]));
    }
    ++k;
    // Stop once the estimated relative forward error of the new iterate,
    // |dx_k|^2 / (|dx_{k-1}| |x|) (contraction |dx_k|/|dx_{k-1}| times the
    // error |dx_k| it removed), is at most tol; norms are system maxima.
    const double dxs = (double)ex.max_f32(dxm);
    const bool conv = (dxs == 0.0) || (dxs * dxs <= tol * xnorm * dprev);
    dprev = dxs;
    if (ex.all(conv)) break;
  }
  return k;
}

} // namespace pentad
