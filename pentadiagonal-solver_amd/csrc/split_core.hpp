// split_core.hpp -- one set of strided lines solved by two workgroups.
//
// Along the outermost axis of a 3-D field every row of a line lies in its own
// 2 MB page, so a workgroup that holds S systems x len rows fetches S x 8 B
// per array and page visit, and the per-CU address-translation cache misses
// on a large share of its lookups.  Splitting each line between two
// workgroups (rows [0, len/2) and [len/2, len)) doubles S for the same
// registers, and with it the bytes each translation buys.
//
// The two halves are solved as one partitioned system (partition_core.hpp):
// each half eliminates its chunks and reduces its separators by block PCR,
// carrying the coupling to the other half as a 2x2 right-hand side on an
// unknown pair X:
//   half 0 (rows 0 .. len/2-1): X = G, the other half's first two rows; its
//     last chunk's separator equations reach G through w_m, u_{m+1}, w_{m+1};
//   half 1 (rows len/2 .. len-1): X = S', the first half's last separator
//     (its last two rows), which its first chunk's interior reaches.
// After PCR each block row reads S_p = r_p + E_p X.  Half 0's last block row
// gives S' = y + V G; half 1's first chunk gives G as an affine function of
// S' and of its own separator S_0 = z + W S', so G = g + H S'.  The halves
// trade (y, V) and (g, H) -- six words per system -- and both solve
// (I - V H) S' = y + V g, with identical operations on identical words, so
// both use the same S' and G.  Each half then has X and finishes its chunks.
//
// The trade goes through device memory, between workgroups that may sit on
// different XCDs (separate L2 caches).  As in ROCm's own decoupled look-back
// for gfx942/gfx950, the words and the flags are written and read with
// relaxed agent-scope atomics, which bypass the non-coherent caches, and the
// order between the words and the flag is kept by waiting for the stores
// (s_waitcnt vmcnt(0)) before the flag is written and for the flag load before
// the words are read -- no agent-scope fence, which would write back and
// invalidate the whole L2.
//
// Forward progress: a workgroup waits only for its partner, and the pair is
// formed from tickets drawn at workgroup start (one counter per class of
// blockIdx mod 8), so a waiting workgroup's partner has either started (and
// publishes before it waits) or will be the next workgroup of its class to
// start.  At most one workgroup per class waits on a partner that has not
// started, every other resident workgroup finishes, and the slots it frees
// let the partner start.  The wait is bounded all the same: after
// kSpinLimit polls the workgroup records a time-out in the header word
// kTimeoutWord and finishes, and every later wait of the same solve sees the
// word and gives up at once (the results are then invalid; the word says so).
#pragma once

#include "partition_core.hpp"

namespace pentad {

// words of the split kernels' scratch, in front of the per-pair flags
constexpr int kSplitClasses = 8;     // ticket counters, one per blockIdx % 8
constexpr int kTimeoutWord = 8;      // set to 1 when a wait gave up
constexpr int kSplitHeaderInts = 64; // 256 bytes: counters, time-out, padding
constexpr int kSpinLimit = 1 << 20;

template <typename T> struct SplitWord;
template <> struct SplitWord<double> {
  using W = unsigned long long;
  __device__ static W put(double v) { return (W)__double_as_longlong(v); }
  __device__ static double get(W w) { return __longlong_as_double((long long)w); }
};
template <> struct SplitWord<float> {
  using W = unsigned int;
  __device__ static W put(float v) { return __float_as_uint(v); }
  __device__ static float get(W w) { return __uint_as_float(w); }
};

// s_waitcnt vmcnt(0): every vector memory operation of this wave has
// completed (expcnt and lgkmcnt left at their no-wait maxima)
__device__ __forceinline__ void wait_vmem() {
  __builtin_amdgcn_s_waitcnt(0 | (0x7 << 4) | (0xf << 8));
}

// The trade between the two workgroups of a line set, and the 2x2 interface
// solve.  mine / theirs: [6][S] words of this workgroup and of its partner;
// lds_own: [6][S] staging of this workgroup's words; lds_x: [2][S], on return
// X of every system; lds_xf: [10][S], half 1's first rows kept over the PCR.
template <typename T> struct PairLink {
  using SW = SplitWord<T>;
  typename SW::W *mine, *theirs;
  int *flag_mine, *flag_theirs, *timeout;
  T *lds_own, *lds_x, *lds_xf; // [6][S], [2][S], [10][S]
  int S, half;

  // own: (y, V) on half 0's last chunk, (g, H) on half 1's first chunk
  // (boundary lanes); every lane of system s gets (X0, X1) back.
  __device__ __forceinline__ void trade(const T (&own)[6], bool boundary, int s, T &X0, T &X1) {
    if (boundary) {
#pragma unroll
      for (int k = 0; k < 6; ++k) lds_own[k * S + s] = own[k];
    }
    __syncthreads();
    if (threadIdx.x < 64) {
      const int lane = (int)threadIdx.x;
      for (int i = lane; i < 6 * S; i += 64)
        __hip_atomic_store(mine + i, SW::put(lds_own[i]), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      // the words are stored before the flag is (order only; the atomics
      // themselves are coherent across the XCDs)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      wait_vmem();
      __hip_atomic_store(flag_mine, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // wait for the partner; give up after kSpinLimit polls, or at once
      // when another workgroup of this solve has given up already
      bool ok = false;
      for (int it = 0; it < kSpinLimit; ++it) {
        const int v = __hip_atomic_load(flag_theirs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__builtin_amdgcn_readfirstlane(v) != 0) {
          ok = true;
          break;
        }
        const int gone = __hip_atomic_load(timeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__builtin_amdgcn_readfirstlane(gone) != 0) break;
        __builtin_amdgcn_s_sleep(1);
      }
      wait_vmem();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      if (!ok && lane == 0)
        __hip_atomic_store(timeout, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // interface solve per system: (y, V) from half 0, (g, H) from half 1
      for (int i = lane; i < S; i += 64) {
        T o[6], t[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          o[k] = lds_own[k * S + i];
          t[k] = SW::get(__hip_atomic_load(theirs + k * S + i, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT));
        }
        const T *yv = half == 0 ? o : t;
        const T *gh = half == 0 ? t : o;
        const M2<T> V{yv[2], yv[3], yv[4], yv[5]}, H{gh[2], gh[3], gh[4], gh[5]};
        const M2<T> VH = mul(V, H);
        const M2<T> Ki = inv(M2<T>{T(1) - VH.a00, -VH.a01, -VH.a10, T(1) - VH.a11});
        const T q0 = yv[0] + V.a00 * gh[0] + V.a01 * gh[1];
        const T q1 = yv[1] + V.a10 * gh[0] + V.a11 * gh[1];
        const T sp0 = Ki.a00 * q0 + Ki.a01 * q1; // S', the first half's last two rows
        const T sp1 = Ki.a10 * q0 + Ki.a11 * q1;
        if (half == 0) {
          lds_x[i] = gh[0] + H.a00 * sp0 + H.a01 * sp1; // G
          lds_x[S + i] = gh[1] + H.a10 * sp0 + H.a11 * sp1;
        } else {
          lds_x[i] = sp0;
          lds_x[S + i] = sp1;
        }
      }
    }
    __syncthreads();
    X0 = lds_x[s];
    X1 = lds_x[S + s];
  }
};

// The partitioned solve of one half line (P chunks of M rows per system, this
// lane's chunk p) with the other half reached through link.  On entry
// s,l,d,u,w,f hold the chunk's rows; on return f holds their solution.
template <typename T, int M, typename Ex>
__device__ __forceinline__ void part_solve_split(const T (&s)[M], const T (&l)[M],
                                                 const T (&d)[M], const T (&u)[M],
                                                 const T (&w)[M], T (&f)[M], int p, int P,
                                                 int half, int sl, Ex &ex, PairLink<T> &link) {
  static_assert(M >= 4, "a chunk needs at least two interior rows");
  constexpr int m = M - 2;
  const bool gfirst = (p == 0), glast = (p == P - 1);
  const bool first = gfirst && half == 0; // the line's first chunk
  const bool last = glast && half == 1;   // the line's last chunk
  const bool edge = glast && half == 0;   // separator rows reach the other half

  T du2[m], dw2[m], zf[m], za[m], zb[m];
  T XL1[5], XL2[5], XF0[5], XF1[5];
  chunk_eliminate<T, M>(s, l, d, u, w, f, first, du2, dw2, zf, za, zb, XL1, XL2, XF0, XF1);

  T Y0[5], Y1[5]; // next chunk of this half: interior rows 0, 1
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
  // half 1's first chunk needs its first two rows again after the PCR: kept
  // in LDS meanwhile, not in registers
  if (half == 1 && gfirst) {
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      link.lds_xf[c * link.S + sl] = XF0[c];
      link.lds_xf[(5 + c) * link.S + sl] = XF1[c];
    }
  }
  const T wa = last ? T(0) : w[m];
  const T ub = last ? T(0) : u[m + 1];
  const T wb = last ? T(0) : w[m + 1];
  M2<T> A, B, C;
  T r0, r1;
  separator_block_row<T, M>(s, l, d, u, f, XL1, XL2, Y0, Y1, edge ? T(0) : wa, edge ? T(0) : ub,
                            edge ? T(0) : wb, A, B, C, r0, r1);
  // the coupling to the other half on the right-hand side: B S_p + A S_{p-1}
  // + C S_{p+1} = r + E X
  M2<T> E{T(0), T(0), T(0), T(0)};
  if (edge) E = M2<T>{-wa, T(0), -ub, -wb};         // (wa, 0; ub, wb) G
  if (half == 1 && gfirst) {                        // A S'
    E = M2<T>{-A.a00, -A.a01, -A.a10, -A.a11};
    A = M2<T>{T(0), T(0), T(0), T(0)};
  }
  {
    const M2<T> Bi = inv(B);
    A = mul(Bi, A);
    C = mul(Bi, C);
    E = mul(Bi, E);
    const T t0 = Bi.a00 * r0 + Bi.a01 * r1;
    const T t1 = Bi.a10 * r0 + Bi.a11 * r1;
    r0 = t0;
    r1 = t1;
  }
  block_pcr_impl<T, true>(A, C, r0, r1, E, P, ex);

  // the words this half trades
  T own[6];
  if (half == 0) { // S' = y + V G on the last chunk
    own[0] = r0;
    own[1] = r1;
    own[2] = E.a00;
    own[3] = E.a01;
    own[4] = E.a10;
    own[5] = E.a11;
  } else if (gfirst) {
    // G = (XF0, XF1) = g_f + Gs S' + Gt S_0 with S_0 = z + W S' (z = r, W = E)
    T xf[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) xf[c] = link.lds_xf[c * link.S + sl];
    const M2<T> Gs{xf[1], xf[2], xf[6], xf[7]}, Gt{xf[3], xf[4], xf[8], xf[9]};
    const M2<T> GtW = mul(Gt, E);
    own[0] = xf[0] + Gt.a00 * r0 + Gt.a01 * r1;
    own[1] = xf[5] + Gt.a10 * r0 + Gt.a11 * r1;
    own[2] = Gs.a00 + GtW.a00;
    own[3] = Gs.a01 + GtW.a01;
    own[4] = Gs.a10 + GtW.a10;
    own[5] = Gs.a11 + GtW.a11;
  } else {
#pragma unroll
    for (int k = 0; k < 6; ++k) own[k] = T(0); // not traded
  }
  T X0, X1;
  link.trade(own, half == 0 ? glast : gfirst, sl, X0, X1);
  const T t0 = r0 + E.a00 * X0 + E.a01 * X1; // S_p
  const T t1 = r1 + E.a10 * X0 + E.a11 * X1;

  T s0, s1;
  {
    T send[2] = {t0, t1}, rp[2], rn[2];
    ex.template swapc<2>(send, rp, rn, 1);
    s0 = rp[0];
    s1 = rp[1];
  }
  if (half == 1 && gfirst) { // the previous separator is S'
    s0 = X0;
    s1 = X1;
  }
  chunk_backsub<T, M>(du2, dw2, zf, za, zb, s0, s1, t0, t1, f);
}

} // namespace pentad
