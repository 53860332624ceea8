// kernels_partition.hip -- register-resident partitioned solves for gfx950.
//
// k_part_x   x direction (contiguous lines).  A wavefront holds 64/L systems,
//            L lanes per system, each lane M consecutive rows loaded with
//            16-byte vector loads.  The reduced block system is exchanged
//            between the L lanes with wavefront shuffles (ds_bpermute).
// k_part_yz  strided directions.  A workgroup holds S adjacent systems and
//            P = len/M chunks of each; thread t owns system t % S, chunk
//            t / S, so at every row the S lanes of a chunk read S consecutive
//            words (coalesced).  The reduced system goes through LDS.
//
// Both keep every row in registers between load and store: HBM traffic is
// the compulsory 5 coefficient reads + 1 RHS read + 1 solution write per row.
#include "partition_core.hpp"
#include "partition_exchange.hpp"
#include "rows_io.hpp"
#include "rotated_store.hpp"
#include "pentad_internal.hpp"

namespace pentad {

// ---------------------------------------------------------------------------
// x direction
// ---------------------------------------------------------------------------
// Addressing through XMap (lines contiguous, see pentad_internal.hpp).
// ROT = false: the solution overwrites the RHS (out == b).
// ROT = true : the solution goes to out at the mapped positions, staged
//              through LDS so that each store instruction writes runs.
//              Workgroups are dealt to XCDs so that workgroups with adjacent
//              systems (whose runs share 128-byte lines) use one L2.
// (No occupancy hint: forcing FP64 rows=8 down to 128 VGPRs for 4 waves per
// SIMD measured 8% slower than the compiler's 138 VGPRs at 3 waves.)
// Q (ROT only): the workgroup solves Q groups of blockDim / L systems one
// after the other into one tile, so that each row of the tile is a run of
// Q * blockDim / L systems (wider runs, fewer partial 128-byte lines).
// M rows of one diagonal stored as C into T (C = float, T = double: the FP32
// coefficients widened on load, exactly)
template <typename T, typename C, int M>
__device__ __forceinline__ void load_coef(const C *__restrict__ src, T (&v)[M]) {
  if constexpr (sizeof(C) == sizeof(T)) {
    RowsIO<T, M>::load(src, v);
  } else {
    C t[M];
    RowsIO<C, M>::load(src, t);
#pragma unroll
    for (int i = 0; i < M; ++i) v[i] = (T)t[i];
  }
}

// the Q passes of k_part_x (and its tile write-out) over exchange Ex
template <typename T, int M, bool ROT, int Q, typename C, typename Ex>
__device__ __forceinline__ void part_x_passes(const C *__restrict__ ds, const C *__restrict__ dl,
                                              const C *__restrict__ d, const C *__restrict__ du,
                                              const C *__restrict__ dw, const T *b, T *out,
                                              int64_t n_sys, int64_t len, const XMap &map,
                                              int64_t blk, int L, int s0shift, int sl, int p,
                                              Ex &ex, unsigned char *smem_raw) {
#pragma unroll 1
  for (int q = 0; q < Q; ++q) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    const bool valid = sys < n_sys;
    const int64_t sc = valid ? sys : n_sys - 1;
    const int64_t slo = map.lo(sc), shi = map.hi(sc);
    const int64_t off = slo * map.in_lo + shi * map.in_hi + (int64_t)p * M;
    const int64_t offc = slo * map.c_lo + shi * map.c_hi + (int64_t)p * M;

    T s[M], l[M], dd[M], u[M], w[M], f[M];
    load_coef<T, C, M>(ds + offc, s);
    load_coef<T, C, M>(dl + offc, l);
    load_coef<T, C, M>(d + offc, dd);
    load_coef<T, C, M>(du + offc, u);
    load_coef<T, C, M>(dw + offc, w);
    RowsIO<T, M>::load(b + off, f);

    part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
    if constexpr (ROT) {
      tile_put<T, M>(reinterpret_cast<T *>(smem_raw), f, p, Q << s0shift, (q << s0shift) + sl);
    } else {
      if (valid) RowsIO<T, M>::store(out + off, f);
    }
  }
  if constexpr (ROT) {
    __syncthreads();
    const int sshift = s0shift + __builtin_ctz(Q);
    tile_flush<T, M>(reinterpret_cast<const T *>(smem_raw), sshift, (int)len, out, n_sys,
                     blk << sshift, map);
  }
}

// C: the coefficients' storage type (float with T = double: one FP64 solve of
// the FP32 coefficients widened on load, the mixed-precision contract met
// where no refinement kernel covers the length, see launch_part_x_map_widened)
// WIDE: more than 64 lanes per system (128 or 256): the lanes of a system
// span wavefronts and exchange through LDS (LdsEx, 10 values per thread),
// placed after the rotated tile.  Long FP64 lines that way keep 8-row chunks
// (three wavefronts per SIMD) where 16-row chunks in one wavefront hold two.
template <typename T, int M, bool ROT, int Q, typename C = T, bool WIDE = false>
__global__ void __launch_bounds__(512)
    k_part_x(const C *__restrict__ ds, const C *__restrict__ dl,
             const C *__restrict__ d, const C *__restrict__ du,
             const C *__restrict__ dw, const T *b, T *out, int64_t n_sys,
             int64_t len, int lshift /* log2 lanes per system */, XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t blk = ROT ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  const int L = 1 << lshift;
  const int s0shift = __builtin_ctz(blockDim.x) - lshift; // systems per pass
  const int sl = (int)threadIdx.x >> lshift;
  const int p = (int)threadIdx.x & (L - 1);
  if constexpr (WIDE) {
    const size_t tb = ROT ? rotated_tile_bytes<T>((int)blockDim.x * Q, M, len) : 0;
    LdsEx ex{smem_raw + ((tb + 15) & ~(size_t)15), sl, p, (int)blockDim.x >> lshift, L};
    part_x_passes<T, M, ROT, Q, C>(ds, dl, d, du, dw, b, out, n_sys, len, map, blk, L,
                                   s0shift, sl, p, ex, smem_raw);
  } else {
    WaveEx ex{(int)(threadIdx.x & 63), p, L};
    part_x_passes<T, M, ROT, Q, C>(ds, dl, d, du, dw, b, out, n_sys, len, map, blk, L,
                                   s0shift, sl, p, ex, smem_raw);
  }
}

// the passes of k_part_xq over exchange Ex
template <typename T, int M, int Q, bool ROT, int NA, typename C, typename Ex>
__device__ __forceinline__ void part_xq_passes(const C *__restrict__ ds, const C *__restrict__ dl,
                                               const C *__restrict__ d, const C *__restrict__ du,
                                               const C *__restrict__ dw, const T *b, T *out,
                                               int64_t n_sys, int64_t len, int lshift,
                                               const XMap &map, Ex &ex,
                                               unsigned char *smem_raw) {
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int L = 1 << lshift;
  const int s0shift = 8 - lshift; // systems per pass (256 threads)
  const int sl = (int)threadIdx.x >> lshift;
  const int p = (int)threadIdx.x & (L - 1);
  auto offset = [&](int q) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    const int64_t sc = sys < n_sys ? sys : n_sys - 1;
    return map.lo(sc) * map.in_lo + map.hi(sc) * map.in_hi + (int64_t)p * M;
  };
  auto load = [&](int q, T (&s)[M], T (&l)[M], T (&dd)[M], T (&u)[M], T (&w)[M],
                  T (&f)[M]) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    const int64_t sc = sys < n_sys ? sys : n_sys - 1;
    const int64_t slo = map.lo(sc), shi = map.hi(sc);
    const int64_t off = slo * map.in_lo + shi * map.in_hi + (int64_t)p * M;
    const int64_t offc = slo * map.c_lo + shi * map.c_hi + (int64_t)p * M;
    load_coef<T, C, M>(ds + offc, s);
    load_coef<T, C, M>(dl + offc, l);
    load_coef<T, C, M>(d + offc, dd);
    load_coef<T, C, M>(du + offc, u);
    load_coef<T, C, M>(dw + offc, w);
    RowsIO<T, M>::load(b + off, f);
  };
  // array k of pass q into dst (k: 0..4 the diagonals, 5 the right-hand side)
  auto load_one = [&](int q, int k, T (&dst)[M]) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    const int64_t sc = sys < n_sys ? sys : n_sys - 1;
    const int64_t slo = map.lo(sc), shi = map.hi(sc);
    if (k == 5) {
      RowsIO<T, M>::load(b + slo * map.in_lo + shi * map.in_hi + (int64_t)p * M, dst);
      return;
    }
    const C *src = k == 0 ? ds : k == 1 ? dl : k == 2 ? d : k == 3 ? du : dw;
    load_coef<T, C, M>(src + slo * map.c_lo + shi * map.c_hi + (int64_t)p * M, dst);
  };
  T s[M], l[M], dd[M], u[M], w[M], f[M];
  load(0, s, l, dd, u, w, f);
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    if constexpr (NA < 6) {
      T nx[NA][M];
#pragma unroll
      for (int k = 0; k < NA; ++k)
        if (q + 1 < Q) load_one(q + 1, k, nx[k]);
      part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
      if constexpr (ROT) {
        tile_put<T, M>(reinterpret_cast<T *>(smem_raw), f, p, Q << s0shift,
                       (q << s0shift) + sl);
      } else {
        if (((blk * Q + q) << s0shift) + sl < n_sys) RowsIO<T, M>::store(out + offset(q), f);
      }
      if (q + 1 < Q) {
        T(*const work[6])[M] = {&s, &l, &dd, &u, &w, &f};
#pragma unroll
        for (int k = NA; k < 6; ++k) load_one(q + 1, k, *work[k]);
#pragma unroll
        for (int k = 0; k < NA; ++k)
#pragma unroll
          for (int i = 0; i < M; ++i) (*work[k])[i] = nx[k][i];
      }
      continue;
    }
    T ns[M], nl[M], nd[M], nu[M], nw[M], nf[M];
    if (q + 1 < Q) load(q + 1, ns, nl, nd, nu, nw, nf);
    part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
    if constexpr (ROT) {
      tile_put<T, M>(reinterpret_cast<T *>(smem_raw), f, p, Q << s0shift, (q << s0shift) + sl);
    } else {
      if (((blk * Q + q) << s0shift) + sl < n_sys) RowsIO<T, M>::store(out + offset(q), f);
    }
    if (q + 1 < Q) {
#pragma unroll
      for (int i = 0; i < M; ++i) {
        s[i] = ns[i];
        l[i] = nl[i];
        dd[i] = nd[i];
        u[i] = nu[i];
        w[i] = nw[i];
        f[i] = nf[i];
      }
    }
  }
  if constexpr (ROT) {
    __syncthreads();
    const int sshift = s0shift + __builtin_ctz(Q);
    tile_flush<T, M>(reinterpret_cast<const T *>(smem_raw), sshift, (int)len, out, n_sys,
                     blk << sshift, map);
  }
}

// Rotated store in Q > 1 passes, software-pipelined: the rows of pass q + 1
// are loaded before pass q is solved, so the loads of one pass overlap the
// arithmetic of the previous one.  The 64 KB tile allows two workgroups of
// 256 threads per CU (2 waves per SIMD), which leaves up to 256 VGPRs for the
// two sets of rows (FP64 M <= 8, FP32 M <= 16; larger chunks keep k_part_x).
// Measured at 512^3: FP64 1.56 -> 1.44 ms, FP32 (4 passes of 8 systems
// instead of 2 passes of 16 in 512-thread workgroups) 0.87 -> 0.73 ms.
// ROT = false: the same pipelined passes solving in place (out == b, no
// tile; measured at 512^3: FP64 1.55 -> 1.44 ms with Q = 2, FP32 0.81 ->
// 0.76 ms with Q = 4).
// NA: arrays of the next pass loaded ahead (of ds, dl, d, du, dw, b in that
// order); the rest are loaded after the current pass's solve.  6 = all; fewer
// where both full sets of rows do not fit the registers.
// WIDE: 128 or 256 lanes per system, exchanged through LDS after the tile
// (as k_part_x).
template <typename T, int M, int Q, bool ROT, int NA = 6, typename C = T, bool WIDE = false>
__global__ void __launch_bounds__(256, 2)
    k_part_xq(const C *__restrict__ ds, const C *__restrict__ dl,
              const C *__restrict__ d, const C *__restrict__ du,
              const C *__restrict__ dw, const T *b, T *out, int64_t n_sys,
              int64_t len, int lshift, XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = 1 << lshift;
  const int sl = (int)threadIdx.x >> lshift;
  const int p = (int)threadIdx.x & (L - 1);
  if constexpr (WIDE) {
    const size_t tb = ROT ? rotated_tile_bytes<T>(256 * Q, M, len) : 0;
    LdsEx ex{smem_raw + ((tb + 15) & ~(size_t)15), sl, p, 256 >> lshift, L};
    part_xq_passes<T, M, Q, ROT, NA, C>(ds, dl, d, du, dw, b, out, n_sys, len, lshift, map,
                                        ex, smem_raw);
  } else {
    WaveEx ex{(int)(threadIdx.x & 63), p, L};
    part_xq_passes<T, M, Q, ROT, NA, C>(ds, dl, d, du, dw, b, out, n_sys, len, lshift, map,
                                        ex, smem_raw);
  }
}

// ---------------------------------------------------------------------------
// strided (y/z) directions
// ---------------------------------------------------------------------------
// C: the coefficients' storage type (float coefficients widened to a double
// solve: the mixed-precision contract met by one FP64 solve, see
// launch_part_yz_widened)
template <typename T, int M, int MAXT, typename C = T>
__global__ void __launch_bounds__(MAXT)
    k_part_yz(const C *__restrict__ ds, const C *__restrict__ dl,
              const C *__restrict__ d, const C *__restrict__ du,
              const C *__restrict__ dw, T *__restrict__ x, Lines ln, int S,
              int P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int t = threadIdx.x;
  const int sl = t % S;
  const int p = t / S;
  // adjacent systems share 128-byte lines: keep neighbouring tiles on one XCD
  const int64_t sys = xcd_tile(blockIdx.x, gridDim.x) * S + sl;
  const bool valid = sys < ln.n_sys;
  const int64_t st = ln.stride;
  const int64_t b = ln.base(valid ? sys : ln.n_sys - 1) + (int64_t)p * M * st;

  T s[M], l[M], dd[M], u[M], w[M], f[M];
  const int len = (int)ln.len;
  if (P * M == len) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const int64_t e = b + i * st;
      s[i] = (T)ds[e];
      l[i] = (T)dl[e];
      dd[i] = (T)d[e];
      u[i] = (T)du[e];
      w[i] = (T)dw[e];
      f[i] = x[e];
    }
  } else {
    // padded length: rows past the end are identity rows, couplings of real
    // rows past the end are zero
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const int r = p * M + i;
      const int64_t e = b + i * st;
      const bool in = r < len;
      s[i] = in ? (T)ds[e] : T(0);
      l[i] = in ? (T)dl[e] : T(0);
      dd[i] = in ? (T)d[e] : T(1);
      u[i] = (r + 1 < len) ? (T)du[e] : T(0);
      w[i] = (r + 2 < len) ? (T)dw[e] : T(0);
      f[i] = in ? x[e] : T(0);
    }
  }

  LdsEx ex{smem_raw, sl, p, S, P};
  part_solve<T, M>(s, l, dd, u, w, f, p, P, ex);
  if (valid) {
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (p * M + i < len) x[b + i * st] = f[i];
  }
}

// ---------------------------------------------------------------------------
// x direction, any length ("transpose"): the S consecutive systems of a
// workgroup are one contiguous span of S * len elements per array.  Each
// array is staged through LDS with coalesced loads (consecutive threads,
// consecutive elements, whatever the alignment), then every lane reads its M
// rows out of LDS -- an in-tile transpose instead of a global one.  Rows past
// a system's end (len <= L * M) are identity rows; the solution goes back
// through the same tile.  HBM traffic is the compulsory 7 words per row.
// WIDE: more than 64 lanes per system (lines above 64 x M rows, up to 512
// lanes in a 512-thread workgroup): the separators are exchanged through
// LDS across the workgroup (LdsEx, system t / L, chunk t % L), in the tile's
// own memory, which no one reads between the last staging read and the
// write of the solution.
// ---------------------------------------------------------------------------
template <typename T, int M, bool WIDE = false>
__global__ void __launch_bounds__(WIDE ? 512 : 256)
    k_tr_x(const T *__restrict__ ds, const T *__restrict__ dl, const T *__restrict__ d,
           const T *__restrict__ du, const T *__restrict__ dw, T *__restrict__ x,
           int64_t n_sys, int len, int lshift) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tile = reinterpret_cast<T *>(smem_raw);
  const int L = 1 << lshift, S = (int)blockDim.x >> lshift;
  const int sl = (int)threadIdx.x >> lshift, p = (int)threadIdx.x & (L - 1);
  const int64_t sys0 = (int64_t)blockIdx.x * S;
  const int ns = (int)((n_sys - sys0) < S ? (n_sys - sys0) : S);
  const int64_t base = sys0 * len;
  const int span = ns * len;
  const int own = sl * len; // this lane's system inside the span
  const bool act = sl < ns;
  auto slot = [&](int e) { return e + e / M; }; // one pad per M elements
  auto stage = [&](const T *__restrict__ src) {
    __syncthreads();
    for (int e = threadIdx.x; e < span; e += blockDim.x) tile[slot(e)] = src[base + e];
    __syncthreads();
  };
  T s[M], l[M], dd[M], u[M], w[M], f[M];
  stage(ds);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    s[i] = (act && r < len && r >= 2) ? tile[slot(own + r)] : T(0);
  }
  stage(dl);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    l[i] = (act && r < len && r >= 1) ? tile[slot(own + r)] : T(0);
  }
  stage(d);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    dd[i] = (act && r < len) ? tile[slot(own + r)] : T(1);
  }
  stage(du);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    u[i] = (act && r + 1 < len) ? tile[slot(own + r)] : T(0);
  }
  stage(dw);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    w[i] = (act && r + 2 < len) ? tile[slot(own + r)] : T(0);
  }
  stage(x);
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    f[i] = (act && r < len) ? tile[slot(own + r)] : T(0);
  }
  if constexpr (WIDE) {
    LdsEx ex{smem_raw, sl, p, S, L};
    part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
  } else {
    WaveEx ex{(int)(threadIdx.x & 63), p, L};
    part_solve<T, M>(s, l, dd, u, w, f, p, L, ex);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int r = p * M + i;
    if (act && r < len) tile[slot(own + r)] = f[i];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < span; e += blockDim.x) x[base + e] = tile[slot(e)];
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
// chunk sizes of the wide x kernels (more than 64 lanes per system): 8 rows
// for FP64 (three wavefronts per SIMD), 8 or 16 for FP32
template <typename T> constexpr bool wide_rows(int M) {
  return M == 8 || (sizeof(T) == 4 && M == 16);
}

namespace {

template <typename T, int M, typename C = T>
hipError_t go_x(const Diags<C> &a, const T *b, T *out, bool rot, int64_t n_sys,
                int64_t len, const XMap &map, int lanes, hipStream_t st) {
  int lshift = 0;
  while ((1 << lshift) < lanes) ++lshift;
  if (lanes > 64) {
    // lanes of a system across wavefronts (part_x_wide_supported):
    // 256-thread workgroups, the exchange after the rotated tile
    if constexpr (wide_rows<T>(M)) {
      constexpr int block = 256;
      const size_t xb = (size_t)10 * block * sizeof(T);
      auto go = [&](auto kernel, int64_t grid, size_t smem) {
        if (smem > 65536) {
          const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)smem);
          if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), smem, st, a.ds, a.dl,
                           a.d, a.du, a.dw, b, out, n_sys, len, lshift, map);
        return hipGetLastError();
      };
      const int sys_per_pass = block >> lshift;
      if (!rot) {
        // in place: two software-pipelined passes per workgroup; one pass
        // where that leaves at most two workgroups per CU (FP64 1024-row
        // lines, 32768 systems: 0.40 -> 0.39 ms, 128 systems: 7.8 against
        // 12.9 us with two passes; four passes 0.46 ms)
        if ((n_sys + sys_per_pass - 1) / sys_per_pass <= 512)
          return go(k_part_x<T, M, false, 1, C, true>,
                    (n_sys + sys_per_pass - 1) / sys_per_pass, xb);
        const int64_t per_wg = (int64_t)sys_per_pass * 2;
        return go(k_part_xq<T, M, 2, false, 6, C, true>, (n_sys + per_wg - 1) / per_wg, xb);
      }
      // rotated: passes per workgroup for the widest tile row up to a
      // 128-byte run that fits the 160 KB of LDS with the exchange, software-
      // pipelined (FP64 1024-row lines, 8 passes: 0.57 -> 0.50 ms)
      auto bytes = [&](int q) {
        return ((rotated_tile_bytes<T>(block * q, M, len) + 15) & ~(size_t)15) + xb;
      };
      int Q = 1;
      while (Q < 8 && sys_per_pass * Q * (int)sizeof(T) < 128 && bytes(2 * Q) <= 163840) Q *= 2;
      const int64_t per_wg = (int64_t)sys_per_pass * Q;
      const int64_t grid = (n_sys + per_wg - 1) / per_wg;
      switch (Q) {
      case 1: return go(k_part_x<T, M, true, 1, C, true>, grid, bytes(1));
      case 2: return go(k_part_xq<T, M, 2, true, 6, C, true>, grid, bytes(2));
      case 4: return go(k_part_xq<T, M, 4, true, 6, C, true>, grid, bytes(4));
      default: return go(k_part_xq<T, M, 8, true, 6, C, true>, grid, bytes(8));
      }
    }
    return hipErrorInvalidValue;
  }
  constexpr bool pipelined = sizeof(T) == 8 ? M <= 8 : M <= 16;
  if (!rot) {
    const int block = 256;
    // a batch that gives at most two one-pass workgroups per CU runs one
    // pass per workgroup: passes in sequence only add latency there
    const bool small = (n_sys * lanes + block - 1) / block <= 512;
    if constexpr (pipelined) if (!small) {
      // four software-pipelined passes per workgroup (FP64 in place at
      // 256^3: 0.215 -> 0.206 ms against two; equal at 512^3)
      constexpr int Q = 4;
      const int64_t per_wg = (int64_t)(block >> lshift) * Q;
      const int64_t grid = (n_sys + per_wg - 1) / per_wg;
      hipLaunchKernelGGL((k_part_xq<T, M, Q, false, 6, C>), dim3((unsigned)grid), dim3(block), 0,
                         st, a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len, lshift, map);
      return hipGetLastError();
    }
    if constexpr (sizeof(T) == 8 && (M == 10 || M == 12)) if (!small) {
      // two passes, two arrays of the next one loaded ahead (as the rotated
      // path below)
      const int64_t per_wg = (int64_t)(block >> lshift) * 2;
      const int64_t grid = (n_sys + per_wg - 1) / per_wg;
      hipLaunchKernelGGL((k_part_xq<T, M, 2, false, 2, C>), dim3((unsigned)grid), dim3(block), 0,
                         st, a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len, lshift, map);
      return hipGetLastError();
    }
    const int64_t grid = (n_sys * lanes + block - 1) / block;
    hipLaunchKernelGGL((k_part_x<T, M, false, 1, C>), dim3((unsigned)grid), dim3(block), 0,
                       st, a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len, lshift, map);
    return hipGetLastError();
  }
  // rotated stores: Q passes per workgroup widen each tile row to a 128-byte
  // run (measured at 512^3, FP64 rows=8 x 64 lanes: 1.89 / 1.76 / 1.64 /
  // 1.73 ms for Q = 1 / 2 / 4 / 8); several passes run software-pipelined
  // in 256-thread workgroups (k_part_xq) where both sets of rows fit
  if constexpr (pipelined) {
    const int run = (256 >> lshift) * (int)sizeof(T);
    if (run < 128) {
      // FP64 takes four passes also where two make a 128-byte run (256^3 ADI
      // step, 32 lanes per system: 0.604 -> 0.583 ms); FP32 keeps two there
      // (0.300 vs 0.306 ms with four)
      const int Q = (run >= 64 && sizeof(T) == 4) ? 2 : 4;
      const int64_t per_wg = (int64_t)(256 >> lshift) * Q;
      const int64_t grid = (n_sys + per_wg - 1) / per_wg;
      const size_t smem = rotated_tile_bytes<T>(256 * Q, M, len);
      auto go = [&](auto kernel) {
        if (smem > 65536) {
          const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)smem);
          if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(256), smem, st, a.ds, a.dl,
                           a.d, a.du, a.dw, b, out, n_sys, len, lshift, map);
        return hipGetLastError();
      };
      return Q == 2 ? go(k_part_xq<T, M, 2, true, 6, C>) : go(k_part_xq<T, M, 4, true, 6, C>);
    }
  }
  // FP64 chunks of 10 / 12 rows: both full sets of rows do not fit, so the
  // next pass loads two arrays ahead and the other four after the solve
  // (384 x 256 x 640 FP64 step, 10-row chunks along z: 2.23 -> 2.06 ms; 3, 4
  // or 5 arrays ahead measured the same)
  if constexpr (sizeof(T) == 8 && (M == 10 || M == 12)) {
    const int run = (256 >> lshift) * (int)sizeof(T);
    if (run < 128) {
      const int Q = run >= 64 ? 2 : 4;
      const int64_t per_wg = (int64_t)(256 >> lshift) * Q;
      const int64_t grid = (n_sys + per_wg - 1) / per_wg;
      const size_t smem = rotated_tile_bytes<T>(256 * Q, M, len);
      auto go = [&](auto kernel) {
        if (smem > 65536) {
          const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                                   (int)smem);
          if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(256), smem, st, a.ds, a.dl,
                           a.d, a.du, a.dw, b, out, n_sys, len, lshift, map);
        return hipGetLastError();
      };
      return Q == 2 ? go(k_part_xq<T, M, 2, true, 2, C>) : go(k_part_xq<T, M, 4, true, 2, C>);
    }
  }
  // one pass, or chunks too large to hold two sets of rows: FP32 in
  // 512-thread workgroups, FP64 in 256 (3 waves per SIMD fit its state)
  const int block = sizeof(T) == 4 ? 512 : 256;
  const int run = (block >> lshift) * (int)sizeof(T); // bytes per pass
  const int Q = run >= 128 ? 1 : (run >= 64 ? 2 : 4);
  const int64_t per_wg = (int64_t)(block >> lshift) * Q;
  const int64_t grid = (n_sys + per_wg - 1) / per_wg;
  const size_t smem = rotated_tile_bytes<T>(block * Q, M, len);
  auto go = [&](auto kernel) {
    if (smem > 65536) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)smem);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), smem, st, a.ds, a.dl, a.d,
                       a.du, a.dw, b, out, n_sys, len, lshift, map);
    return hipGetLastError();
  };
  if (Q == 1) return go(k_part_x<T, M, true, 1, C>);
  if (Q == 2) return go(k_part_x<T, M, true, 2, C>);
  return go(k_part_x<T, M, true, 4, C>);
}

// Largest workgroup each (type, M) instantiation is compiled for.  With W
// waves per SIMD a lane may use 512/W VGPRs; a 1024-thread workgroup means
// 4 waves per SIMD (128 VGPRs), which only the small chunks fit without
// spilling (checked with -Rpass-analysis=kernel-resource-usage).
constexpr int yz_maxt_rt(int elem_bytes, int rows) {
  return (elem_bytes == 8) ? (rows <= 4 ? 1024 : 512)
                           : (rows <= 12 ? 1024 : 512);
}
template <typename T, int M> constexpr int yz_maxt() {
  return yz_maxt_rt((int)sizeof(T), M);
}

template <typename T, int M, typename C = T>
hipError_t go_yz(const Diags<C> &a, T *x, const Lines &ln, int S,
                 hipStream_t st) {
  constexpr int MAXT = yz_maxt<T, M>();
  const int P = (int)((ln.len + M - 1) / M);
  const int block = S * P;
  if (block > MAXT) return hipErrorInvalidConfiguration;
  const size_t smem = (size_t)10 * S * P * sizeof(T);
  const int64_t grid = (ln.n_sys + S - 1) / S;
  if (smem > 65536) {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_part_yz<T, M, MAXT, C>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_part_yz<T, M, MAXT, C>), dim3((unsigned)grid),
                     dim3(block), smem, st, a.ds, a.dl, a.d, a.du, a.dw, x, ln,
                     S, P);
  return hipGetLastError();
}

} // namespace

bool part_x_supported(int elem_bytes, int rows, int lanes) {
  if (lanes < 1 || lanes > 64 || (lanes & (lanes - 1)) != 0) return false;
  return rows == 4 || rows == 6 || rows == 8 || rows == 10 || rows == 12 ||
         rows == 16;
}

bool part_x_wide_supported(int elem_bytes, int rows, int lanes) {
  if (lanes != 128 && lanes != 256) return false;
  return elem_bytes == 8 ? wide_rows<double>(rows) : wide_rows<float>(rows);
}

template <typename T>
hipError_t launch_part_x_map(const Diags<T> &a, const T *b, T *out, bool rot,
                             int64_t n_sys, int64_t len, const XMap &map, int rows,
                             int lanes, hipStream_t st) {
  if (!(part_x_supported(sizeof(T), rows, lanes) || part_x_wide_supported(sizeof(T), rows, lanes)) ||
      rows * lanes != len)
    return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_x<T, 4>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 6: return go_x<T, 6>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 8: return go_x<T, 8>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 10: return go_x<T, 10>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 12: return go_x<T, 12>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 16: return go_x<T, 16>(a, b, out, rot, n_sys, len, map, lanes, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_part_x_map_widened(const Diags<float> &a, const double *b, double *out,
                                     bool rot, int64_t n_sys, int64_t len, const XMap &map,
                                     int rows, int lanes, hipStream_t st) {
  if (!(part_x_supported(8, rows, lanes) || part_x_wide_supported(8, rows, lanes)) ||
      rows * lanes != len)
    return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_x<double, 4, float>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 6: return go_x<double, 6, float>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 8: return go_x<double, 8, float>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 10: return go_x<double, 10, float>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 12: return go_x<double, 12, float>(a, b, out, rot, n_sys, len, map, lanes, st);
  case 16: return go_x<double, 16, float>(a, b, out, rot, n_sys, len, map, lanes, st);
  }
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_part_x(const Diags<T> &a, T *x, const Lines &ln, int rows,
                         int lanes, hipStream_t st) {
  return launch_part_x_map<T>(a, x, x, false, ln.n_sys, ln.len, xmap_inplace(ln), rows,
                              lanes, st);
}

bool part_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 12 && rows != 16)
    return false;
  if (len < 3) return false;
  const int64_t P = (len + rows - 1) / rows; // a partial last chunk is padded
  const int64_t block = P * sys_per_wg;     // (a partial last wavefront is fine)
  return block <= yz_maxt_rt(elem_bytes, rows);
}

// LDS-staged x solve: rows 8 or 16, lanes a power of two <= 512, any len
// with lanes * rows >= len
bool tr_x_supported(int rows, int lanes, int64_t len) {
  if (rows != 8 && rows != 16) return false;
  if (lanes < 1 || lanes > 512 || (lanes & (lanes - 1))) return false;
  return len >= 3 && len <= (int64_t)rows * lanes;
}

template <typename T>
hipError_t launch_tr_x(const Diags<T> &a, T *x, const Lines &ln, int rows, int lanes,
                       hipStream_t st) {
  if (!tr_x_supported(rows, lanes, ln.len) || ln.dir != 0) return hipErrorInvalidValue;
  int lshift = 0;
  while ((1 << lshift) < lanes) ++lshift;
  const bool wide = lanes > 64;
  const int block = wide ? 512 : 256, S = block >> lshift;
  const int64_t grid = (ln.n_sys + S - 1) / S;
  const int64_t span = (int64_t)S * ln.len;
  size_t smem = (size_t)(span + span / rows + 1) * sizeof(T);
  if (wide) { // the LDS exchange (10 values per thread) shares the tile
    const size_t ex = (size_t)10 * block * sizeof(T);
    if (ex > smem) smem = ex;
    const void *fn = rows == 8 ? reinterpret_cast<const void *>(&k_tr_x<T, 8, true>)
                               : reinterpret_cast<const void *>(&k_tr_x<T, 16, true>);
    const hipError_t e =
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    if (rows == 8)
      hipLaunchKernelGGL((k_tr_x<T, 8, true>), dim3((unsigned)grid), dim3(block), smem, st,
                         a.ds, a.dl, a.d, a.du, a.dw, x, ln.n_sys, (int)ln.len, lshift);
    else
      hipLaunchKernelGGL((k_tr_x<T, 16, true>), dim3((unsigned)grid), dim3(block), smem, st,
                         a.ds, a.dl, a.d, a.du, a.dw, x, ln.n_sys, (int)ln.len, lshift);
    return hipGetLastError();
  }
  if (rows == 8)
    hipLaunchKernelGGL((k_tr_x<T, 8>), dim3((unsigned)grid), dim3(block), smem, st, a.ds,
                       a.dl, a.d, a.du, a.dw, x, ln.n_sys, (int)ln.len, lshift);
  else
    hipLaunchKernelGGL((k_tr_x<T, 16>), dim3((unsigned)grid), dim3(block), smem, st, a.ds,
                       a.dl, a.d, a.du, a.dw, x, ln.n_sys, (int)ln.len, lshift);
  return hipGetLastError();
}
template hipError_t launch_tr_x<float>(const Diags<float> &, float *, const Lines &, int, int,
                                       hipStream_t);
template hipError_t launch_tr_x<double>(const Diags<double> &, double *, const Lines &, int,
                                        int, hipStream_t);

template <typename T>
hipError_t launch_part_yz(const Diags<T> &a, T *x, const Lines &ln, int rows,
                          int sys_per_wg, hipStream_t st) {
  if (!part_yz_supported(sizeof(T), rows, sys_per_wg, ln.len))
    return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_yz<T, 4>(a, x, ln, sys_per_wg, st);
  case 6: return go_yz<T, 6>(a, x, ln, sys_per_wg, st);
  case 8: return go_yz<T, 8>(a, x, ln, sys_per_wg, st);
  case 12: return go_yz<T, 12>(a, x, ln, sys_per_wg, st);
  case 16: return go_yz<T, 16>(a, x, ln, sys_per_wg, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_part_yz_widened(const Diags<float> &a, double *x, const Lines &ln, int rows,
                                  int sys_per_wg, hipStream_t st) {
  if (!part_yz_supported(8, rows, sys_per_wg, ln.len)) return hipErrorInvalidValue;
  switch (rows) {
  case 4: return go_yz<double, 4, float>(a, x, ln, sys_per_wg, st);
  case 6: return go_yz<double, 6, float>(a, x, ln, sys_per_wg, st);
  case 8: return go_yz<double, 8, float>(a, x, ln, sys_per_wg, st);
  case 12: return go_yz<double, 12, float>(a, x, ln, sys_per_wg, st);
  case 16: return go_yz<double, 16, float>(a, x, ln, sys_per_wg, st);
  }
  return hipErrorInvalidValue;
}

template hipError_t launch_part_x<float>(const Diags<float> &, float *,
                                         const Lines &, int, int, hipStream_t);
template hipError_t launch_part_x<double>(const Diags<double> &, double *,
                                          const Lines &, int, int,
                                          hipStream_t);
template hipError_t launch_part_x_map<float>(const Diags<float> &, const float *,
                                             float *, bool, int64_t, int64_t,
                                             const XMap &, int, int, hipStream_t);
template hipError_t launch_part_x_map<double>(const Diags<double> &,
                                              const double *, double *, bool,
                                              int64_t, int64_t, const XMap &, int,
                                              int, hipStream_t);
template hipError_t launch_part_yz<float>(const Diags<float> &, float *,
                                          const Lines &, int, int,
                                          hipStream_t);
template hipError_t launch_part_yz<double>(const Diags<double> &, double *,
                                           const Lines &, int, int,
                                           hipStream_t);

} // namespace pentad
