// kernels_ir.hip -- FP32 + FP64-refinement batched solves for gfx950.
//
// k_ir_x    x direction: L lanes of a wavefront per system, M rows per lane,
//           FP32 coefficients via 16-byte loads, FP64 RHS/solution via
//           16-byte loads.  HBM traffic per row: 5 x 4 B + 8 B in, 8 B out.
// k_ir_xq   the same lines with the solution written rotated (ADI layouts):
//           Q passes into one tile whose rows are whole 128-byte lines,
//           coefficients in registers.
// k_ir_yz   strided directions: S systems x P chunks per workgroup, LDS
//           exchange, XCD-aware tile order so that workgroups reading the
//           same 128-byte lines run on the same XCD (same L2).
// k_ir_lane any shape: one lane per system, FP32 factors in the caller's
//           scratch buffer (fallback for lengths the partitioned kernels do
//           not cover).
#include "ir_core.hpp"
#include "rows_io.hpp"
#include "rotated_store.hpp"
#include "pentad_internal.hpp"

namespace pentad {

// right-hand side rows of one chunk, rows past the end of the system read 0
struct VarRhsX {
  const double *p;
  int nv;
  __device__ __forceinline__ double operator()(int i) const { return i < nv ? p[i] : 0.0; }
};

// Largest step count over the launch.  Every wavefront reaching one address
// with an atomic serialises (hundreds of thousands per launch), so the
// current value is read first and the atomic issued only when it grows --
// after the first few wavefronts, almost never.
__device__ __forceinline__ void note_steps(int *steps_out, int k) {
  if (steps_out == nullptr) return;
  const int cur = __hip_atomic_load(steps_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (k > cur) atomicMax(steps_out, k);
}

// ---------------------------------------------------------------------------
// ROT as in k_part_x: false = in place on x ([sys][row]); true = RHS from b,
// solution to out in [row][sys] layout, staged through LDS.
// The coefficients live in LDS ([5][M][64] per wavefront) for the whole
// refinement; the RHS is re-read from memory for each residual (the RHS is
// not written before the final store, in place or not).
// Occupancy: 3 waves per SIMD (<= 168 VGPRs) where the state fits without
// spilling (M <= 8), else 2.
template <int M, int R, bool ROT>
__global__ void __launch_bounds__(256, (M <= 8 ? 3 : 2))
    k_ir_x(const float *__restrict__ ds, const float *__restrict__ dl,
           const float *__restrict__ d, const float *__restrict__ du,
           const float *__restrict__ dw, const double *bin,
           double *out, int64_t n_sys, int64_t len, int max_steps,
           double tol, int *__restrict__ steps_out, XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int L = 1 << R;
  const int64_t blk = ROT ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  const int64_t g = blk * blockDim.x + threadIdx.x;
  const int64_t sys = g >> R;
  const int p = (int)(g & (L - 1));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool valid = sys < n_sys;
  const int64_t sc = valid ? sys : n_sys - 1;
  const int64_t slo = map.lo(sc), shi = map.hi(sc);
  const int64_t off = slo * map.in_lo + shi * map.in_hi + (int64_t)p * M;
  const int64_t offc = slo * map.c_lo + shi * map.c_hi + (int64_t)p * M;
  float *alds = reinterpret_cast<float *>(smem_raw) + wave * (5 * M * 64) + lane;
  {
    float t[M];
    const float *src[5] = {ds, dl, d, du, dw};
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      RowsIO<float, M>::load(src[k] + offc, t);
      ir_zero_out_of_range<M>(k, t, p == 0, p == L - 1);
#pragma unroll
      for (int i = 0; i < M; ++i) alds[(k * M + i) * 64] = t[i];
    }
  }
  double xs[M];
  WaveEx ex{lane, p, L};
  const LdsCoef<M> A{alds};
  const MemRhs b{bin + off, 1};
  const int k = ir_solve<M, ir_stored_rounds(R)>(A, b, xs, p, L, ex, max_steps, tol);
  if constexpr (ROT) {
    stage_rotated_store<double, M>(reinterpret_cast<double *>(smem_raw), xs, p, R,
                                   (int)len, out, n_sys, blk, map);
  } else {
    if (valid) RowsIO<double, M>::store(out + off, xs);
  }
  if (lane == 0) note_steps(steps_out, k);
}

// Rotated store (ADI layouts): Q passes of 256 >> R systems each solve into
// one LDS tile, so that every row of the tile is a 128-byte line written
// whole by this workgroup (4 FP64 systems per pass at 64 lanes: Q = 4).
// The coefficients stay in registers (the LDS holds the 66 KB tile, two
// workgroups per CU, 2 waves per SIMD): a tile of 4 systems with the
// coefficients in LDS runs 3 waves per SIMD but writes 32-byte pieces of
// lines shared with three other workgroups, which the memory system merges
// poorly (1.01 ms against 0.91 ms for this kernel at 512^3).
// BLOCK = 512 (one workgroup per CU, the same 2 waves per SIMD) serves
// chunks of 10 rows: a 16-system tile of 640-row lines takes 82 KB of LDS,
// more than half a CU's.  SP = false recomputes the separator spikes in every
// apply instead of keeping them (ir_core.hpp), which the larger chunk needs
// to stay in registers; SR is the number of stored PCR rounds (IrFactors).
// IL: the five coefficient arrays interleaved per group of IL rows in one
// array (IL = 4: ds holds [s0..3, l0..3, d0..3, u0..3, w0..3] per 4 rows;
// IL = 2: [s0, s1, l0, l1, d0, d1, u0, u1, w0, w1] per 2 rows; the other four
// pointers are unused): one coefficient stream instead of five, the layout
// the ADI driver writes when it permutes the coefficients.  Needs M % IL ==
// 0 and line starts at multiples of IL rows; IL = 0: five arrays.
// X16: the 16-system tile without pads (tile_put_x16), for tiles that
// would not fit two workgroups per CU with them.
template <int M, int R, int Q, int BLOCK = 256, bool SP = true, int SR = ir_stored_rounds(R),
          int IL = 0, bool X16 = false>
__global__ void __launch_bounds__(BLOCK, 512 / BLOCK)
    k_ir_xq(const float *__restrict__ ds, const float *__restrict__ dl,
            const float *__restrict__ d, const float *__restrict__ du,
            const float *__restrict__ dw, const double *bin, double *out,
            int64_t n_sys, int64_t len, int max_steps, double tol,
            int *__restrict__ steps_out, XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int L = 1 << R;
  constexpr int s0shift = (BLOCK == 512 ? 9 : 8) - R; // systems per pass
  const int64_t blk = xcd_tile(blockIdx.x, gridDim.x);
  const int sl = (int)threadIdx.x >> R;
  const int p = (int)(threadIdx.x & (L - 1));
  const int lane = threadIdx.x & 63;
  WaveEx ex{lane, p, L};
  int kmax = 0;
#pragma unroll 1
  for (int q = 0; q < Q; ++q) {
    const int64_t sys = ((blk * Q + q) << s0shift) + sl;
    const int64_t sc = sys < n_sys ? sys : n_sys - 1;
    const int64_t slo = map.lo(sc), shi = map.hi(sc);
    const int64_t off = slo * map.in_lo + shi * map.in_hi + (int64_t)p * M;
    const int64_t offc = slo * map.c_lo + shi * map.c_hi + (int64_t)p * M;
    float s[M], l[M], dd[M], u[M], w[M];
    if constexpr (IL == 2) {
      static_assert(M % 2 == 0, "interleaved coefficients come in groups of 2 rows");
      const float2 *g = reinterpret_cast<const float2 *>(ds) + (offc >> 1) * 5;
#pragma unroll
      for (int k = 0; k < M / 2; ++k) {
        const float2 a0 = g[5 * k], a1 = g[5 * k + 1], a2 = g[5 * k + 2], a3 = g[5 * k + 3],
                     a4 = g[5 * k + 4];
        s[2 * k] = a0.x; s[2 * k + 1] = a0.y;
        l[2 * k] = a1.x; l[2 * k + 1] = a1.y;
        dd[2 * k] = a2.x; dd[2 * k + 1] = a2.y;
        u[2 * k] = a3.x; u[2 * k + 1] = a3.y;
        w[2 * k] = a4.x; w[2 * k + 1] = a4.y;
      }
    } else if constexpr (IL == 4) {
      static_assert(M % 4 == 0, "interleaved coefficients come in groups of 4 rows");
      const float4 *g = reinterpret_cast<const float4 *>(ds) + (offc >> 2) * 5;
#pragma unroll
      for (int k = 0; k < M / 4; ++k) {
        const float4 a0 = g[5 * k], a1 = g[5 * k + 1], a2 = g[5 * k + 2], a3 = g[5 * k + 3],
                     a4 = g[5 * k + 4];
        s[4 * k] = a0.x; s[4 * k + 1] = a0.y; s[4 * k + 2] = a0.z; s[4 * k + 3] = a0.w;
        l[4 * k] = a1.x; l[4 * k + 1] = a1.y; l[4 * k + 2] = a1.z; l[4 * k + 3] = a1.w;
        dd[4 * k] = a2.x; dd[4 * k + 1] = a2.y; dd[4 * k + 2] = a2.z; dd[4 * k + 3] = a2.w;
        u[4 * k] = a3.x; u[4 * k + 1] = a3.y; u[4 * k + 2] = a3.z; u[4 * k + 3] = a3.w;
        w[4 * k] = a4.x; w[4 * k + 1] = a4.y; w[4 * k + 2] = a4.z; w[4 * k + 3] = a4.w;
      }
    } else {
      RowsIO<float, M>::load(ds + offc, s);
      RowsIO<float, M>::load(dl + offc, l);
      RowsIO<float, M>::load(d + offc, dd);
      RowsIO<float, M>::load(du + offc, u);
      RowsIO<float, M>::load(dw + offc, w);
    }
    ir_zero_out_of_range<M>(0, s, p == 0, p == L - 1);
    ir_zero_out_of_range<M>(1, l, p == 0, p == L - 1);
    ir_zero_out_of_range<M>(3, u, p == 0, p == L - 1);
    ir_zero_out_of_range<M>(4, w, p == 0, p == L - 1);
    const RegCoef<M> A{s, l, dd, u, w};
    const MemRhs b{bin + off, 1};
    double xs[M];
    const int k = ir_solve<M, SR, SP>(A, b, xs, p, L, ex, max_steps, tol);
    kmax = k > kmax ? k : kmax;
    if constexpr (X16) {
      static_assert((Q << s0shift) == 16, "X16 tiles hold 16 systems");
      tile_put_x16<double, M>(reinterpret_cast<double *>(smem_raw), xs, p, (q << s0shift) + sl);
    } else {
      tile_put<double, M>(reinterpret_cast<double *>(smem_raw), xs, p, Q << s0shift,
                          (q << s0shift) + sl);
    }
  }
  __syncthreads();
  constexpr int sshift = s0shift + (Q == 4 ? 2 : (Q == 2 ? 1 : 0));
  if constexpr (X16)
    tile_flush_x16<double, M>(reinterpret_cast<const double *>(smem_raw), (int)len, out, n_sys,
                              blk << sshift, map);
  else
    tile_flush<double, M>(reinterpret_cast<const double *>(smem_raw), sshift, (int)len, out,
                          n_sys, blk << sshift, map);
  if (lane == 0) note_steps(steps_out, kmax);
}

template <int M, int R, int MAXT>
__global__ void __launch_bounds__(MAXT)
    k_ir_yz(const float *__restrict__ ds, const float *__restrict__ dl,
            const float *__restrict__ d, const float *__restrict__ du,
            const float *__restrict__ dw, double *__restrict__ x, Lines ln,
            int S, int P, int max_steps, double tol, int *__restrict__ steps_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int t = threadIdx.x;
  const int sl = t % S;
  const int p = t / S;
  const int64_t tile = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t sys = tile * S + sl;
  const bool valid = sys < ln.n_sys;
  const int64_t st = ln.stride;
  const int64_t b0 = ln.base(valid ? sys : ln.n_sys - 1) + (int64_t)p * M * st;
  float s[M], l[M], dd[M], u[M], w[M];
  double b[M], xs[M];
  const int len = (int)ln.len;
  if (P * M == len) {
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const int64_t e = b0 + i * st;
      s[i] = ds[e];
      l[i] = dl[e];
      dd[i] = d[e];
      u[i] = du[e];
      w[i] = dw[e];
      b[i] = x[e];
    }
  } else {
    // padded length: identity rows past the end, couplings past it zero
#pragma unroll
    for (int i = 0; i < M; ++i) {
      const int r = p * M + i;
      const int64_t e = b0 + i * st;
      const bool in = r < len;
      s[i] = in ? ds[e] : 0.0f;
      l[i] = in ? dl[e] : 0.0f;
      dd[i] = in ? d[e] : 1.0f;
      u[i] = (r + 1 < len) ? du[e] : 0.0f;
      w[i] = (r + 2 < len) ? dw[e] : 0.0f;
      b[i] = in ? x[e] : 0.0;
    }
  }
  ir_zero_out_of_range<M>(0, s, p == 0, p == P - 1);
  ir_zero_out_of_range<M>(1, l, p == 0, p == P - 1);
  ir_zero_out_of_range<M>(3, u, p == 0, p == P - 1);
  ir_zero_out_of_range<M>(4, w, p == 0, p == P - 1);
  LdsEx ex{smem_raw, sl, p, S, P};
  const RegCoef<M> A{s, l, dd, u, w};
  const RegRhs<M> bv{b};
  const int k = ir_solve<M, ir_stored_rounds(R)>(A, bv, xs, p, P, ex, max_steps, tol);
  if (valid) {
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (p * M + i < len) x[b0 + i * st] = xs[i];
  }
  if (t == 0) note_steps(steps_out, k);
}

// ---------------------------------------------------------------------------
// x direction, any length: the S = blockDim / L consecutive systems of the
// workgroup are one contiguous span of S * len elements per array; the span
// of each coefficient array is read with coalesced loads straight into the
// per-wavefront LDS layout [5][M][64] the refinement reads (chunk rows past
// a system's end stay identity rows, couplings past either end are zero).
// The right-hand side is re-read per residual from the caller's x, which is
// written only at the end.
// ---------------------------------------------------------------------------
template <int M, int R>
__global__ void __launch_bounds__(256, 2) // 3 waves per SIMD would spill for R >= 4
    k_ir_tr_x(const float *__restrict__ ds, const float *__restrict__ dl,
              const float *__restrict__ d, const float *__restrict__ du,
              const float *__restrict__ dw, double *x, int64_t n_sys, int len,
              int max_steps, double tol, int *__restrict__ steps_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  constexpr int L = 1 << R;
  constexpr int REG = 5 * M * 64;            // floats per wavefront region
  float *lds = reinterpret_cast<float *>(smem_raw);
  const int S = (int)blockDim.x >> R;
  const int64_t sys0 = (int64_t)blockIdx.x * S;
  const int ns = (int)((n_sys - sys0) < S ? (n_sys - sys0) : S);
  const int64_t base = sys0 * len;
  const int span = ns * len;
  // identity everywhere first (padding rows, absent systems)
  for (int e = threadIdx.x; e < ((int)blockDim.x >> 6) * REG; e += blockDim.x) {
    const int k = (e % REG) / (M * 64);
    lds[e] = (k == 2) ? 1.0f : 0.0f;
  }
  __syncthreads();
  const float *src[5] = {ds, dl, d, du, dw};
  const float inv_len = 1.0f / (float)len;
#pragma unroll 1
  for (int k = 0; k < 5; ++k) {
    for (int e = threadIdx.x; e < span; e += blockDim.x) {
      int sl = (int)((float)e * inv_len);   // e / len, then corrected
      sl -= (sl * len > e);
      sl += ((sl + 1) * len <= e);
      const int r = e - sl * len;
      const int t = sl * L + r / M, i = r % M; // owning thread, row in its chunk
      float v = src[k][base + e];
      if ((k == 0 && r < 2) || (k == 1 && r < 1) || (k == 3 && r + 1 >= len) ||
          (k == 4 && r + 2 >= len))
        v = 0.0f;
      lds[(t >> 6) * REG + (k * M + i) * 64 + (t & 63)] = v;
    }
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sl = (int)threadIdx.x >> R, p = (int)threadIdx.x & (L - 1);
  const bool act = sl < ns;
  const int r0 = p * M;
  const int nv = act ? (len - r0 <= 0 ? 0 : (len - r0 >= M ? M : len - r0)) : 0;
  double *xrow = x + base + (act ? (int64_t)sl * len + r0 : 0);
  WaveEx ex{lane, p, L};
  const LdsCoef<M> A{lds + wave * REG + lane};
  const VarRhsX b{xrow, nv};
  double xs[M];
  const int k = ir_solve<M, ir_stored_rounds(R)>(A, b, xs, p, L, ex, max_steps, tol);
#pragma unroll
  for (int i = 0; i < M; ++i)
    if (i < nv) xrow[i] = xs[i];
  if (lane == 0) note_steps(steps_out, k);
}

// ---------------------------------------------------------------------------
// fallback: one lane per system, factors in scratch [5][len][n_sys] (FP32)
// plus a copy of b [len][n_sys] (FP64).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
    k_ir_lane(const float *__restrict__ ds, const float *__restrict__ dl,
              const float *__restrict__ d, const float *__restrict__ du,
              const float *__restrict__ dw, double *__restrict__ x, Lines ln,
              float *__restrict__ fs, double *__restrict__ bs, int max_steps,
              double tol, int *__restrict__ steps_out) {
  const int64_t sys = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (sys >= ln.n_sys) return;
  const int64_t n = ln.len, st = ln.stride, ns = ln.n_sys;
  const int64_t b0 = ln.base(sys);
  float *du2 = fs + sys, *dw2 = du2 + n * ns, *ddl = dw2 + n * ns,
        *ivs = ddl + n * ns, *z = ivs + n * ns;
  double *bb = bs + sys;
  // factor + first solve (FP32)
  double xmax = 0.0;
  {
    float u1 = 0, w1 = 0, z1 = 0, u0 = 0, w0 = 0, z0 = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t e = b0 + i * st;
      const double bi = x[e];
      bb[i * ns] = bi;
      const float s = (i >= 2) ? ds[e] : 0.0f;
      const float lv = (i >= 1) ? dl[e] : 0.0f;
      const float dl_ = lv - s * u1;
      const float iv = recip(d[e] - s * w1 - dl_ * u0);
      // couplings past the end of the line are out of range: never read
      const float un = (i + 1 < n) ? (du[e] - dl_ * w0) * iv : 0.0f;
      const float wn = (i + 2 < n) ? dw[e] * iv : 0.0f;
      const float zn = ((float)bi - s * z1 - dl_ * z0) * iv;
      du2[i * ns] = un;
      dw2[i * ns] = wn;
      ddl[i * ns] = dl_;
      ivs[i * ns] = iv;
      z[i * ns] = zn;
      u1 = u0; w1 = w0; z1 = z0;
      u0 = un; w0 = wn; z0 = zn;
    }
    float xa = 0, xb = 0;
    for (int64_t i = n - 1; i >= 0; --i) {
      const float xi = z[i * ns] - du2[i * ns] * xa - dw2[i * ns] * xb;
      x[b0 + i * st] = (double)xi;
      xmax = fmax(xmax, fabs((double)xi));
      xb = xa;
      xa = xi;
    }
  }
  int k = 0;
  double dprev = xmax; // dx_0 := x_0
  while (k < max_steps) {
    // residual + forward substitution with the stored factors
    float z1 = 0, z0 = 0;
    for (int64_t i = 0; i < n; ++i) {
      const int64_t e = b0 + i * st;
      double acc = (double)d[e] * x[e];
      if (i >= 2) acc += (double)ds[e] * x[e - 2 * st];
      if (i >= 1) acc += (double)dl[e] * x[e - st];
      if (i + 1 < n) acc += (double)du[e] * x[e + st];
      if (i + 2 < n) acc += (double)dw[e] * x[e + 2 * st];
      const float r = (float)(bb[i * ns] - acc);
      const float s = (i >= 2) ? ds[e] : 0.0f;
      const float zn = (r - s * z1 - ddl[i * ns] * z0) * ivs[i * ns];
      z[i * ns] = zn;
      z1 = z0;
      z0 = zn;
    }
    float xa = 0, xb = 0;
    double dxm = 0.0;
    for (int64_t i = n - 1; i >= 0; --i) {
      const float xi = z[i * ns] - du2[i * ns] * xa - dw2[i * ns] * xb;
      x[b0 + i * st] += (double)xi;
      dxm = fmax(dxm, fabs((double)xi));
      xb = xa;
      xa = xi;
    }
    ++k;
    // estimated forward error of the new iterate: |dx_k|^2 / |dx_{k-1}|
    const bool conv = (dxm == 0.0) || (dxm * dxm <= tol * xmax * dprev);
    dprev = dxm;
    if (conv) break;
  }
  note_steps(steps_out, k);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {

int ceil_log2(int64_t v) {
  int r = 0;
  while (((int64_t)1 << r) < v) ++r;
  return r;
}

template <int M, int R>
hipError_t go_ir_x(const Diags<float> &a, const double *b, double *out, bool rot,
                   int64_t n_sys, int64_t len, const XMap &map, int max_steps,
                   double tol, int *steps, hipStream_t st) {
  const int block = 256; // 512 measured slower (fewer resident workgroups)
  // passes per workgroup so that a tile row is a whole 128-byte line
  if constexpr (M <= 8) if (rot) {
    constexpr int run = (256 >> R) * 8;
    constexpr int Q = run >= 128 ? 1 : 4;
    const int64_t per_wg = (int64_t)(256 >> R) * Q;
    const int64_t grid = (n_sys + per_wg - 1) / per_wg;
    const size_t smem = rotated_tile_bytes<double>(block * Q, M, len);
    if (smem > 65536) {
      const hipError_t e = hipFuncSetAttribute(
          reinterpret_cast<const void *>(&k_ir_xq<M, R, Q>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_ir_xq<M, R, Q>), dim3((unsigned)grid), dim3(block), smem, st, a.ds,
                       a.dl, a.d, a.du, a.dw, b, out, n_sys, len, max_steps, tol, steps, map);
    return hipGetLastError();
  }
  // 10 rows per lane: spikes recomputed, two stored PCR rounds (255 VGPRs,
  // no spill; three stored rounds or kept spikes spill).  640-row lines (64
  // lanes): 256-thread workgroups, four passes into an unpadded 16-system
  // tile (80 KB; 82 KB with the pads, too much for two per CU), two
  // workgroups per CU as for 8-row chunks (384x256x640 hyperdiffusion
  // direction: 0.453 -> 0.436 ms against the 512-thread kernel below, one
  // workgroup per CU; the same results bit for bit)
  if constexpr (M == 10 && R == 6) if (rot) {
    constexpr int Q = 4;
    const int64_t per_wg = 16;
    const int64_t grid = (n_sys + per_wg - 1) / per_wg;
    const size_t smem = sizeof(double) * 16 * (size_t)len;
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_ir_xq<M, R, Q, 256, false, 2, 0, true>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ir_xq<M, R, Q, 256, false, 2, 0, true>), dim3((unsigned)grid),
                       dim3(block), smem, st, a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len,
                       max_steps, tol, steps, map);
    return hipGetLastError();
  }
  // other 10-row lines: 512-thread workgroups
  if constexpr (M == 10) if (rot) {
    constexpr int B = 512;
    constexpr int run = (B >> R) * 8;
    constexpr int Q = run >= 128 ? 1 : (run >= 64 ? 2 : 4);
    const int64_t per_wg = (int64_t)(B >> R) * Q;
    const int64_t grid = (n_sys + per_wg - 1) / per_wg;
    const size_t smem = rotated_tile_bytes<double>(B * Q, M, len);
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_ir_xq<M, R, Q, B, false, 2>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_ir_xq<M, R, Q, B, false, 2>), dim3((unsigned)grid), dim3(B), smem, st,
                       a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len, max_steps, tol, steps,
                       map);
    return hipGetLastError();
  }
  const int64_t threads = n_sys << R;
  const int64_t grid = (threads + block - 1) / block;
  const size_t smem_a = (size_t)(block / 64) * 5 * M * 64 * sizeof(float);
  if (rot) {
    const size_t tile = rotated_tile_bytes<double>(block, M, len);
    const size_t smem = smem_a > tile ? smem_a : tile;
    if (smem > 65536) {
      const hipError_t e = hipFuncSetAttribute(
          reinterpret_cast<const void *>(&k_ir_x<M, R, true>),
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((k_ir_x<M, R, true>), dim3((unsigned)grid), dim3(block), smem, st,
                       a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len, max_steps, tol, steps,
                       map);
    return hipGetLastError();
  }
  if (smem_a > 65536) {
    const hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&k_ir_x<M, R, false>),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem_a);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_ir_x<M, R, false>), dim3((unsigned)grid), dim3(block), smem_a, st,
                     a.ds, a.dl, a.d, a.du, a.dw, b, out, n_sys, len, max_steps, tol, steps,
                     map);
  return hipGetLastError();
}

// the rotated-store kernel over interleaved coefficients (k_ir_xq, IL):
// groups of 4 rows for 4- and 8-row chunks, of 2 rows for 6-row chunks and
// for the 10-row chunks of 640-row lines (the unpadded 16-system tile)
template <int M, int R>
hipError_t go_ir_x_il(const float *coef, const double *b, double *out, int64_t n_sys,
                      int64_t len, const XMap &map, int max_steps, double tol, int *steps,
                      hipStream_t st) {
  const int block = 256;
  auto go = [&](auto kernel, int64_t per_wg, size_t smem) {
    if (smem > 65536) {
      const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               (int)smem);
      if (e != hipSuccess) return e;
    }
    const int64_t grid = (n_sys + per_wg - 1) / per_wg;
    hipLaunchKernelGGL(kernel, dim3((unsigned)grid), dim3(block), smem, st, coef, coef, coef,
                       coef, coef, b, out, n_sys, len, max_steps, tol, steps, map);
    return hipGetLastError();
  };
  if constexpr (M == 10) {
    if constexpr (R == 6)
      return go(k_ir_xq<M, R, 4, 256, false, 2, 2, true>, 16, sizeof(double) * 16 * (size_t)len);
    else
      return hipErrorInvalidValue;
  } else if constexpr (M <= 8) {
    constexpr int G = M % 4 == 0 ? 4 : 2;
    constexpr int run = (256 >> R) * 8;
    constexpr int Q = run >= 128 ? 1 : 4;
    constexpr int SR = ir_stored_rounds(R);
    return go(k_ir_xq<M, R, Q, 256, true, SR, G>, (int64_t)(256 >> R) * Q,
              rotated_tile_bytes<double>(block * Q, M, len));
  } else {
    return hipErrorInvalidValue;
  }
}

// IR keeps ~ (5 FP32 + 2 FP64 + 6 FP32 factors) per row plus 12 floats per
// PCR round in registers: workgroups of up to 512 threads (2 waves/SIMD).
constexpr int IR_YZ_MAXT = 512;

template <int M, int R>
hipError_t go_ir_yz(const Diags<float> &a, double *x, const Lines &ln, int S,
                    int max_steps, double tol, int *steps, hipStream_t st) {
  const int P = (int)((ln.len + M - 1) / M); // a partial last chunk is padded
  const int block = S * P;
  if (block > IR_YZ_MAXT) return hipErrorInvalidConfiguration;
  const size_t smem = (size_t)8 * S * P * sizeof(float) > (size_t)4 * S * P * sizeof(double)
                          ? (size_t)8 * S * P * sizeof(float)
                          : (size_t)4 * S * P * sizeof(double);
  const int64_t grid = (ln.n_sys + S - 1) / S;
  hipLaunchKernelGGL((k_ir_yz<M, R, IR_YZ_MAXT>), dim3((unsigned)grid), dim3(block),
                     smem, st, a.ds, a.dl, a.d, a.du, a.dw, x, ln, S, P, max_steps,
                     tol, steps);
  return hipGetLastError();
}

} // namespace

bool ir_x_supported(int rows, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 10) return false;
  if (len % rows) return false;
  const int64_t L = len / rows;
  return L >= 2 && L <= 64 && (L & (L - 1)) == 0;
}

bool ir_yz_supported(int rows, int sys_per_wg, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 10) return false;
  if (len < 3) return false;
  const int64_t P = (len + rows - 1) / rows; // a partial last chunk is padded
  if (P < 1 || P > 128) return false;
  const int64_t block = P * sys_per_wg;
  return block <= IR_YZ_MAXT;
}

hipError_t launch_ir_x_map(const Diags<float> &a, const double *b, double *out,
                           bool rot, int64_t n_sys, int64_t len, const XMap &map,
                           int rows, int max_steps, double tol, int *steps,
                           hipStream_t st) {
  if (!ir_x_supported(rows, len)) return hipErrorInvalidValue;
  const int R = ceil_log2(len / rows);
#define IRX(MM, RR) \
  if (rows == MM && R == RR) \
    return go_ir_x<MM, RR>(a, b, out, rot, n_sys, len, map, max_steps, tol, steps, st);
  IRX(4, 1) IRX(4, 2) IRX(4, 3) IRX(4, 4) IRX(4, 5) IRX(4, 6)
  IRX(6, 1) IRX(6, 2) IRX(6, 3) IRX(6, 4) IRX(6, 5) IRX(6, 6)
  IRX(8, 1) IRX(8, 2) IRX(8, 3) IRX(8, 4) IRX(8, 5) IRX(8, 6)
  IRX(10, 1) IRX(10, 2) IRX(10, 3) IRX(10, 4) IRX(10, 5) IRX(10, 6)
#undef IRX
  return hipErrorInvalidValue;
}

int ir_x_interleave_group(int rows, int64_t len) {
  if (!ir_x_supported(rows, len)) return 0;
  if (rows == 4 || rows == 8) return len % 4 == 0 ? 4 : 0;
  if (rows == 6) return 2;
  if (rows == 10) return len == 640 ? 2 : 0; // 64 lanes: the unpadded-tile kernel
  return 0;
}

hipError_t launch_ir_x_map_il(const float *coef, const double *b, double *out, int64_t n_sys,
                              int64_t len, const XMap &map, int rows, int max_steps, double tol,
                              int *steps, hipStream_t st) {
  // every line must start on a group of rows
  const int G = ir_x_interleave_group(rows, len);
  if (G == 0 || map.c_lo % G || map.c_hi % G || (reinterpret_cast<uintptr_t>(coef) & 15u))
    return hipErrorInvalidValue;
  const int R = ceil_log2(len / rows);
#define IRXI(MM, RR) \
  if (rows == MM && R == RR) \
    return go_ir_x_il<MM, RR>(coef, b, out, n_sys, len, map, max_steps, tol, steps, st);
  IRXI(4, 1) IRXI(4, 2) IRXI(4, 3) IRXI(4, 4) IRXI(4, 5) IRXI(4, 6)
  IRXI(6, 1) IRXI(6, 2) IRXI(6, 3) IRXI(6, 4) IRXI(6, 5) IRXI(6, 6)
  IRXI(8, 1) IRXI(8, 2) IRXI(8, 3) IRXI(8, 4) IRXI(8, 5) IRXI(8, 6)
  IRXI(10, 6)
#undef IRXI
  return hipErrorInvalidValue;
}

hipError_t launch_ir_x(const Diags<float> &a, double *x, const Lines &ln, int rows,
                       int max_steps, double tol, int *steps, hipStream_t st) {
  return launch_ir_x_map(a, x, x, false, ln.n_sys, ln.len, xmap_inplace(ln), rows,
                         max_steps, tol, steps, st);
}

hipError_t launch_ir_yz(const Diags<float> &a, double *x, const Lines &ln, int rows,
                        int sys_per_wg, int max_steps, double tol, int *steps,
                        hipStream_t st) {
  if (!ir_yz_supported(rows, sys_per_wg, ln.len)) return hipErrorInvalidValue;
  const int R = ceil_log2((ln.len + rows - 1) / rows);
#define IRYZ(MM, RR) \
  if (rows == MM && R == RR) return go_ir_yz<MM, RR>(a, x, ln, sys_per_wg, max_steps, tol, steps, st);
  IRYZ(4, 0) IRYZ(4, 1) IRYZ(4, 2) IRYZ(4, 3) IRYZ(4, 4) IRYZ(4, 5) IRYZ(4, 6) IRYZ(4, 7)
  IRYZ(6, 0) IRYZ(6, 1) IRYZ(6, 2) IRYZ(6, 3) IRYZ(6, 4) IRYZ(6, 5) IRYZ(6, 6) IRYZ(6, 7)
  IRYZ(8, 0) IRYZ(8, 1) IRYZ(8, 2) IRYZ(8, 3) IRYZ(8, 4) IRYZ(8, 5) IRYZ(8, 6) IRYZ(8, 7)
  IRYZ(10, 0) IRYZ(10, 1) IRYZ(10, 2) IRYZ(10, 3) IRYZ(10, 4) IRYZ(10, 5) IRYZ(10, 6) IRYZ(10, 7)
#undef IRYZ
  return hipErrorInvalidValue;
}

bool ir_tr_x_supported(int rows, int64_t len) {
  if (rows != 8) return false;
  int lanes = 1;
  while ((int64_t)lanes * rows < len) lanes <<= 1;
  return len >= 3 && lanes <= 64;
}

hipError_t launch_ir_tr_x(const Diags<float> &a, double *x, const Lines &ln, int rows,
                          int max_steps, double tol, int *steps, hipStream_t st) {
  if (!ir_tr_x_supported(rows, ln.len) || ln.dir != 0) return hipErrorInvalidValue;
  int R = 0;
  while (((int64_t)8 << R) < ln.len) ++R;
  const int block = 256, S = block >> R;
  const int64_t grid = (ln.n_sys + S - 1) / S;
  const size_t smem = (size_t)(block / 64) * 5 * 8 * 64 * sizeof(float);
#define IRT(RR) \
  if (R == RR) { \
    hipLaunchKernelGGL((k_ir_tr_x<8, RR>), dim3((unsigned)grid), dim3(block), smem, st, a.ds, \
                       a.dl, a.d, a.du, a.dw, x, ln.n_sys, (int)ln.len, max_steps, tol, steps); \
    return hipGetLastError(); \
  }
  IRT(0) IRT(1) IRT(2) IRT(3) IRT(4) IRT(5) IRT(6)
#undef IRT
  return hipErrorInvalidValue;
}

size_t ir_lane_scratch_bytes(const Lines &ln) {
  const size_t n = (size_t)ln.n_sys * (size_t)ln.len;
  return n * (5 * sizeof(float) + sizeof(double));
}

hipError_t launch_ir_lane(const Diags<float> &a, double *x, const Lines &ln,
                          void *scratch, int max_steps, double tol, int *steps,
                          hipStream_t st) {
  const size_t n = (size_t)ln.n_sys * (size_t)ln.len;
  double *bs = static_cast<double *>(scratch);
  float *fs = reinterpret_cast<float *>(bs + n);
  const int block = 256;
  const int64_t grid = (ln.n_sys + block - 1) / block;
  hipLaunchKernelGGL(k_ir_lane, dim3((unsigned)grid), dim3(block), 0, st, a.ds, a.dl,
                     a.d, a.du, a.dw, x, ln, fs, bs, max_steps, tol, steps);
  return hipGetLastError();
}

} // namespace pentad
