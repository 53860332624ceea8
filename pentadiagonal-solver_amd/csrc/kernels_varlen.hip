// kernels_varlen.hip -- many systems of different lengths in one launch.
//
// Systems are packed back to back: system k owns rows offs[k]..offs[k+1]-1
// of every array.  The host sorts the systems by length (longest first) and
// cuts the sorted list into work items; an item is S systems that share one
// chunk count P (a power of two, P * M >= the longest length of the item,
// S * P = VL_BLOCK threads).  Thread t of the workgroup owns chunk
// p = t % P of system t / P: the chunks of one system are consecutive
// lanes, so its rows are read as one contiguous span.  Rows past a system's
// length are virtual identity rows (d = 1, no couplings, zero right-hand
// side), which leaves the real rows' solution unchanged; couplings of real
// rows past the end are zeroed.  The workgroups are persistent and take the
// items in list order from a device counter (reset by the launcher).  The
// host lists the longest quarter of the items alternating with the
// shortest, then the rest longest first (pentad_varlen.cpp, claim_order):
// long items start first and short ones fill in behind them, while the
// mixed sizes keep the CUs from loading and solving in step.  (Dealing item
// b + kG to workgroup b instead left a tail: 1035 single-system items on
// 256 one-workgroup CUs took 0.21 ms, 1024 took 0.15 ms.)
//
// Exchanges: P <= 64 keeps a system inside one wavefront (DPP / shuffles);
// longer systems exchange through LDS across the workgroup.
#include "ir_core.hpp"
#include "partition_core.hpp"
#include "partition_exchange.hpp"
#include "pentad_internal.hpp"
#include "rows_io.hpp"

namespace pentad {

// exchange through LDS between the chunks of one system that are the
// consecutive threads [sys * P, sys * P + P) of the workgroup
struct VarEx {
  unsigned char *raw; // K * N elements
  int t, p, P, N;
  bool synced = false; // the last workgroup-wide step was all() (as LdsEx)
  float *red = nullptr; // N / 64 floats of LDS apart from raw (max_f32), or none
  template <int K, typename T>
  __device__ __forceinline__ void swapc(const T *send, T *from_prev, T *from_next,
                                        int h) {
    T *lds = reinterpret_cast<T *>(raw);
    if (!synced) __syncthreads();
    synced = false;
#pragma unroll
    for (int k = 0; k < K; ++k) lds[k * N + t] = send[k];
    __syncthreads();
    const int op = (p - h) >= 0 ? t - h : t;
    const int on = (p + h) < P ? t + h : t;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      from_prev[k] = lds[k * N + op];
      from_next[k] = lds[k * N + on];
    }
  }
  __device__ __forceinline__ bool all(bool pred) {
    const bool r = __syncthreads_and(pred) != 0;
    synced = true;
    return r;
  }
  // With red and systems of whole wavefronts (P >= 64): the maximum of each
  // wavefront by DPP, one slot per wavefront in LDS, and every thread takes
  // the maximum of its system's P / 64 slots -- two barriers, where the
  // exchange rounds below take two per round (18 for 512 chunks).
  __device__ __forceinline__ float max_f32(float v) {
    if (red != nullptr && P >= 64) {
      v = WaveEx{t & 63, t & 63, 64}.max_f32(v);
      if (!synced) __syncthreads(); // the slots' previous readers are done
      if ((t & 63) == 0) red[t >> 6] = v;
      __syncthreads();
      const int w0 = (t - p) >> 6; // the system's first wavefront
      float m = red[w0];
      for (int k = 1; k < (P >> 6); ++k) m = fmaxf(m, red[w0 + k]);
      synced = false;
      return m;
    }
    for (int h = 1; h < P; h <<= 1) {
      float a[1] = {v}, bp[1], bn[1];
      swapc<1>(a, bp, bn, h);
      v = fmaxf(v, fmaxf(bp[0], bn[0]));
    }
    return v;
  }
};

// The exchange of k_varlen: within a wavefront (WaveEx) when a system's
// chunks fit one (P <= 64), else through LDS (VarEx).  The choice is the
// item's, so it is uniform over the workgroup and every lane reaches the LDS
// branch's barriers; one inlined solve serves both kinds of item (a solve
// inlined per exchange kind spilled 148 bytes per lane in FP64, this one
// 12).  The refinement kernel keeps one solve per kind: with this exchange it
// spilled more (320 against 292 bytes).
struct MixEx {
  WaveEx wv;
  VarEx lx;
  bool lds;
  template <int K, typename T>
  __device__ __forceinline__ void swapc(const T *send, T *from_prev, T *from_next, int h) {
    if (lds)
      lx.template swapc<K>(send, from_prev, from_next, h);
    else
      wv.template swapc<K>(send, from_prev, from_next, h);
  }
  __device__ __forceinline__ bool all(bool pred) { return lds ? lx.all(pred) : wv.all(pred); }
};

// right-hand side rows of a chunk; rows past the end of the system are 0
struct VarRhs {
  const double *p;
  int nv; // real rows in this chunk
  __device__ __forceinline__ double operator()(int i) const { return i < nv ? p[i] : 0.0; }
};

namespace {

// the next item for the whole workgroup (thread 0 claims it); every
// thread sees the same value (read into a scalar register, so the item's
// fields and the exchange choice are uniform branches), and all leave the
// loop together
__device__ __forceinline__ int claim_item(int *next, int *slot) {
  if (threadIdx.x == 0) *slot = atomicAdd(next, 1);
  __syncthreads();
  const int it = __builtin_amdgcn_readfirstlane(*slot);
  __syncthreads();
  return it;
}

struct Item {
  int first, count, pshift;
};
__device__ __forceinline__ Item item_at(const int2 *items, int it) {
  const int2 v = items[it];
  return Item{v.x, v.y & 0xFFFF, v.y >> 16};
}

// a chunk starting at element e of an array of C can use 16-byte vectors
template <typename C, int M>
__device__ __forceinline__ bool chunk_vec(int64_t e) {
  return RowsIO<C, M>::VB == 16 && ((e * (int64_t)sizeof(C)) & 15) == 0;
}

// rows [r0, r0 + M) of a system of length len starting at element off;
// VEC: 16-byte loads for whole aligned chunks (the refinement kernel, whose
// registers are full, keeps the scalar path: its vector path spills more)
template <typename T, typename C, int M, bool VEC>
__device__ __forceinline__ void load_rows(const C *__restrict__ ds, const C *__restrict__ dl,
                                          const C *__restrict__ d, const C *__restrict__ du,
                                          const C *__restrict__ dw, int64_t off, int64_t r0,
                                          int64_t len, T (&s)[M], T (&l)[M], T (&dd)[M],
                                          T (&u)[M], T (&w)[M]) {
  if (VEC && chunk_vec<C, M>(off + r0) && r0 + M <= len) {
    // whole chunk inside the system, 16-byte aligned: vector loads, then
    // zero the couplings that reach past either end
    C t[M];
    const int64_t e = off + r0;
    RowsIO<C, M>::load(ds + e, t);
#pragma unroll
    for (int i = 0; i < M; ++i) s[i] = (r0 + i >= 2) ? (T)t[i] : T(0);
    RowsIO<C, M>::load(dl + e, t);
#pragma unroll
    for (int i = 0; i < M; ++i) l[i] = (r0 + i >= 1) ? (T)t[i] : T(0);
    RowsIO<C, M>::load(d + e, t);
#pragma unroll
    for (int i = 0; i < M; ++i) dd[i] = (T)t[i];
    RowsIO<C, M>::load(du + e, t);
#pragma unroll
    for (int i = 0; i < M; ++i) u[i] = (r0 + i + 1 < len) ? (T)t[i] : T(0);
    RowsIO<C, M>::load(dw + e, t);
#pragma unroll
    for (int i = 0; i < M; ++i) w[i] = (r0 + i + 2 < len) ? (T)t[i] : T(0);
    return;
  }
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int64_t r = r0 + i, e = off + r;
    const bool in = r < len;
    s[i] = (in && r >= 2) ? (T)ds[e] : T(0);
    l[i] = (in && r >= 1) ? (T)dl[e] : T(0);
    dd[i] = in ? (T)d[e] : T(1);
    u[i] = (r + 1 < len) ? (T)du[e] : T(0);
    w[i] = (r + 2 < len) ? (T)dw[e] : T(0);
  }
}

} // namespace

// C: the coefficients' storage type (float with T = double: the FP32
// coefficients widened to FP64 on load, see launch_varlen_widened)
template <typename T, int M, typename C = T>
__global__ void __launch_bounds__(VL_BLOCK)
    k_varlen(const C *__restrict__ ds, const C *__restrict__ dl, const C *__restrict__ d,
             const C *__restrict__ du, const C *__restrict__ dw, T *__restrict__ x,
             const int64_t *__restrict__ offs, const int *__restrict__ perm,
             const int2 *__restrict__ items, int n_items, int *__restrict__ next) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int slot;
  const int t = threadIdx.x;
  for (int it = claim_item(next, &slot); it < n_items; it = claim_item(next, &slot)) {
    const Item I = item_at(items, it);
    const int P = 1 << I.pshift;
    const int sl = t >> I.pshift, p = t & (P - 1);
    const bool active = sl < I.count;
    int64_t off = 0, len = 0;
    if (active) {
      const int sys = perm[I.first + sl];
      off = offs[sys];
      len = offs[sys + 1] - off;
    }
    const int64_t r0 = (int64_t)p * M;
    T s[M], l[M], dd[M], u[M], w[M], f[M];
    load_rows<T, C, M, true>(ds, dl, d, du, dw, off, r0, len, s, l, dd, u, w);
    const bool whole = r0 + M <= len && chunk_vec<T, M>(off + r0);
    if (whole) {
      RowsIO<T, M>::load(x + off + r0, f);
    } else {
#pragma unroll
      for (int i = 0; i < M; ++i) f[i] = (r0 + i < len) ? x[off + r0 + i] : T(0);
    }
    {
      MixEx ex{WaveEx{t & 63, p, P}, VarEx{smem_raw, t, p, P, VL_BLOCK}, P > 64};
      part_solve<T, M>(s, l, dd, u, w, f, p, P, ex);
    }
    if (whole) {
      RowsIO<T, M>::store(x + off + r0, f);
    } else {
#pragma unroll
      for (int i = 0; i < M; ++i)
        if (r0 + i < len) x[off + r0 + i] = f[i];
    }
  }
}

// FP32 coefficients, FP64 right-hand side / solution, refinement on chip
// (ir_core.hpp with the block PCR redone in every apply).
template <int M>
__global__ void __launch_bounds__(VL_BLOCK)
    k_varlen_ir(const float *__restrict__ ds, const float *__restrict__ dl,
                const float *__restrict__ d, const float *__restrict__ du,
                const float *__restrict__ dw, double *x, const int64_t *__restrict__ offs,
                const int *__restrict__ perm, const int2 *__restrict__ items, int n_items,
                int *__restrict__ next, int max_steps, double tol,
                int *__restrict__ steps_out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int slot;
  const int t = threadIdx.x;
  for (int it = claim_item(next, &slot); it < n_items; it = claim_item(next, &slot)) {
    const Item I = item_at(items, it);
    const int P = 1 << I.pshift;
    const int sl = t >> I.pshift, p = t & (P - 1);
    const bool active = sl < I.count;
    int64_t off = 0, len = 0;
    if (active) {
      const int sys = perm[I.first + sl];
      off = offs[sys];
      len = offs[sys + 1] - off;
    }
    const int64_t r0 = (int64_t)p * M;
    float s[M], l[M], dd[M], u[M], w[M];
    load_rows<float, float, M, false>(ds, dl, d, du, dw, off, r0, len, s, l, dd, u, w);
    const int64_t nv = len - r0;
    const VarRhs b{x + off + r0, nv <= 0 ? 0 : (nv >= M ? M : (int)nv)};
    const RegCoef<M> A{s, l, dd, u, w};
    double xs[M];
    int k;
    if (P <= 64) {
      WaveEx ex{t & 63, p, P};
      k = ir_solve<M, 0, false>(A, b, xs, p, P, ex, max_steps, tol);
    } else {
      VarEx ex{smem_raw, t, p, P, VL_BLOCK};
      ex.red = reinterpret_cast<float *>(smem_raw + (size_t)10 * VL_BLOCK * sizeof(float));
      k = ir_solve<M, 0, false>(A, b, xs, p, P, ex, max_steps, tol);
    }
#pragma unroll
    for (int i = 0; i < M; ++i)
      if (r0 + i < len) x[off + r0 + i] = xs[i];
    if (active && p == 0 && steps_out != nullptr) {
      const int cur = __hip_atomic_load(steps_out, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (k > cur) atomicMax(steps_out, k);
    }
  }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {

// workgroups that are resident at once (more would only find the list
// empty), at most one per item
template <typename K>
int64_t vl_grid(int n_items, K kernel, size_t smem) {
  int dev = 0, cus = 256, per_cu = 1;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, VL_BLOCK, smem) == hipSuccess &&
      nb > 0)
    per_cu = nb;
  const int64_t cap = (int64_t)cus * per_cu;
  return n_items < cap ? n_items : cap;
}

} // namespace

template <typename T>
hipError_t launch_varlen(const Diags<T> &a, T *x, const int64_t *offs, const int *perm,
                         const int2 *items, int n_items, int *next, hipStream_t st) {
  if (n_items <= 0) return hipSuccess;
  if (next == nullptr) return hipErrorInvalidValue;
  const hipError_t e = launch_fill_int(next, 0, st);
  if (e != hipSuccess) return e;
  const size_t smem = (size_t)10 * VL_BLOCK * sizeof(T);
  const int64_t grid = vl_grid(n_items, k_varlen<T, VL_ROWS>, smem);
  hipLaunchKernelGGL((k_varlen<T, VL_ROWS>), dim3((unsigned)grid), dim3(VL_BLOCK), smem, st,
                     a.ds, a.dl, a.d, a.du, a.dw, x, offs, perm, items, n_items, next);
  return hipGetLastError();
}

hipError_t launch_varlen_ir(const Diags<float> &a, double *x, const int64_t *offs,
                            const int *perm, const int2 *items, int n_items, int *next,
                            int max_steps, double tol, int *steps, hipStream_t st) {
  if (n_items <= 0) return hipSuccess;
  if (next == nullptr) return hipErrorInvalidValue;
  const hipError_t e = launch_fill_int(next, 0, st);
  if (e != hipSuccess) return e;
  // largest exchange: 10 floats or 4 doubles per thread; then the slots of
  // the system maxima (one float per wavefront)
  const size_t smem = (size_t)10 * VL_BLOCK * sizeof(float) + (VL_BLOCK / 64) * sizeof(float);
  const int64_t grid = vl_grid(n_items, k_varlen_ir<VL_ROWS>, smem);
  hipLaunchKernelGGL((k_varlen_ir<VL_ROWS>), dim3((unsigned)grid), dim3(VL_BLOCK), smem, st,
                     a.ds, a.dl, a.d, a.du, a.dw, x, offs, perm, items, n_items, next,
                     max_steps, tol, steps);
  return hipGetLastError();
}

// FP32 coefficients, FP64 right-hand side and solution: one FP64 solve of
// the coefficients widened on load.  That is the system the refinement
// (launch_varlen_ir) converges to, reached in one pass over the same 36 B per
// row; the refinement kernel, whose FP32 factors, FP64 iterate and residual
// do not fit its registers (292 bytes per lane spilled), took 0.46 ms on the
// 8192-system batch where this takes about the FP64 kernel's 0.29 ms.
hipError_t launch_varlen_widened(const Diags<float> &a, double *x, const int64_t *offs,
                                 const int *perm, const int2 *items, int n_items, int *next,
                                 hipStream_t st) {
  if (n_items <= 0) return hipSuccess;
  if (next == nullptr) return hipErrorInvalidValue;
  const hipError_t e = launch_fill_int(next, 0, st);
  if (e != hipSuccess) return e;
  const size_t smem = (size_t)10 * VL_BLOCK * sizeof(double);
  const int64_t grid = vl_grid(n_items, k_varlen<double, VL_ROWS, float>, smem);
  hipLaunchKernelGGL((k_varlen<double, VL_ROWS, float>), dim3((unsigned)grid), dim3(VL_BLOCK),
                     smem, st, a.ds, a.dl, a.d, a.du, a.dw, x, offs, perm, items, n_items,
                     next);
  return hipGetLastError();
}

template hipError_t launch_varlen<float>(const Diags<float> &, float *, const int64_t *,
                                         const int *, const int2 *, int, int *, hipStream_t);
template hipError_t launch_varlen<double>(const Diags<double> &, double *, const int64_t *,
                                          const int *, const int2 *, int, int *, hipStream_t);

} // namespace pentad
