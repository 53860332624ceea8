// kernels_shared.hip -- shared-factorisation solves ("shared-fact").
//
// k_sf_factor  one workgroup of P <= 512 threads factors system 0 of a direction
//              into the table of shared_core.hpp (in the caller's scratch).
// k_sf_x       x-like lines (contiguous rows): L lanes of a wavefront per
//              system, M rows per lane; the table is staged once into LDS
//              and the workgroup loops over systems (persistent), so each
//              system costs its right-hand side in and its solution out.
//              ROT = true writes the solution through the rotated store
//              (ADI layouts).
// k_sf_yz      strided lines: S systems x P chunks per workgroup, LDS
//              exchange, persistent over tiles of S adjacent systems.
#include "pentad_internal.hpp"
#include "rotated_store.hpp"
#include "rows_io.hpp"
#include "shared_core.hpp"

namespace pentad {

namespace {

template <typename T, int M>
__device__ __forceinline__ void zero_out_of_range(T (&s)[M], T (&l)[M], T (&u)[M],
                                                  T (&w)[M], bool first, bool last) {
  if (first) {
    s[0] = s[1] = T(0);
    l[0] = T(0);
  }
  if (last) {
    u[M - 1] = T(0);
    w[M - 2] = w[M - 1] = T(0);
  }
}

// stage the table (K x P values) into LDS
template <typename T>
__device__ __forceinline__ void stage_table(T *dst, const T *__restrict__ tab, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = tab[i];
  __syncthreads();
}

} // namespace

template <typename T, int M>
__global__ void __launch_bounds__(512)
    k_sf_factor(const T *__restrict__ ds, const T *__restrict__ dl,
                const T *__restrict__ d, const T *__restrict__ du,
                const T *__restrict__ dw, int64_t base, int64_t stride, int P,
                T *__restrict__ tab) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int p = threadIdx.x;
  T s[M], l[M], dd[M], u[M], w[M];
#pragma unroll
  for (int i = 0; i < M; ++i) {
    const int64_t e = base + (int64_t)(p * M + i) * stride;
    s[i] = ds[e];
    l[i] = dl[e];
    dd[i] = d[e];
    u[i] = du[e];
    w[i] = dw[e];
  }
  zero_out_of_range<T, M>(s, l, u, w, p == 0, p == P - 1);
  LdsEx ex{smem_raw, 0, p, 1, P};
  sf_factor<T, M>(s, l, dd, u, w, p, P, ex, tab);
}

template <typename T, int M, bool ROT>
__global__ void __launch_bounds__(512)
    k_sf_x(const T *__restrict__ tab, int K, const T *b, T *out, int64_t n_sys,
           int64_t len, int lshift, XMap map) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int L = 1 << lshift;
  const int S = (int)blockDim.x >> lshift;
  T *tl = reinterpret_cast<T *>(smem_raw);
  stage_table(tl, tab, K * L);
  T *tile = tl + ((K * L + 1) & ~1); // 16-byte aligned after the table
  const int sl = (int)threadIdx.x >> lshift, p = (int)threadIdx.x & (L - 1);
  WaveEx ex{(int)(threadIdx.x & 63), p, L};
  const int64_t ntiles = (n_sys + S - 1) / S;
  for (int64_t t = ROT ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x; t < ntiles;
       t += gridDim.x) {
    const int64_t sys = t * S + sl;
    const bool valid = sys < n_sys;
    const int64_t sc = valid ? sys : n_sys - 1;
    const int64_t off = map.lo(sc) * map.in_lo + map.hi(sc) * map.in_hi + (int64_t)p * M;
    T f[M];
    RowsIO<T, M>::load(b + off, f);
    sf_apply<T, M>(tl + p, L, f, ex);
    if constexpr (ROT) {
      stage_rotated_store<T, M>(tile, f, p, lshift, (int)len, out, n_sys, t, map);
    } else {
      if (valid) RowsIO<T, M>::store(out + off, f);
    }
  }
}

template <typename T, int M>
__global__ void __launch_bounds__(512)
    k_sf_yz(const T *__restrict__ tab, int K, T *__restrict__ x, Lines ln, int S, int P) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tl = reinterpret_cast<T *>(smem_raw);
  stage_table(tl, tab, K * P);
  unsigned char *exbuf = reinterpret_cast<unsigned char *>(tl + ((K * P + 1) & ~1));
  const int t = threadIdx.x;
  const int sl = t % S, p = t / S;
  LdsEx ex{exbuf, sl, p, S, P};
  const int64_t st = ln.stride;
  const int64_t ntiles = (ln.n_sys + S - 1) / S;
  for (int64_t tile = xcd_tile(blockIdx.x, gridDim.x); tile < ntiles; tile += gridDim.x) {
    const int64_t sys = tile * S + sl;
    const bool valid = sys < ln.n_sys;
    const int64_t b0 = ln.base(valid ? sys : ln.n_sys - 1) + (int64_t)p * M * st;
    T f[M];
#pragma unroll
    for (int i = 0; i < M; ++i) f[i] = x[b0 + i * st];
    sf_apply<T, M>(tl + p, P, f, ex);
    if (valid) {
#pragma unroll
      for (int i = 0; i < M; ++i) x[b0 + i * st] = f[i];
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

constexpr int SF_YZ_MAXT = 512;

int sf_K(int M, int P) { return 7 * (M - 2) + 10 + 12 * sf_rounds(P); }

// persistent grid: enough workgroups to fill every CU a few times over
int64_t persistent_grid(int64_t tiles, size_t smem) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  const int per_cu = smem > 80 * 1024 ? 1 : (smem > 53 * 1024 ? 2 : 4);
  const int64_t cap = (int64_t)cus * per_cu;
  return tiles < cap ? tiles : cap;
}

template <typename K>
hipError_t allow_smem(K kernel, size_t smem) {
  if (smem <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
}

template <typename T, int M>
hipError_t go_factor(const Diags<T> &a, int64_t base, int64_t stride, int P, T *tab,
                     hipStream_t st) {
  const size_t smem = (size_t)8 * P * sizeof(T);
  hipLaunchKernelGGL((k_sf_factor<T, M>), dim3(1), dim3(P), smem, st, a.ds, a.dl, a.d,
                     a.du, a.dw, base, stride, P, tab);
  return hipGetLastError();
}

template <typename T, int M>
hipError_t go_sf_x(const T *tab, const T *b, T *out, bool rot, int64_t n_sys, int64_t len,
                   const XMap &map, hipStream_t st) {
  const int L = (int)(len / M);
  int lshift = 0;
  while ((1 << lshift) < L) ++lshift;
  const int block = 512, S = block / L; // 8 waves share one staged table
  const int K = sf_K(M, L);
  const size_t tab_bytes = (size_t)((K * L + 1) & ~1) * sizeof(T);
  const size_t smem = tab_bytes + (rot ? rotated_tile_bytes<T>(block, M, len) : 0);
  const int64_t grid = persistent_grid((n_sys + S - 1) / S, smem);
  hipError_t e;
  if (rot) {
    if ((e = allow_smem(&k_sf_x<T, M, true>, smem)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_sf_x<T, M, true>), dim3((unsigned)grid), dim3(block), smem, st,
                       tab, K, b, out, n_sys, len, lshift, map);
  } else {
    if ((e = allow_smem(&k_sf_x<T, M, false>, smem)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_sf_x<T, M, false>), dim3((unsigned)grid), dim3(block), smem, st,
                       tab, K, b, out, n_sys, len, lshift, map);
  }
  return hipGetLastError();
}

template <typename T, int M>
hipError_t go_sf_yz(const T *tab, T *x, const Lines &ln, int S, hipStream_t st) {
  const int P = (int)(ln.len / M);
  const int K = sf_K(M, P);
  const size_t smem = (size_t)((K * P + 1) & ~1) * sizeof(T) + (size_t)2 * P * S * sizeof(T);
  const int64_t grid = persistent_grid((ln.n_sys + S - 1) / S, smem);
  hipError_t e;
  if ((e = allow_smem(&k_sf_yz<T, M>, smem)) != hipSuccess) return e;
  hipLaunchKernelGGL((k_sf_yz<T, M>), dim3((unsigned)grid), dim3(S * P), smem, st, tab, K,
                     x, ln, S, P);
  return hipGetLastError();
}

} // namespace

size_t sf_table_bytes(int elem_bytes, int rows, int64_t len) {
  const int P = (int)(len / rows);
  return (size_t)((sf_K(rows, P) * P + 1) & ~1) * (size_t)elem_bytes;
}

bool sf_x_supported(int elem_bytes, int rows, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 16) return false;
  if (len % rows) return false;
  const int64_t L = len / rows;
  if (L < 1 || L > 64 || (L & (L - 1))) return false;
  // table (+ a rotated tile) must fit the workgroup's LDS
  return sf_table_bytes(elem_bytes, rows, len) +
             (size_t)elem_bytes * (512 / L * len + L) <= 150 * 1024;
}

bool sf_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len) {
  if (rows != 4 && rows != 6 && rows != 8 && rows != 16) return false;
  if (len % rows) return false;
  const int64_t P = len / rows;
  if (P < 1 || P > SF_YZ_MAXT) return false;
  const int64_t block = P * sys_per_wg;
  if (block > SF_YZ_MAXT || block % 64 != 0) return false;
  return sf_table_bytes(elem_bytes, rows, len) + (size_t)2 * block * elem_bytes <= 150 * 1024;
}

#define SF_ROWS(X) X(4) X(6) X(8) X(16)

template <typename T>
hipError_t launch_sf_factor(const Diags<T> &a, const Lines &ln, int rows, T *tab,
                            hipStream_t st) {
  const int P = (int)(ln.len / rows);
  if (P < 1 || P > 512 || (int64_t)P * rows != ln.len) return hipErrorInvalidValue;
  const int64_t base = ln.base(0);
#define F(MM) \
  if (rows == MM) return go_factor<T, MM>(a, base, ln.stride, P, tab, st);
  SF_ROWS(F)
#undef F
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_sf_x_map(const T *tab, const T *b, T *out, bool rot, int64_t n_sys,
                           int64_t len, const XMap &map, int rows, hipStream_t st) {
  if (!sf_x_supported((int)sizeof(T), rows, len)) return hipErrorInvalidValue;
#define F(MM) \
  if (rows == MM) return go_sf_x<T, MM>(tab, b, out, rot, n_sys, len, map, st);
  SF_ROWS(F)
#undef F
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_sf_yz(const T *tab, T *x, const Lines &ln, int rows, int sys_per_wg,
                        hipStream_t st) {
  if (!sf_yz_supported((int)sizeof(T), rows, sys_per_wg, ln.len)) return hipErrorInvalidValue;
#define F(MM) \
  if (rows == MM) return go_sf_yz<T, MM>(tab, x, ln, sys_per_wg, st);
  SF_ROWS(F)
#undef F
  return hipErrorInvalidValue;
}
#undef SF_ROWS

template hipError_t launch_sf_factor<float>(const Diags<float> &, const Lines &, int, float *,
                                            hipStream_t);
template hipError_t launch_sf_factor<double>(const Diags<double> &, const Lines &, int,
                                             double *, hipStream_t);
template hipError_t launch_sf_x_map<float>(const float *, const float *, float *, bool,
                                           int64_t, int64_t, const XMap &, int, hipStream_t);
template hipError_t launch_sf_x_map<double>(const double *, const double *, double *, bool,
                                            int64_t, int64_t, const XMap &, int, hipStream_t);
template hipError_t launch_sf_yz<float>(const float *, float *, const Lines &, int, int,
                                        hipStream_t);
template hipError_t launch_sf_yz<double>(const double *, double *, const Lines &, int, int,
                                         hipStream_t);

} // namespace pentad
