// kernels_split.hip -- strided lines shared by two workgroups (split_core.hpp).
//
// k_part_yz_split: a line set of S adjacent systems is solved by a pair of
// 512-thread workgroups, each holding half of every line (P chunks of M rows
// per system, thread t owns system t % S, chunk t / S as in k_part_yz).  The
// pair is formed from tickets; the two halves trade six words per system
// once, after their block PCR.  HBM traffic stays the compulsory 5 + 1 reads
// and 1 write per row, plus 2 x 6 words and 2 flags per system.
//
// Scratch (caller's buffer, split_scratch_bytes): kSplitHeaderInts ints
// (ticket counters per class, the time-out word), 2 flags per line set
// (padded to 256 bytes), then [line set][half][6][S] words.  k_split_reset
// zeroes the counters and flags before every solve (a kernel, so that the
// launch may be captured into a graph).
#include "split_core.hpp"
#include "partition_exchange.hpp"
#include "pentad_internal.hpp"

namespace pentad {

__global__ void k_split_reset(int *w, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = 0;
}

// FP32 (16-row chunks) at 4 waves per SIMD, 128 VGPRs with 8 spilled: two
// workgroups per CU, so one computes while the other waits for its partner
// (512^3 z 1.01 -> 0.985 ms); FP64 needs 248 VGPRs, one workgroup per CU.
template <typename T, int M, typename C = T>
__global__ void __launch_bounds__(512, sizeof(T) == 4 ? 4 : 2)
    k_part_yz_split(const C *__restrict__ ds, const C *__restrict__ dl,
                    const C *__restrict__ d, const C *__restrict__ du,
                    const C *__restrict__ dw, T *__restrict__ x, Lines ln, int S, int P,
                    int64_t n_sets, int *head, int *flags,
                    typename SplitWord<T>::W *words) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ int ticket;
  const int t = threadIdx.x;
  const int cls = (int)(blockIdx.x % kSplitClasses);
  if (t == 0) ticket = atomicAdd(head + cls, 1);
  __syncthreads();
  const int tk = ticket;
  const int half = tk & 1;
  const int64_t set = (int64_t)(tk >> 1) * kSplitClasses + cls;
  if (set >= n_sets) return; // both workgroups of a pair leave here
  const int sl = t % S;
  const int p = t / S;
  const int64_t sys = set * S + sl;
  const bool valid = sys < ln.n_sys;
  const int64_t st = ln.stride;
  const int64_t b =
      ln.base(valid ? sys : ln.n_sys - 1) + ((int64_t)half * P + p) * M * st;

  // (the host launches whole chunks only, 2 P M == len; the branch keeps
  // the compiler from hoisting every load ahead of the solve, which costs
  // 28 VGPRs and spills the FP64 16-row kernel)
  T s[M], l[M], dd[M], u[M], w[M], f[M];
  if (2 * P * M == ln.len) {
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
#pragma unroll
    for (int i = 0; i < M; ++i) { // identity rows
      s[i] = l[i] = u[i] = w[i] = f[i] = T(0);
      dd[i] = T(1);
    }
  }

  LdsEx ex{smem_raw, sl, p, S, P};
  T *tail = reinterpret_cast<T *>(smem_raw + (size_t)14 * S * P * sizeof(T));
  const int64_t slot = set * 2;
  PairLink<T> link{words + (slot + half) * 6 * S,
                   words + (slot + 1 - half) * 6 * S,
                   flags + slot + half,
                   flags + slot + 1 - half,
                   head + kTimeoutWord,
                   tail,
                   tail + 6 * S,
                   tail + 8 * S,
                   S,
                   half};
  part_solve_split<T, M>(s, l, dd, u, w, f, p, P, half, sl, ex, link);
  if (valid) {
#pragma unroll
    for (int i = 0; i < M; ++i) x[b + i * st] = f[i];
  }
}

namespace {

int64_t split_sets(const Lines &ln, int S) { return (ln.n_sys + S - 1) / S; }

size_t split_flag_bytes(int64_t n_sets) {
  return ((size_t)n_sets * 2 * sizeof(int) + 255) & ~(size_t)255;
}

template <typename T, int M, typename C>
hipError_t go_split(const Diags<C> &a, T *x, const Lines &ln, int S, void *scratch,
                    hipStream_t st) {
  const int P = (int)(ln.len / (2 * M));
  const int64_t n_sets = split_sets(ln, S);
  const int64_t grid = 2 * kSplitClasses * ((n_sets + kSplitClasses - 1) / kSplitClasses);
  if (grid > INT32_MAX) return hipErrorInvalidConfiguration;
  int *head = static_cast<int *>(scratch);
  int *flags = head + kSplitHeaderInts;
  auto *words = reinterpret_cast<typename SplitWord<T>::W *>(
      static_cast<unsigned char *>(scratch) + kSplitHeaderInts * sizeof(int) +
      split_flag_bytes(n_sets));
  const int64_t nz = kSplitHeaderInts + 2 * n_sets;
  hipLaunchKernelGGL(k_split_reset, dim3((unsigned)std::min<int64_t>((nz + 255) / 256, 1024)),
                     dim3(256), 0, st, head, nz);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const size_t smem = ((size_t)14 * S * P + (size_t)18 * S) * sizeof(T);
  if (smem > 65536) {
    e = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_part_yz_split<T, M, C>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((k_part_yz_split<T, M, C>), dim3((unsigned)grid), dim3(S * P), smem, st,
                     a.ds, a.dl, a.d, a.du, a.dw, x, ln, S, P, n_sets, head, flags, words);
  return hipGetLastError();
}

} // namespace

// S systems per line set, M rows per chunk, two halves of P = len / (2 M)
// chunks: S x P threads (at most 512) per workgroup.  FP64 chunks of 8 or 16
// rows, FP32 of 16 (the register budget of 512-thread workgroups).
bool part_yz_split_supported(int elem_bytes, int rows, int S, int64_t len) {
  if (!(rows == 16 || (elem_bytes == 8 && rows == 8))) return false;
  if (S != 32 && S != 64) return false;
  if (len % (2 * rows) != 0) return false;
  const int64_t P = len / (2 * rows);
  return P >= 2 && S * P <= 512;
}

size_t part_yz_split_scratch_bytes(const Lines &ln, int S, int elem_bytes) {
  const int64_t n_sets = split_sets(ln, S);
  return (size_t)kSplitHeaderInts * sizeof(int) + split_flag_bytes(n_sets) +
         (size_t)n_sets * 2 * 6 * S * elem_bytes;
}

template <typename T>
hipError_t launch_part_yz_split(const Diags<T> &a, T *x, const Lines &ln, int rows, int S,
                                void *scratch, hipStream_t st) {
  if (scratch == nullptr || !part_yz_split_supported((int)sizeof(T), rows, S, ln.len))
    return hipErrorInvalidValue;
  if constexpr (sizeof(T) == 8) {
    if (rows == 8) return go_split<T, 8, T>(a, x, ln, S, scratch, st);
  }
  return go_split<T, 16, T>(a, x, ln, S, scratch, st);
}
template hipError_t launch_part_yz_split<float>(const Diags<float> &, float *, const Lines &,
                                                int, int, void *, hipStream_t);
template hipError_t launch_part_yz_split<double>(const Diags<double> &, double *,
                                                 const Lines &, int, int, void *,
                                                 hipStream_t);

// FP32 coefficients widened on load into the FP64 solve (the mixed-precision
// contract's automatic choice along y / z)
hipError_t launch_part_yz_split_widened(const Diags<float> &a, double *x, const Lines &ln,
                                        int rows, int S, void *scratch, hipStream_t st) {
  if (scratch == nullptr || !part_yz_split_supported(8, rows, S, ln.len))
    return hipErrorInvalidValue;
  if (rows == 8) return go_split<double, 8, float>(a, x, ln, S, scratch, st);
  return go_split<double, 16, float>(a, x, ln, S, scratch, st);
}

} // namespace pentad
