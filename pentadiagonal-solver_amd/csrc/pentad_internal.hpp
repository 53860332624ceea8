// pentad_internal.hpp -- types shared by the gfx950 kernels and the host
// dispatch of libpentadsolver.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace pentad {

// ---------------------------------------------------------------------------
// Line decomposition of a batched solve along one axis.
//
// Element (i0, i1, ...) lives at i0 + d0*i1 + d0*d1*i2 ...; t_dims[0] is the
// fastest axis.  A "line" (one independent system) runs along solvedim.
// Semantics follow src/cuda/singlenode/pentad_cuda.cu:1420-1453 (and the CPU
// pentad_orig_cpu.cpp:158-195), with 64-bit counts and offsets throughout:
//   solvedim 0     : n_sys = prod(dims[1:]), line j starts at j*len, stride 1
//   solvedim sd > 0: n_in = prod(dims[:sd]), n_out = prod(dims[sd+1:]);
//                    line (o, j) starts at o*n_in*len + j, stride n_in.
// The outermost direction is the n_out == 1 case of the same formula.
// ---------------------------------------------------------------------------
struct Lines {
  int64_t n_sys = 0;  // number of independent systems
  int64_t len = 0;    // rows per system
  int64_t stride = 1; // element distance between consecutive rows of a line
  int64_t n_in = 1;   // lines that are adjacent in memory (stride-1 run)
  int dir = 0;        // 0 = x (contiguous lines), 1 = middle, 2 = outermost
  // Padded (pitched) layouts: lstep != 0 puts line `sys` at
  // (sys / n_in) * gstep + (sys % n_in) * lstep.  lstep == 0 is the dense
  // layout of decompose() below.
  int64_t lstep = 0, gstep = 0;

  // first element of line `sys`
  __host__ __device__ __forceinline__ int64_t base(int64_t sys) const {
    if (lstep != 0) {
      const int64_t o = sys / n_in;
      return o * gstep + (sys - o * n_in) * lstep;
    }
    if (dir == 0) return sys * len;
    if (n_in == n_sys) return sys; // outermost: lines are adjacent
    const int64_t o = sys / n_in;
    return o * n_in * len + (sys - o * n_in);
  }
  bool dense() const { return lstep == 0; }
};

// Returns 0 on success, a negative code when the shape is invalid.
int decompose(const int *dims, int ndim, int solvedim, Lines *out);
// The same over a padded array: element (i, j, k) of a field with extents
// dims[] is stored at i + pads[0] * (j + pads[1] * k), pads[k] >= dims[k]
// (the layout of src/cpu/singlenode/pentad_cpu.cpp:611-613, :457, :569).
// pads == nullptr, or pads that change no address, give the dense layout.
int decompose_padded(const int *dims, const int *pads, int ndim, int solvedim, Lines *out);

// ---------------------------------------------------------------------------
// Kernel argument bundle: the five diagonals (read only) and the in/out RHS.
// Naming follows the reference API (include/pentadsolver.hpp:93-104):
// ds = 2nd lower, dl = lower, d = main, du = upper, dw = 2nd upper.
// ---------------------------------------------------------------------------
template <typename T> struct Diags {
  const T *ds, *dl, *d, *du, *dw;
};

// Launchers, one per kernel family.  Each returns hipSuccess or the launch
// error; none synchronises.  Implemented in the kernels_*.hip files.
template <typename T>
hipError_t launch_thomas_lane(const Diags<T> &a, T *x, const Lines &ln,
                              T *scratch, hipStream_t st);

// Partitioned solve, x direction: `lanes` lanes of one wavefront per system,
// each lane owning `rows` consecutive rows (lanes * rows == len).
template <typename T>
hipError_t launch_part_x(const Diags<T> &a, T *x, const Lines &ln, int rows,
                         int lanes, hipStream_t st);
bool part_x_supported(int elem_bytes, int rows, int lanes);
// The same solve with 128 or 256 lanes per system (lines of 1024 / 2048 rows
// in 8-row chunks, FP32 also 16-row): the lanes of a system span wavefronts
// and exchange through LDS.  launch_part_x / launch_part_x_map accept these.
bool part_x_wide_supported(int elem_bytes, int rows, int lanes);
// Addressing of a batch of contiguous lines for the x-like kernels (used by
// the ADI driver, whose field and coefficient layouts differ per direction).
// System s splits as s_lo = s % S1, s_hi = s / S1 (64-bit):
//   RHS line start          s_lo * in_lo  + s_hi * in_hi
//   coefficient line start  s_lo * c_lo   + s_hi * c_hi
//   solution element row r  s_lo * out_lo + r * out_row + s_hi * out_hi
// (rows of a line are contiguous on input).  A workgroup owns consecutive
// s, so with out_lo = 1 its solutions form runs along s_lo in the output.
struct XMap {
  int64_t S1 = 1;
  int64_t in_lo = 0, in_hi = 0;
  int64_t c_lo = 0, c_hi = 0;
  int64_t out_lo = 0, out_row = 1, out_hi = 0;
  __host__ __device__ __forceinline__ int64_t lo(int64_t s) const { return s % S1; }
  __host__ __device__ __forceinline__ int64_t hi(int64_t s) const { return s / S1; }
};
// lines [sys][row] contiguous, coefficients alike, in place (padded x
// lines: n_in lines lstep apart per group, groups gstep apart)
inline XMap xmap_inplace(const Lines &ln) {
  XMap m;
  m.out_row = 1;
  if (!ln.dense()) {
    m.S1 = ln.n_in > 0 ? ln.n_in : 1;
    m.in_lo = m.c_lo = m.out_lo = ln.lstep;
    m.in_hi = m.c_hi = m.out_hi = ln.gstep;
    return m;
  }
  m.S1 = ln.n_sys > 0 ? ln.n_sys : 1;
  m.in_lo = m.c_lo = m.out_lo = ln.len;
  return m;
}
// The same solves with the mapped addressing; rot = false writes the
// solution over the RHS (out == b), rot = true writes it at the mapped
// output positions (out != b), staged through LDS.
template <typename T>
hipError_t launch_part_x_map(const Diags<T> &a, const T *b, T *out, bool rot,
                             int64_t n_sys, int64_t len, const XMap &map, int rows,
                             int lanes, hipStream_t st);

// The strided partitioned solve in FP64 arithmetic over FP32-stored
// coefficients (widened on load): FP32 coefficients, FP64 right-hand side and
// solution, the result of the widened system to FP64 rounding.
hipError_t launch_part_yz_widened(const Diags<float> &a, double *x, const Lines &ln, int rows,
                                  int sys_per_wg, hipStream_t st);
// Partitioned solve, strided (y/z) directions: a workgroup owns `sys_per_wg`
// adjacent systems and splits each into len/rows chunks, one lane per
// (system, chunk); the reduced system is solved through LDS.
template <typename T>
hipError_t launch_part_yz(const Diags<T> &a, T *x, const Lines &ln, int rows,
                          int sys_per_wg, hipStream_t st);
bool part_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len);
// The same with each line shared by two workgroups (the first and second
// half of its rows; kernels_split.hip): S systems per line set, scratch of
// part_yz_split_scratch_bytes in the caller's buffer.  For the outermost
// direction, where every row of a line lies in its own page.
bool part_yz_split_supported(int elem_bytes, int rows, int S, int64_t len);
size_t part_yz_split_scratch_bytes(const Lines &ln, int S, int elem_bytes);
template <typename T>
hipError_t launch_part_yz_split(const Diags<T> &a, T *x, const Lines &ln, int rows, int S,
                                void *scratch, hipStream_t st);
hipError_t launch_part_yz_split_widened(const Diags<float> &a, double *x, const Lines &ln,
                                        int rows, int S, void *scratch, hipStream_t st);

// LDS-staged x solve for any length (rows 8 or 16, lanes a power of two,
// len <= rows * lanes; rows past the end are identity rows): "transpose".
bool tr_x_supported(int rows, int lanes, int64_t len);
template <typename T>
hipError_t launch_tr_x(const Diags<T> &a, T *x, const Lines &ln, int rows, int lanes,
                       hipStream_t st);

// Writes one int with a kernel (usable inside captured graphs, where a
// memset node is avoided).
hipError_t launch_fill_int(int *p, int v, hipStream_t st);

// Variable-length batch: systems packed back to back (offsets), solved in
// one persistent launch over work items of S systems x P chunks of
// VL_ROWS rows (S * P = VL_BLOCK); lengths 3 .. VL_BLOCK * VL_ROWS.  The
// workgroups take items in list order from the device counter `next`,
// which the launcher resets first (two solves sharing one counter must not
// overlap).
constexpr int VL_BLOCK = 512;
constexpr int VL_ROWS = 16;
template <typename T>
hipError_t launch_varlen(const Diags<T> &a, T *x, const int64_t *offs, const int *perm,
                         const int2 *items, int n_items, int *next, hipStream_t st);
hipError_t launch_varlen_ir(const Diags<float> &a, double *x, const int64_t *offs,
                            const int *perm, const int2 *items, int n_items, int *next,
                            int max_steps, double tol, int *steps, hipStream_t st);
hipError_t launch_varlen_widened(const Diags<float> &a, double *x, const int64_t *offs,
                                 const int *perm, const int2 *items, int n_items, int *next,
                                 hipStream_t st);

// Shared factorisation (every system of the direction has system 0's
// coefficients): k_sf_factor writes the factor table of system 0 into `tab`
// (sf_table_bytes), then the solve kernels stream only right-hand sides.
size_t sf_table_bytes(int elem_bytes, int rows, int64_t len);
bool sf_x_supported(int elem_bytes, int rows, int64_t len);
bool sf_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len);
template <typename T>
hipError_t launch_sf_factor(const Diags<T> &a, const Lines &ln, int rows, T *tab,
                            hipStream_t st);
// x-like lines through XMap; rot as in launch_part_x_map
template <typename T>
hipError_t launch_sf_x_map(const T *tab, const T *b, T *out, bool rot, int64_t n_sys,
                           int64_t len, const XMap &map, int rows, hipStream_t st);
template <typename T>
hipError_t launch_sf_yz(const T *tab, T *x, const Lines &ln, int rows, int sys_per_wg,
                        hipStream_t st);

// Structured coefficients: along every line the system is
// id * I + s * [w0 .. w4] (ds, dl, d - id, du, dw), couplings past either end
// dropped, with one scale s per line.  x-like kernels index the scale as
// lo * sa + hi * sb of the system's XMap split; the strided kernels by the
// system index of the Lines decomposition.
template <typename T> struct Stencil {
  const T *scale = nullptr;
  int64_t sa = 1, sb = 0;
  T w[5] = {};
  T id = T(1);
};
template <typename T>
hipError_t launch_stencil_x_map(const Stencil<T> &stc, const T *b, T *out, bool rot,
                                int64_t n_sys, int64_t len, const XMap &map, int rows,
                                int lanes, hipStream_t st);
bool stencil_yz_supported(int elem_bytes, int rows, int sys_per_wg, int64_t len);
// FP32 solve + FP64 refinement over FP64 structured coefficients, x-like
// lines with the rotated store (rows 4, 6, 8, 10 with lanes = len / rows a
// power of two <= 64): converges to the FP64 operator's solution
hipError_t launch_stencil_ir_x_map(const Stencil<double> &stc, const double *b, double *out,
                                   int64_t n_sys, int64_t len, const XMap &map, int rows,
                                   int max_steps, double tol, int *steps, hipStream_t st);
template <typename T>
hipError_t launch_stencil_yz(const Stencil<T> &stc, T *x, const Lines &ln, int rows,
                             int sys_per_wg, hipStream_t st);

// FP32 coefficients + FP64 RHS/solution with on-chip iterative refinement.
// steps (device int, may be null) receives atomicMax of the refinement
// steps taken.
bool ir_x_supported(int rows, int64_t len);
bool ir_yz_supported(int rows, int sys_per_wg, int64_t len);
hipError_t launch_ir_x(const Diags<float> &a, double *x, const Lines &ln, int rows,
                       int max_steps, double tol, int *steps, hipStream_t st);
hipError_t launch_ir_x_map(const Diags<float> &a, const double *b, double *out,
                           bool rot, int64_t n_sys, int64_t len, const XMap &map,
                           int rows, int max_steps, double tol, int *steps,
                           hipStream_t st);
hipError_t launch_ir_yz(const Diags<float> &a, double *x, const Lines &ln, int rows,
                        int sys_per_wg, int max_steps, double tol, int *steps,
                        hipStream_t st);
// The rotated-store refinement kernel reading the five coefficient arrays
// interleaved in one array (the layout launch_interleave writes) per group
// of ir_x_interleave_group(rows, len) rows: 4 for 4- and 8-row chunks
// (lines of a multiple of 4 rows), 2 for 6-row chunks and for 640-row lines
// of 10-row chunks, 0 where the kernel has no interleaved form; line starts
// at multiples of the group, coef 16-byte aligned.
// The x kernels over FP32 coefficients widened to FP64 on load, FP64 b and
// out: one FP64 solve of the system the refinement converges to, for the
// lengths no refinement kernel covers (the rows / lanes of
// launch_part_x_map<double>).
hipError_t launch_part_x_map_widened(const Diags<float> &a, const double *b, double *out,
                                     bool rot, int64_t n_sys, int64_t len, const XMap &map,
                                     int rows, int lanes, hipStream_t st);
int ir_x_interleave_group(int rows, int64_t len);
hipError_t launch_ir_x_map_il(const float *coef, const double *b, double *out, int64_t n_sys,
                              int64_t len, const XMap &map, int rows, int max_steps, double tol,
                              int *steps, hipStream_t st);
// out[5 G (i / G) + G k + i % G] = src[k][i], k < 5, i < n, for a group of
// G = 2 or 4 rows (n % G == 0)
hipError_t launch_interleave(const float *const src[5], float *out, int64_t n, int group,
                             hipStream_t st);
// x direction, any length <= 8 * 64: coefficient spans staged through LDS
// (rows 8, lanes the next power of two, identity rows past the end)
bool ir_tr_x_supported(int rows, int64_t len);
hipError_t launch_ir_tr_x(const Diags<float> &a, double *x, const Lines &ln, int rows,
                          int max_steps, double tol, int *steps, hipStream_t st);
size_t ir_lane_scratch_bytes(const Lines &ln);
hipError_t launch_ir_lane(const Diags<float> &a, double *x, const Lines &ln,
                          void *scratch, int max_steps, double tol, int *steps,
                          hipStream_t st);

// ---------------------------------------------------------------------------
// Distributed line solve, one rank's share (kernels_dist.hip).  A line is
// split into P slabs of m >= 5 rows, slab k on rank k; a record holds
// DIST_REC words per line, stored [word][line].
// ---------------------------------------------------------------------------
constexpr int DIST_REC = 23;
// phase 1: this rank's record (first / last: rank 0 / rank P-1)
template <typename T>
hipError_t launch_dist_forward(const Diags<T> &a, const T *x, const Lines &ln, bool first,
                               bool last, T *rec, hipStream_t st);
// phase 2 on one device: the P records of n_sys lines, consecutive in recs;
// seps gets per rank [4][n_sys]: the previous rank's separator and its own
// (P <= 64)
template <typename T>
hipError_t launch_dist_reduced(const T *recs, int64_t n_sys, int P, T *seps, hipStream_t st);
// phase 3, first half: separator couplings moved to the interior's end rows,
// separator rows written (sep: this rank's [4][n_sys])
template <typename T>
hipError_t launch_dist_halo(const Diags<T> &a, T *x, const Lines &ln, bool first, const T *sep,
                            hipStream_t st);

// ---------------------------------------------------------------------------
// 2-D cavity consumer: pressure-correction assembly and solve
// (kernels_cavity.hip).  Inputs row-major with the reference's extents:
// up = u' (ny+2) x (nx+1), vp = v' (ny+1) x (nx+2), uv (ny+2) x (nx+2);
// the field and the arrays have rows of nx_pad elements.
// ---------------------------------------------------------------------------
struct Cavity {
  int nx = 0, ny = 0, nx_pad = 0;
  double alpha_u = 0, dx = 0, dy = 0;
  const double *up = nullptr, *vp = nullptr, *uv = nullptr;
};
// x, then the x set (ds, dl, d, du, dw), then the y set (p[6] == nullptr:
// no y set)
struct CavityOut {
  double *p[11] = {};
};
bool cavity_plan_x(int nx, int *rows, int *lanes);
bool cavity_plan_y(int ny, int *rows, int *sys_per_wg);
hipError_t launch_cavity_assemble(const Cavity &c, const CavityOut &o, hipStream_t st);
bool cavity_y_supported(int ny, int rows, int sys_per_wg);
// assembly fused into the x solve (8-row chunks, `lanes` per row), then the
// y solve (my-row chunks, S columns per workgroup), in place on x
hipError_t launch_cavity_solve(const Cavity &c, double *x, int mx, int lanes, int my, int S,
                               hipStream_t st);

} // namespace pentad
