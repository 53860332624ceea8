/*
 * pentadsolver.h -- C ABI of libpentadsolver, the gfx950 batched
 * pentadiagonal solver.
 *
 * The first block is the drop-in surface of the reference's single-node
 * library: the same names, argument order and meaning as
 * include/pentadsolver.hpp:36-131 of SidhantKrAgrawal/pentadiagonal-solver,
 * with device pointers as in its GPU build.  A caller written against that
 * header links against this library unchanged.  The second block adds
 * entry points the reference does not have; each says so.
 *
 * Conventions (reference include/pentadsolver.hpp:88-106):
 *  - ds, dl, d, du, dw: 2nd lower, lower, main, upper, 2nd upper diagonal;
 *    x: right-hand side on input, solution on output (in place).
 *  - t_dims[0] is the fastest dimension, t_ndim <= 3 is not required here
 *    (any t_ndim >= 1 works); 0 <= t_solvedim < t_ndim.
 *  - Every array holds prod(t_dims) elements in device memory.
 *  - Out-of-range couplings (ds of rows 0-1, dl of row 0, du of row N-1,
 *    dw of rows N-2..N-1 of every system) are ignored.
 *  - Launches are asynchronous on the handle's stream (the null stream for
 *    a NULL handle); the library never synchronises inside a solve.
 */
#ifndef PENTADSOLVER_H_INCLUDED
#define PENTADSOLVER_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Handle (include/pentadsolver_handle.hpp:4-5: `void *` in single-node
 * builds).  NULL is valid everywhere and selects a process-wide default
 * context on the current device and the null stream. */
typedef void *pentadsolver_handle_t;

/* Status codes: same values as the reference's pentadStatus_t
 * (include/pentad_common.hpp:26-37). */
typedef enum {
  PENTAD_STATUS_SUCCESS = 0,
  PENTAD_STATUS_NOT_INITIALIZED = 1,
  PENTAD_STATUS_ALLOC_FAILED = 2,
  PENTAD_STATUS_INVALID_VALUE = 3,
  PENTAD_STATUS_ARCH_MISMATCH = 4,
  PENTAD_STATUS_MAPPING_ERROR = 5,
  PENTAD_STATUS_EXECUTION_FAILED = 6,
  PENTAD_STATUS_INTERNAL_ERROR = 7,
  PENTAD_STATUS_MATRIX_TYPE_NOT_SUPPORTED = 8,
  PENTAD_STATUS_ZERO_PIVOT = 9
} pentadStatus_t;

/* ======================================================================= */
/* Reference surface                                                        */
/* ======================================================================= */

/* Replaces pentadsolver_create (include/pentadsolver.hpp:36-38; a no-op in
 * the reference's single-node builds, src/cuda/singlenode/pentad_cuda.cu:1546).
 * Here it allocates a context holding the stream, the tuning table, the
 * executed-kernel record and the graph cache.  communicator, ndims and
 * num_procs are accepted and ignored (single device). */
void pentadsolver_create(pentadsolver_handle_t *handle, void *communicator,
                         int ndims, const int *num_procs);

/* Replaces pentadsolver_destroy (include/pentadsolver.hpp:40-41). Sets
 * *handle to NULL. */
void pentadsolver_destroy(pentadsolver_handle_t *handle);

/* Replace pentadsolver_{D,S}_gpsv_batch_buffer_size_ext
 * (include/pentadsolver.hpp:64-73).  Bytes of device scratch the solve
 * along t_solvedim needs with the algorithm the dispatcher will choose for
 * this shape.  0 when the chosen kernel keeps its intermediates on chip.
 * (The reference always returned 9*n or 3*n elements.) */
size_t pentadsolver_D_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t handle, const double *ds, const double *dl,
    const double *d, const double *du, const double *dw, const double *x,
    const int *t_dims, int t_ndim, int t_solvedim);
size_t pentadsolver_S_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t handle, const float *ds, const float *dl,
    const float *d, const float *du, const float *dw, const float *x,
    const int *t_dims, int t_ndim, int t_solvedim);

/* Replace pentadsolver_{D,S}_gpsv_batch (include/pentadsolver.hpp:107-118).
 * Solves every system along t_solvedim in place.  t_buffer must hold at
 * least the bytes reported by the matching _buffer_size_ext (may be NULL
 * when that is 0).  Errors are reported on stderr, as the reference does
 * (src/cuda/singlenode/pentad_cuda.cu:1528-1539), and through
 * pentadsolver_get_last_status. */
void pentadsolver_D_gpsv_batch(pentadsolver_handle_t handle, const double *ds,
                               const double *dl, const double *d,
                               const double *du, const double *dw, double *x,
                               const int *t_dims, int t_ndim, int t_solvedim,
                               void *t_buffer);
void pentadsolver_S_gpsv_batch(pentadsolver_handle_t handle, const float *ds,
                               const float *dl, const float *d,
                               const float *du, const float *dw, float *x,
                               const int *t_dims, int t_ndim, int t_solvedim,
                               void *t_buffer);

/* Padded arrays: the reference's CPU library addresses element (i, j, k)
 * at i + pads[0] * (j + pads[1] * k) with pads[0] = t_dims[0] rounded up to
 * its SIMD width (src/cpu/singlenode/pentad_cpu.cpp:611-613, used at :457,
 * :569), and its 2-D consumer allocates nx_pad-wide rows
 * (apps/2D_cavity_cpu.cpp:70-72, apps/preproc.hpp:248, :320).  These take that layout
 * explicitly: t_pads[k] >= t_dims[k] are the extents in memory (t_ndim <= 3;
 * t_pads = NULL is the dense layout).  Elements in the padding are neither
 * read nor written.  Same solve, routing and errors as
 * pentadsolver_{D,S}_gpsv_batch; returns the status.  The buffer-size query
 * takes the same shape.  The x kernels with vector loads need pads[0] (and
 * pads[0] * pads[1]) to be multiples of 16 bytes; other pitches run the
 * workgroup-wide partitioned kernel with scalar loads (no scratch). */
size_t pentadsolver_D_gpsv_batch_padded_buffer_size(
    pentadsolver_handle_t handle, const double *ds, const double *dl,
    const double *d, const double *du, const double *dw, const double *x,
    const int *t_dims, const int *t_pads, int t_ndim, int t_solvedim);
size_t pentadsolver_S_gpsv_batch_padded_buffer_size(
    pentadsolver_handle_t handle, const float *ds, const float *dl,
    const float *d, const float *du, const float *dw, const float *x,
    const int *t_dims, const int *t_pads, int t_ndim, int t_solvedim);
pentadStatus_t pentadsolver_D_gpsv_batch_padded(
    pentadsolver_handle_t handle, const double *ds, const double *dl,
    const double *d, const double *du, const double *dw, double *x,
    const int *t_dims, const int *t_pads, int t_ndim, int t_solvedim,
    void *t_buffer);
pentadStatus_t pentadsolver_S_gpsv_batch_padded(
    pentadsolver_handle_t handle, const float *ds, const float *dl,
    const float *d, const float *du, const float *dw, float *x,
    const int *t_dims, const int *t_pads, int t_ndim, int t_solvedim,
    void *t_buffer);

/* The 2-D cavity consumer's pressure-correction step on the GPU
 * (apps/2D_cavity_cpu.cpp:195-219).  Device inputs, row-major with the
 * reference's extents: u_prime (ny+2) x (nx+1), v_prime (ny+1) x (nx+2),
 * uv (= A_p) (ny+2) x (nx+2); the field x and the arrays have ny rows of
 * nx_pad elements (nx_pad >= nx; nx, ny >= 3).
 *
 * _coef replaces pressure_coef_matrix (apps/preproc.hpp:248-525): the
 * right-hand side into x and the five diagonals into the x set and, when
 * s_y..w_y are not NULL, the y set (the reference fills both with the same
 * values).  Point (i, j) goes to (i-1) * nx_pad + (j-1); the values are the
 * reference's bit for bit.  (When nx < nx_pad the reference's own index
 * expressions put the right column and the bottom row elsewhere,
 * preproc.hpp:298, :381, :465, :495; its consumer only runs nx == nx_pad,
 * where they agree.)
 *
 * _solve is that assembly followed by the x solve and the y solve of
 * apps/2D_cavity_cpu.cpp:201-219, with the coefficients generated inside the
 * solve kernels: the solution overwrites x (its input values are ignored),
 * no coefficient array is written or read.  nx <= 2048, ny <= 8192 (larger
 * grids: _coef + pentadsolver_D_gpsv_batch_padded).  No reference
 * counterpart as one call. */
pentadStatus_t pentadsolver_D_cavity_pressure_coef(
    pentadsolver_handle_t handle, int nx, int ny, int nx_pad, double alphaU,
    double dx, double dy, const double *u_prime, const double *v_prime,
    const double *uv, double *x, double *s_x, double *l_x, double *d_x,
    double *u_x, double *w_x, double *s_y, double *l_y, double *d_y,
    double *u_y, double *w_y);
pentadStatus_t pentadsolver_D_cavity_pressure_solve(
    pentadsolver_handle_t handle, int nx, int ny, int nx_pad, double alphaU,
    double dx, double dy, const double *u_prime, const double *v_prime,
    const double *uv, double *x);

/* Replaces pentadsolver_kernel_that_ran (src/cuda/singlenode/
 * pentad_cuda.cu:1159-1161): name of the kernel the most recent solve along
 * direction dir (0 = x, 1 = y/middle, 2 = z/outermost) actually launched,
 * in any handle.  Names: "naive", "transpose", "thomas-pcr", "shared-fact",
 * "ir", "varlen", or "none". */
const char *pentadsolver_kernel_that_ran(int dir);

/* ======================================================================= */
/* Extensions (no counterpart in the reference)                             */
/* ======================================================================= */

/* Status of the most recent call made with this handle (NULL: default
 * context).  No reference counterpart: the reference returns void and only
 * prints. */
pentadStatus_t pentadsolver_get_last_status(pentadsolver_handle_t handle);

/* Bind a HIP stream (hipStream_t passed as void *) to the handle; all later
 * work of the handle is enqueued on it.  No reference counterpart (the
 * reference uses the legacy default stream only). */
pentadStatus_t pentadsolver_set_stream(pentadsolver_handle_t handle,
                                       void *stream);

/* Algorithm override per direction (0 = x, 1 = y, 2 = z).  Same names as
 * the reference's PENTA_XALGO / PENTA_YALGO / PENTA_ZALGO environment
 * variables (src/cuda/singlenode/pentad_cuda.cu:1455-1489, :1163-1174):
 * "auto", "naive", "transpose", "thomas-pcr", "shared-fact" (and "legacy"
 * for y/z, an alias of "naive"), plus one name of this library's own,
 * "thomas-pcr-split": along y / z each line shared by two workgroups (the
 * automatic choice for the outermost direction of lines over 384 rows whose
 * rows lie 256 KB or more apart, when the caller passes the scratch that
 * pentadsolver_*_gpsv_batch_buffer_size_ext reports; without it the solve
 * runs one workgroup per line).  A handle setting wins over the
 * environment; NULL or "" clears it.  The environment is read on every call
 * unless pentadsolver_freeze_env(handle, 1) was called. */
pentadStatus_t pentadsolver_set_algo(pentadsolver_handle_t handle, int dir,
                                     const char *algo);
pentadStatus_t pentadsolver_freeze_env(pentadsolver_handle_t handle,
                                       int frozen);

/* Shape parameters of the partitioned ("thomas-pcr") kernels along dir:
 * rows per lane, and lanes per system (x) or systems per workgroup (y/z).
 * rows = 0 returns to automatic choice (tuning table, then defaults).  A
 * combination the kernels do not support for a given length is ignored for
 * that solve (the defaults apply).  No reference counterpart (the reference
 * reads PENTA_PCR_LANES, src/cuda/singlenode/pentad_cuda.cu:661-665). */
pentadStatus_t pentadsolver_set_kernel_params(pentadsolver_handle_t handle,
                                              int dir, int rows, int width);

/* Mixed precision: coefficients in FP32, right-hand side and solution in
 * FP64.  The bulk solve runs in FP32; the residual r = b - A x (A widened
 * exactly to FP64) and the update x += dx are the only FP64 steps, and the
 * refinement of each system stops once the estimated relative forward error
 * of the iterate, max|dx_k|^2 / (max|dx_{k-1}| * max|x|) (dx_0 = the FP32
 * solution), is at most tol, or after max_refine corrections.  The result solves widen(A) x = b to FP64
 * accuracy for well-conditioned systems.  Factorisation, residuals and
 * updates stay on chip: HBM traffic is 20 B of coefficients + 8 B in + 8 B
 * out per row.  d_refine_steps (device int*, may be NULL) receives
 * atomicMax(*d_refine_steps, steps taken); the caller zeroes it.  t_buffer
 * as for pentadsolver_D_gpsv_batch, sized by the _ir_buffer_size_ext call.
 * Along y / z the automatic choice is one FP64 solve of the coefficients
 * widened on load instead: the same 20 + 8 + 8 B per row, the widened
 * system's solution to FP64 rounding in one pass (what the refinement
 * converges to), no refinement sweep (steps 0, tol and max_refine unused),
 * because every sweep of the strided refinement kernel exchanges through
 * LDS; pentadsolver_set_algo(h, dir, "thomas-pcr") keeps the refinement
 * kernel there.  No reference counterpart (the reference solves in one
 * precision). */
size_t pentadsolver_DS_gpsv_batch_ir_buffer_size_ext(
    pentadsolver_handle_t handle, const float *ds, const float *dl,
    const float *d, const float *du, const float *dw, const double *x,
    const int *t_dims, int t_ndim, int t_solvedim);
void pentadsolver_DS_gpsv_batch_ir(pentadsolver_handle_t handle,
                                   const float *ds, const float *dl,
                                   const float *d, const float *du,
                                   const float *dw, double *x,
                                   const int *t_dims, int t_ndim,
                                   int t_solvedim, int max_refine, double tol,
                                   int *d_refine_steps, void *t_buffer);

/* Autotuner: times every kernel configuration that can solve this shape
 * (rows per lane x lanes per system for x; rows per lane x systems per
 * workgroup for y/z; the one-lane-per-system kernel when t_buffer holds its
 * scratch), reps launches each on the handle's stream, and stores the
 * fastest in the handle's table for (precision, direction, length).  x is
 * restored before returning.  Synchronises the stream and allocates a
 * temporary copy of x.  Returns the number of configurations timed (< 0 on
 * error); report (may be NULL) receives one "config time" line each.
 * With PENTA_AUTOTUNE=1 in the environment, pentadsolver_{D,S}_gpsv_batch
 * runs this tuning itself the first time a (precision, direction, length)
 * is solved with the automatic choice on a handle (that call synchronises
 * and allocates once; later calls of the key are as before).
 * No reference counterpart (the reference uses fixed dispatch rules and
 * PENTA_PCR_LANES, src/cuda/singlenode/pentad_cuda.cu:613-697). */
int pentadsolver_D_autotune(pentadsolver_handle_t handle, const double *ds,
                            const double *dl, const double *d,
                            const double *du, const double *dw, double *x,
                            const int *t_dims, int t_ndim, int t_solvedim,
                            void *t_buffer, size_t buffer_bytes, int reps,
                            char *report, size_t report_len);
int pentadsolver_S_autotune(pentadsolver_handle_t handle, const float *ds,
                            const float *dl, const float *d, const float *du,
                            const float *dw, float *x, const int *t_dims,
                            int t_ndim, int t_solvedim, void *t_buffer,
                            size_t buffer_bytes, int reps, char *report,
                            size_t report_len);
int pentadsolver_DS_autotune_ir(pentadsolver_handle_t handle, const float *ds,
                                const float *dl, const float *d,
                                const float *du, const float *dw, double *x,
                                const int *t_dims, int t_ndim, int t_solvedim,
                                int max_refine, double tol, void *t_buffer,
                                size_t buffer_bytes, int reps, char *report,
                                size_t report_len);
void pentadsolver_clear_tuning(pentadsolver_handle_t handle);

/* Bandwidth probe with the solve's access pattern: out[i] = in0[i] + ... +
 * in5[i] followed by `fmas` FMAs, n elements of elem_bytes (4 or 8), on the
 * handle's stream.  Counterpart of the reference's
 * apps/roofline_probe_cuda.cu:27-49 (there a stand-alone app). */
pentadStatus_t pentadsolver_probe_stream(pentadsolver_handle_t handle,
                                         const void *const in[6], void *out,
                                         size_t n, int elem_bytes, int fmas);

/* Bandwidth probe with the byte mix of the FP32 + FP64 refinement solve: n
 * rows (n a multiple of 4), each 20 B of FP32 coefficients and 8 B of FP64 b
 * read and 8 B of FP64 x written (x = b + the coefficients' sum), on the
 * handle's stream.  streams = 7: coef[0..4] are five arrays of n floats;
 * streams = 3: coef[0] holds the five interleaved per 4 rows (5n floats,
 * 80 B per 4 rows); streams = 2: coef[0] holds coefficients and b
 * interleaved per 4 rows (7n floats, 112 B per 4 rows) and b is not read
 * (any valid pointer).  All pointers 16-byte aligned.  No reference
 * counterpart (a variant of apps/roofline_probe_cuda.cu:27-49). */
pentadStatus_t pentadsolver_probe_ir_mix(pentadsolver_handle_t handle, int streams,
                                         const float *const coef[5], const double *b,
                                         double *x, size_t n);

/* The FP32 + FP64 refinement kernel's I/O without its solve, per 8-row chunk
 * (one per thread): 10 float4 of coefficients interleaved per 4 rows (coef:
 * 5n floats, the ADI driver's layout) and 4 double2 of b read at the lane
 * strides of the solve kernel, then `rounds` rounds of 8 FP64 FMAs standing
 * in for the solve's arithmetic.  layout = 0: each wavefront's 512 rows of
 * x written with coalesced 16-byte stores, one chunk per thread at full
 * occupancy (n a multiple of 2048); layout = 1: the headline kernel's workgroup structure (256 threads,
 * two workgroups per CU, four passes of four 512-row systems into a
 * 16-system LDS tile written rotated as 128-byte runs 4 KB apart), n =
 * 512^3 exactly.  No reference counterpart (a variant of
 * apps/roofline_probe_cuda.cu:27-49). */
pentadStatus_t pentadsolver_probe_ir_layout(pentadsolver_handle_t handle, int layout,
                                            int rounds, const float *coef, const double *b,
                                            double *x, size_t n);
/* Bandwidth probe with the FP64 strided solve's I/O: along t_solvedim (1 = y
 * or 2 = z of a 3-D grid t_dims, t_dims[t_solvedim] = 512, the product of
 * the other two a multiple of 64) every line's rows of the five diagonals
 * and of x are read and x += ds + dl + d + du + dw is written in place, by
 * 512-thread workgroups of 16 * split adjacent lines x 512 / split rows
 * (split 1: whole lines, the drop-in kernel's structure; 2 or 4: each line
 * shared by 2 or 4 workgroups), every row loaded before any is stored, as
 * in the solve.  No reference counterpart (a variant of
 * apps/roofline_probe_cuda.cu:27-49). */
pentadStatus_t pentadsolver_probe_strided_io(pentadsolver_handle_t handle, int split,
                                             const double *const coef[5], double *x,
                                             const int *t_dims, int t_solvedim);

/* ADI sweep driver.  One step solves along every direction in turn (x, then
 * y, then z for ndim = 3) in place on a field x in the standard layout
 * (t_dims[0] fastest), the loop of the reference's apps/adi_cuda.cu:167-176.
 * The driver keeps its own copies of the coefficients, laid out once per
 * direction so that every direction reads contiguous lines and writes its
 * solution in the layout the next direction reads; between the steps of one
 * call x holds the field in an internal layout, and after the last step it
 * is back in the caller's layout.  x must be 16-byte aligned.
 *   mode 0: FP64 coefficients and field; mode 1: FP32 both;
 *   mode 2: FP32 coefficients, FP64 field, FP32 solve + FP64 refinement
 *           (max_refine, tol as in pentadsolver_DS_gpsv_batch_ir);
 *   mode 3 / 4: shared factorisation in FP64 / FP32 (the reference's
 *           "shared-fact", src/cuda/singlenode/pentad_cuda.cu:908-1195):
 *           each direction's system 0 is factored once at create and every
 *           step streams only the field; coeff_elem_bytes must be 8 / 4,
 *           and every length must be one the shared-fact kernels cover.
 *   coeffs: coeff_sets * 5 device pointers (ds, dl, d, du, dw per set) in
 *   the standard layout with elements of coeff_elem_bytes (4 or 8); one set
 *   serves every direction (as the reference's apps do), or ndim sets give
 *   each direction its own.
 * create allocates (coefficients: 5 * ndim copies; two field-sized
 * buffers) and synchronises the handle's stream.  No reference counterpart
 * as a library call. */
typedef void *pentadsolver_adi_t;
pentadStatus_t pentadsolver_adi_create(pentadsolver_handle_t handle,
                                       pentadsolver_adi_t *adi, int mode,
                                       int ndim, const int *dims,
                                       const void *const *coeffs,
                                       int coeff_sets, int coeff_elem_bytes,
                                       int max_refine, double tol);
/* ADI driver over structured coefficients (see
 * pentadsolver_D_gpsv_stencil_batch): along every line of direction d the
 * system is id_weights[d] * I + s * weights[d*5 .. d*5+4] (ds, dl, d - id,
 * du, dw), couplings past the ends of the line dropped, with s the line's
 * entry of scales[d].  scales: ndim device arrays, direction d's holding one
 * value per line laid out over the remaining axes in the standard order
 * (fastest first), copied at create; weights, id_weights on the host.
 *   mode 0: FP64 field and scales; mode 1: FP32 both; mode 2: FP64 field
 *   and scales, every solve in FP32 with FP64 refinement whose residual
 *   uses the FP64 operator, so the result is the FP64 operator's solution
 *   (refined to tol 1e-15, at most 5 corrections, unless
 *   pentadsolver_adi_set_refine says otherwise; lengths must split into 2..64
 *   chunks of 4, 6, 8 or 10 rows).
 * The kernels generate the diagonals from the scales, so a step streams only
 * the field (16 B per row and direction in FP64).  Every length must split
 * into 2..64 chunks of 4..16 rows.  Step, capture, replay, timing and
 * destroy are the pentadsolver_adi_* calls.  No reference counterpart. */
pentadStatus_t pentadsolver_adi_create_stencil(pentadsolver_handle_t handle,
                                               pentadsolver_adi_t *adi, int mode,
                                               int ndim, const int *dims,
                                               const void *const *scales,
                                               const double *weights,
                                               const double *id_weights);
/* nsteps ADI steps on x, launched on the handle's stream (asynchronous). */
pentadStatus_t pentadsolver_adi_step(pentadsolver_adi_t adi, void *x,
                                     int nsteps);
/* Capture nsteps steps on x as one hipGraph (stream capture of the same
 * launches); pentadsolver_adi_graph_launch replays the whole capture reps
 * times on the handle's stream.  Capture again to change x or nsteps. */
pentadStatus_t pentadsolver_adi_graph_capture(pentadsolver_adi_t adi, void *x,
                                              int nsteps);
pentadStatus_t pentadsolver_adi_graph_launch(pentadsolver_adi_t adi, int reps);
/* reps + 2 steps on x with HIP events around each direction's launch on
 * the handle's stream (synchronises); ms[d] = mean duration of direction d
 * over the reps steady-state steps.  The per-direction pass of the
 * reference's apps/adi_cuda.cu:216-256. */
pentadStatus_t pentadsolver_adi_direction_ms(pentadsolver_adi_t adi, void *x,
                                             int reps, float *ms);
/* Stopping rule of the refinement modes (pentadsolver_adi_create mode 2,
 * pentadsolver_adi_create_stencil mode 2): at most max_refine corrections,
 * stop at estimated relative forward error tol (as in
 * pentadsolver_DS_gpsv_batch_ir).  Applies to later steps and captures. */
pentadStatus_t pentadsolver_adi_set_refine(pentadsolver_adi_t adi, int max_refine,
                                           double tol);
/* Largest number of refinement steps any line took since the last call
 * (refinement modes; synchronises the stream), then resets the counter. */
int pentadsolver_adi_refine_steps(pentadsolver_adi_t adi);
/* Kernel configuration used along direction dir. */
const char *pentadsolver_adi_config(pentadsolver_adi_t adi, int dir);
void pentadsolver_adi_destroy(pentadsolver_adi_t *adi);

/* Structured coefficients: solve in place along t_solvedim (shape contract
 * of pentadsolver_D_gpsv_batch) the systems
 *     id_weight * I + scale[line] * [w0 w1 w2 w3 w4]
 * i.e. ds = s*w0, dl = s*w1, d = id_weight + s*w2, du = s*w3, dw = s*w4,
 * with the couplings that reach past either end of a line dropped.
 * scale (device): one value per line, laid out over the remaining axes in
 * the standard order (fastest first); weights (host, 5 values) and
 * id_weight as doubles for both precisions.  Implicit hyperdiffusion
 * I + nu dt/h^4 [1,-4,6,-4,1] is weights {1,-4,6,-4,1}, id_weight 1.
 * The diagonals are generated inside the kernels, so a solve reads the
 * right-hand side and writes the solution only (16 B per row in FP64 against
 * 56 B for five diagonal arrays).  Asynchronous on the handle's stream; no
 * scratch.  No reference counterpart: the reference assembles and reads
 * five diagonal arrays for every solve (its apps/preproc.hpp:248 builds
 * them). */
pentadStatus_t pentadsolver_D_gpsv_stencil_batch(pentadsolver_handle_t handle,
                                                 const double *scale,
                                                 const double *weights,
                                                 double id_weight, double *x,
                                                 const int *t_dims, int t_ndim,
                                                 int t_solvedim);
pentadStatus_t pentadsolver_S_gpsv_stencil_batch(pentadsolver_handle_t handle,
                                                 const float *scale,
                                                 const double *weights,
                                                 double id_weight, float *x,
                                                 const int *t_dims, int t_ndim,
                                                 int t_solvedim);

/* Name and tuning parameters of the kernel the last solve of this handle
 * along dir launched, e.g. "thomas-pcr rows=8 lanes=32". */
const char *pentadsolver_kernel_config(pentadsolver_handle_t handle, int dir);

/* Variable-length batch: n_sys independent systems packed back to back,
 * system k owning elements offsets[k] .. offsets[k+1]-1 of every array
 * (offsets: host array of n_sys + 1 entries; lengths 3 .. 8192).  create
 * builds a device schedule once (systems sorted longest first and grouped
 * into work items whose lanes per system follow the length; the longest
 * quarter of the items is claimed alternating with the shortest, the rest
 * longest first); each solve is
 * then one persistent launch on the handle's stream with no allocation and
 * no host synchronisation (capturable in a hipGraph).  The workgroups take
 * work items from a counter held in the plan, so two solves of one plan
 * must not run at the same time (use one plan per concurrent stream).  The
 * plan keeps the handle, which must outlive it.  No reference counterpart
 * (the reference batches systems of one length only). */
typedef void *pentadsolver_varlen_t;
pentadStatus_t pentadsolver_varlen_create(pentadsolver_handle_t handle,
                                          pentadsolver_varlen_t *plan,
                                          const int64_t *offsets, int64_t n_sys);
void pentadsolver_varlen_destroy(pentadsolver_varlen_t *plan);
/* number of work items of the schedule */
int pentadsolver_varlen_items(pentadsolver_varlen_t plan);
/* Lanes per system by length bucket: a system of c = ceil(len / 16) chunks is
 * in bucket b = ceil(log2 c) (0..9) and is solved by 2^min(b + spread[b], 9)
 * lanes (each holding up to 16 rows).  spread[10], 0 by default.  The
 * autotuners time spreads 0, 1 and 2 of every bucket present, each on the
 * caller's arrays (x is not modified: the timed solves run on a copy), then
 * the whole schedule with the fastest spread per bucket against the whole
 * default one, and keep the faster in the plan.  report (may be NULL) gets
 * one "bucket b (chunks <= c) spread s t ms" line per trial and two "whole
 * default|tuned t ms" lines.  They allocate (the copy) and synchronise;
 * return 0 or a negative code.  No reference counterpart. */
pentadStatus_t pentadsolver_varlen_set_spread(pentadsolver_varlen_t plan, const int *spread);
pentadStatus_t pentadsolver_varlen_get_spread(pentadsolver_varlen_t plan, int *spread);
int pentadsolver_D_varlen_autotune(pentadsolver_varlen_t plan, const double *ds,
                                   const double *dl, const double *d, const double *du,
                                   const double *dw, const double *x, int reps,
                                   char *report, size_t report_len);
int pentadsolver_S_varlen_autotune(pentadsolver_varlen_t plan, const float *ds,
                                   const float *dl, const float *d, const float *du,
                                   const float *dw, const float *x, int reps,
                                   char *report, size_t report_len);
int pentadsolver_DS_varlen_autotune_ir(pentadsolver_varlen_t plan, const float *ds,
                                       const float *dl, const float *d, const float *du,
                                       const float *dw, const double *x, int reps,
                                       char *report, size_t report_len);
pentadStatus_t pentadsolver_D_gpsv_varlen_batch(pentadsolver_varlen_t plan,
                                                const double *ds, const double *dl,
                                                const double *d, const double *du,
                                                const double *dw, double *x);
pentadStatus_t pentadsolver_S_gpsv_varlen_batch(pentadsolver_varlen_t plan,
                                                const float *ds, const float *dl,
                                                const float *d, const float *du,
                                                const float *dw, float *x);
/* FP32 coefficients, FP64 x.  By default one FP64 solve of the coefficients
 * widened on load: the widened system's solution to FP64 rounding in one
 * pass (what the refinement converges to), no refinement sweep (steps 0,
 * *d_refine_steps left as it is, tol and max_refine unused); the kernel
 * config reads "varlen widened ...".  pentadsolver_varlen_set_refine(plan, 1)
 * selects the FP32 solve + FP64 refinement of pentadsolver_DS_gpsv_batch_ir
 * instead ("varlen ir ..."); set_refine(plan, 0) goes back.  get_refine
 * returns the setting (-1 for a NULL plan).  No reference counterpart. */
pentadStatus_t pentadsolver_varlen_set_refine(pentadsolver_varlen_t plan, int refine);
int pentadsolver_varlen_get_refine(pentadsolver_varlen_t plan);
pentadStatus_t pentadsolver_DS_gpsv_varlen_batch_ir(
    pentadsolver_varlen_t plan, const float *ds, const float *dl, const float *d,
    const float *du, const float *dw, double *x, int max_refine, double tol,
    int *d_refine_steps);

/* Distributed line solve, one rank's share.  Every line along t_solvedim is
 * split across nranks ranks: rank r holds rows r*m .. r*m+m-1 of each line in
 * its own arrays, whose shape is t_dims (t_dims[t_solvedim] = m >= 5, the
 * same on every rank; t_pads as for pentadsolver_D_gpsv_batch_padded, NULL
 * for the dense layout, t_ndim <= 3).  This replaces the reference's
 * distributed pentadsolver_gpsv_batch (src/cpu/mpi/pentad_cpu.cpp:834-856,
 * src/cuda/mpi/pentad_cuda.cu), which runs all phases with MPI inside one
 * call; here each phase is a call and the transport between them is the
 * caller's:
 *  1. _dist_forward on every rank: slab-local elimination (reads the
 *     coefficients and x, writes nothing but the record).  record receives
 *     PENTAD_DIST_RECORD_WORDS * n_sys elements (n_sys = lines per rank),
 *     stored [word][line]: the rank's partial 2x2 block row in its
 *     separator (the last two rows of its slab of each line) and the
 *     first two rows of its slab as functions of the separators around it.
 *  2. _dist_reduced: the block tridiagonal system in the separators, one
 *     block row per rank and line, by block cyclic reduction.  records
 *     holds the nranks records back to back (rank r at r *
 *     PENTAD_DIST_RECORD_WORDS * n_sys); seps receives 4 * n_sys elements
 *     per rank, back to back: the previous rank's separator (zero on rank
 *     0) and the rank's own, [4][n_sys].  nranks <= 64.  On several GPUs
 *     this phase is an exchange of records between neighbours followed by
 *     ceil(log2 nranks) rounds of the same reduction across ranks; the call
 *     runs it for all ranks on one device.
 *  3. _dist_backward on every rank with its 4 * n_sys separator values:
 *     x (the right-hand side on input, untouched by phases 1 and 2) becomes
 *     the rank's part of the solution.  The interior of each slab is solved
 *     as an ordinary batch of lines of m - 2 rows; t_buffer holds the
 *     scratch _dist_backward_buffer_size reports (0 for the on-chip
 *     kernels).
 * Asynchronous on the handle's stream; return the status.  No reference
 * counterpart as separate calls. */
#define PENTAD_DIST_RECORD_WORDS 23
pentadStatus_t pentadsolver_D_dist_forward(pentadsolver_handle_t handle, const double *ds,
                                           const double *dl, const double *d, const double *du,
                                           const double *dw, const double *x,
                                           const int *t_dims, const int *t_pads, int t_ndim,
                                           int t_solvedim, int rank, int nranks,
                                           double *record);
pentadStatus_t pentadsolver_S_dist_forward(pentadsolver_handle_t handle, const float *ds,
                                           const float *dl, const float *d, const float *du,
                                           const float *dw, const float *x, const int *t_dims,
                                           const int *t_pads, int t_ndim, int t_solvedim,
                                           int rank, int nranks, float *record);
pentadStatus_t pentadsolver_D_dist_reduced(pentadsolver_handle_t handle, const double *records,
                                           int64_t n_sys, int nranks, double *seps);
pentadStatus_t pentadsolver_S_dist_reduced(pentadsolver_handle_t handle, const float *records,
                                           int64_t n_sys, int nranks, float *seps);
size_t pentadsolver_D_dist_backward_buffer_size(pentadsolver_handle_t handle, const double *ds,
                                                const double *dl, const double *d,
                                                const double *du, const double *dw,
                                                const double *x, const int *t_dims,
                                                const int *t_pads, int t_ndim, int t_solvedim);
size_t pentadsolver_S_dist_backward_buffer_size(pentadsolver_handle_t handle, const float *ds,
                                                const float *dl, const float *d,
                                                const float *du, const float *dw,
                                                const float *x, const int *t_dims,
                                                const int *t_pads, int t_ndim, int t_solvedim);
pentadStatus_t pentadsolver_D_dist_backward(pentadsolver_handle_t handle, const double *ds,
                                            const double *dl, const double *d, const double *du,
                                            const double *dw, double *x, const int *t_dims,
                                            const int *t_pads, int t_ndim, int t_solvedim,
                                            int rank, int nranks, const double *sep,
                                            void *t_buffer);
pentadStatus_t pentadsolver_S_dist_backward(pentadsolver_handle_t handle, const float *ds,
                                            const float *dl, const float *d, const float *du,
                                            const float *dw, float *x, const int *t_dims,
                                            const int *t_pads, int t_ndim, int t_solvedim,
                                            int rank, int nranks, const float *sep,
                                            void *t_buffer);

/* Allocation-free FP64 solve: pentadsolver_D_gpsv_batch (same kernels,
 * same results bit for bit, same errors -- except outermost lines whose
 * two-workgroup trade does not fit the array, which run one workgroup per
 * line as with a NULL t_buffer) with the scratch taken from a fixed
 * array in the library's own device memory instead of t_buffer, so a call
 * allocates nothing and can be recorded under HIP graph capture.  No
 * reference counterpart: it is for callers that cannot allocate (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Limit: a solve whose scratch need (what
 * pentadsolver_D_gpsv_batch_buffer_size_ext reports; 0 for the on-chip
 * kernels, 24 B per row for the one-lane kernel, the factor table for
 * shared-fact; the two-workgroup kernel's trade words, about 3 B per row,
 * when they fit the array, else the one-workgroup kernel runs and needs
 * none) exceeds PENTAD_NOALLOC_SCRATCH_BYTES is refused with
 * PENTAD_STATUS_INVALID_VALUE and a message on stderr, and nothing is
 * launched.  Every GPU the library runs on holds one such array until the
 * process ends (the device address is looked up at pentadsolver_create, or
 * at the first call on a device).  All calls on one GPU share that GPU's
 * array: calls that overlap in time on one GPU (other streams, other host
 * threads) race and give wrong results; calls queued on one stream do not. */
#define PENTAD_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
pentadStatus_t pentadsolver_D_gpsv_batch_noalloc(pentadsolver_handle_t handle,
                                                 const double *ds, const double *dl,
                                                 const double *d, const double *du,
                                                 const double *dw, double *x,
                                                 const int *t_dims, int t_ndim,
                                                 int t_solvedim);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PENTADSOLVER_H_INCLUDED */
