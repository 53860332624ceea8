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
 * for y/z, an alias of "naive").  A handle setting wins over the
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

/* Name and tuning parameters of the kernel the last solve of this handle
 * along dir launched, e.g. "thomas-pcr rows=8 lanes=32". */
const char *pentadsolver_kernel_config(pentadsolver_handle_t handle, int dir);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* PENTADSOLVER_H_INCLUDED */
