// pentad_dist.cpp -- C ABI of the distributed line solve, one rank's share.
//
// The reference's distributed solver (src/cpu/mpi/pentad_cpu.cpp:834-856,
// src/cuda/mpi/pentad_cuda.cu) is one pentadsolver_gpsv_batch call per rank
// on its local arrays, with MPI inside: local elimination, an exchange of
// per-line records with the neighbours, a reduced solve over the ranks, and
// local back-substitution.  Here the phases are separate calls so that the
// transport between them belongs to the caller (RCCL point-to-point on
// several GPUs); on one GPU the records of every rank sit in one buffer and
// the reduced phase runs on it directly.  The kernels are in
// kernels_dist.hip.
#include "pentad_api_internal.hpp"

#include <cstdio>

namespace pentad {
namespace {

template <typename T>
pentadStatus_t dist_forward(Context *c, const T *ds, const T *dl, const T *d, const T *du,
                            const T *dw, const T *x, const int *dims, const int *pads,
                            int ndim, int solvedim, int rank, int nranks, T *rec) {
  Lines ln;
  if (decompose_padded(dims, pads, ndim, solvedim, &ln) != 0 || ln.len < 5 || nranks < 1 ||
      rank < 0 || rank >= nranks) {
    std::fprintf(stderr,
                 "[penta] dist forward: invalid arguments (solvedim %d, rank %d of %d; every "
                 "rank needs at least 5 rows of each line)\n",
                 solvedim, rank, nranks);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (!ds || !dl || !d || !du || !dw || !x || !rec) {
    std::fprintf(stderr, "[penta] dist forward: null array pointer\n");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  const Diags<T> a{ds, dl, d, du, dw};
  const hipError_t e =
      launch_dist_forward<T>(a, x, ln, rank == 0, rank == nranks - 1, rec, c->stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] dist forward launch FAILED: %s\n", hipGetErrorString(e));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  return PENTAD_STATUS_SUCCESS;
}

template <typename T>
pentadStatus_t dist_reduced(Context *c, const T *recs, int64_t n_sys, int nranks, T *seps) {
  if (!recs || !seps || n_sys < 1 || nranks < 1 || nranks > 64) {
    std::fprintf(stderr, "[penta] dist reduced: invalid arguments (%lld lines, %d ranks; at "
                         "most 64 ranks)\n",
                 (long long)n_sys, nranks);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  const hipError_t e = launch_dist_reduced<T>(recs, n_sys, nranks, seps, c->stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] dist reduced launch FAILED: %s\n", hipGetErrorString(e));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  return PENTAD_STATUS_SUCCESS;
}

// the interior of every line: rows 0 .. m-3 of the rank's arrays, a padded
// view with the solve axis cut to m-2 rows
bool interior_view(const int *dims, const int *pads, int ndim, int solvedim, int *idims,
                   int *ipads) {
  if (ndim < 1 || ndim > 3 || solvedim < 0 || solvedim >= ndim) return false;
  for (int i = 0; i < ndim; ++i) {
    idims[i] = dims[i];
    ipads[i] = pads ? pads[i] : dims[i];
  }
  idims[solvedim] -= 2;
  return idims[solvedim] >= 3;
}

template <typename T>
pentadStatus_t dist_backward(Context *c, const T *ds, const T *dl, const T *d, const T *du,
                             const T *dw, T *x, const int *dims, const int *pads, int ndim,
                             int solvedim, int rank, int nranks, const T *sep, void *buffer) {
  Lines ln;
  int idims[3], ipads[3];
  if (decompose_padded(dims, pads, ndim, solvedim, &ln) != 0 || ln.len < 5 || nranks < 1 ||
      rank < 0 || rank >= nranks || !interior_view(dims, pads, ndim, solvedim, idims, ipads)) {
    std::fprintf(stderr,
                 "[penta] dist backward: invalid arguments (solvedim %d, rank %d of %d, at "
                 "most 3 dimensions, every rank needs at least 5 rows of each line)\n",
                 solvedim, rank, nranks);
    return PENTAD_STATUS_INVALID_VALUE;
  }
  if (!ds || !dl || !d || !du || !dw || !x || !sep) {
    std::fprintf(stderr, "[penta] dist backward: null array pointer\n");
    return PENTAD_STATUS_INVALID_VALUE;
  }
  const Diags<T> a{ds, dl, d, du, dw};
  const hipError_t e = launch_dist_halo<T>(a, x, ln, rank == 0, sep, c->stream);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[penta] dist backward launch FAILED: %s\n", hipGetErrorString(e));
    return PENTAD_STATUS_EXECUTION_FAILED;
  }
  // the interior: an ordinary batch of lines, the planner's fastest kernel
  return gpsv_batch_pitched<T>(c, ds, dl, d, du, dw, x, idims, ipads, ndim, solvedim, buffer,
                               SIZE_MAX);
}

template <typename T>
size_t dist_backward_buffer_size(Context *c, const T *ds, const T *dl, const T *d,
                                 const T *du, const T *dw, const T *x, const int *dims,
                                 const int *pads, int ndim, int solvedim) {
  int idims[3], ipads[3];
  if (dims == nullptr || !interior_view(dims, pads, ndim, solvedim, idims, ipads)) return 0;
  Lines ln;
  if (decompose_padded(idims, ipads, ndim, solvedim, &ln) != 0) return 0;
  const void *ptrs[6] = {ds, dl, d, du, dw, x};
  return plan_scratch_bytes(plan_solve(c, ln, (int)sizeof(T), ptrs), ln, (int)sizeof(T));
}

} // namespace
} // namespace pentad

using namespace pentad;

extern "C" {

pentadStatus_t pentadsolver_D_dist_forward(pentadsolver_handle_t h, const double *ds,
                                           const double *dl, const double *d, const double *du,
                                           const double *dw, const double *x,
                                           const int *t_dims, const int *t_pads, int t_ndim,
                                           int t_solvedim, int rank, int nranks,
                                           double *record) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = dist_forward<double>(c, ds, dl, d, du, dw, x, t_dims, t_pads, t_ndim,
                                                t_solvedim, rank, nranks, record);
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_S_dist_forward(pentadsolver_handle_t h, const float *ds,
                                           const float *dl, const float *d, const float *du,
                                           const float *dw, const float *x, const int *t_dims,
                                           const int *t_pads, int t_ndim, int t_solvedim,
                                           int rank, int nranks, float *record) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = dist_forward<float>(c, ds, dl, d, du, dw, x, t_dims, t_pads, t_ndim,
                                               t_solvedim, rank, nranks, record);
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_D_dist_reduced(pentadsolver_handle_t h, const double *records,
                                           int64_t n_sys, int nranks, double *seps) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = dist_reduced<double>(c, records, n_sys, nranks, seps);
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_S_dist_reduced(pentadsolver_handle_t h, const float *records,
                                           int64_t n_sys, int nranks, float *seps) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = dist_reduced<float>(c, records, n_sys, nranks, seps);
  set_status(c, s);
  return s;
}

size_t pentadsolver_D_dist_backward_buffer_size(pentadsolver_handle_t h, const double *ds,
                                                const double *dl, const double *d,
                                                const double *du, const double *dw,
                                                const double *x, const int *t_dims,
                                                const int *t_pads, int t_ndim, int t_solvedim) {
  return dist_backward_buffer_size<double>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_pads,
                                           t_ndim, t_solvedim);
}

size_t pentadsolver_S_dist_backward_buffer_size(pentadsolver_handle_t h, const float *ds,
                                                const float *dl, const float *d,
                                                const float *du, const float *dw,
                                                const float *x, const int *t_dims,
                                                const int *t_pads, int t_ndim, int t_solvedim) {
  return dist_backward_buffer_size<float>(ctx_of(h), ds, dl, d, du, dw, x, t_dims, t_pads,
                                          t_ndim, t_solvedim);
}

pentadStatus_t pentadsolver_D_dist_backward(pentadsolver_handle_t h, const double *ds,
                                            const double *dl, const double *d, const double *du,
                                            const double *dw, double *x, const int *t_dims,
                                            const int *t_pads, int t_ndim, int t_solvedim,
                                            int rank, int nranks, const double *sep,
                                            void *t_buffer) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = dist_backward<double>(c, ds, dl, d, du, dw, x, t_dims, t_pads, t_ndim,
                                                 t_solvedim, rank, nranks, sep, t_buffer);
  set_status(c, s);
  return s;
}

pentadStatus_t pentadsolver_S_dist_backward(pentadsolver_handle_t h, const float *ds,
                                            const float *dl, const float *d, const float *du,
                                            const float *dw, float *x, const int *t_dims,
                                            const int *t_pads, int t_ndim, int t_solvedim,
                                            int rank, int nranks, const float *sep,
                                            void *t_buffer) {
  Context *c = ctx_of(h);
  const pentadStatus_t s = dist_backward<float>(c, ds, dl, d, du, dw, x, t_dims, t_pads, t_ndim,
                                                t_solvedim, rank, nranks, sep, t_buffer);
  set_status(c, s);
  return s;
}

} // extern "C"
