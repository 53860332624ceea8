// pentadsolver.hpp -- C++ drop-in header for libpentadsolver.
//
// Same declarations as the reference's include/pentadsolver.hpp:36-131: the
// C entry points of pentadsolver.h plus the C++ overloads
// pentadsolver_gpsv_batch_buffer_size_ext(...) and pentadsolver_gpsv_batch(...)
// for double and float (reference :75-84, :120-131).  The overloads are
// exported by the library with C++ linkage, like the reference's.
#ifndef PENTADSOLVER_HPP_INCLUDED
#define PENTADSOLVER_HPP_INCLUDED

#include "pentadsolver.h"
#include "pentadsolver_handle.hpp"

#include <cstddef>

[[nodiscard]] size_t pentadsolver_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t handle, const double *ds, const double *dl,
    const double *d, const double *du, const double *dw, const double *x,
    const int *t_dims, int t_ndim, int t_solvedim);
[[nodiscard]] size_t pentadsolver_gpsv_batch_buffer_size_ext(
    pentadsolver_handle_t handle, const float *ds, const float *dl,
    const float *d, const float *du, const float *dw, const float *x,
    const int *t_dims, int t_ndim, int t_solvedim);

void pentadsolver_gpsv_batch(pentadsolver_handle_t handle, const double *ds,
                             const double *dl, const double *d,
                             const double *du, const double *dw, double *x,
                             const int *t_dims, int t_ndim, int t_solvedim,
                             void *t_buffer);
void pentadsolver_gpsv_batch(pentadsolver_handle_t handle, const float *ds,
                             const float *dl, const float *d, const float *du,
                             const float *dw, float *x, const int *t_dims,
                             int t_ndim, int t_solvedim, void *t_buffer);

#endif // PENTADSOLVER_HPP_INCLUDED
