// pentadsolver_handle.hpp -- handle type, as in the reference's
// include/pentadsolver_handle.hpp:4-5 for single-node builds: an opaque
// pointer.  (The reference's MPI handle struct, :6-43, belongs to its
// multi-process backends, which this library does not provide.)
#ifndef PENTADSOLVER_HANDLE_HPP_INCLUDED
#define PENTADSOLVER_HANDLE_HPP_INCLUDED

#include "pentadsolver.h" // typedef void *pentadsolver_handle_t

#endif // PENTADSOLVER_HANDLE_HPP_INCLUDED
