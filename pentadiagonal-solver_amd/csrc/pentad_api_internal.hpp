// pentad_api_internal.hpp -- handle contents and dispatch helpers shared by
// the host sources of libpentadsolver.  Not part of the public interface.
#pragma once

#include "pentad_internal.hpp"
#include "pentadsolver.h"

#include <map>
#include <mutex>
#include <string>

namespace pentad {

// Which kernel a solve launches and with what shape parameters.
struct Plan {
  enum Kind {
    Naive = 0, PartX = 1, PartYZ = 2, IrLane = 3, IrX = 4, IrYZ = 5, SfX = 6, SfYZ = 7,
    TrX = 8, IrTrX = 9, IrWideYZ = 10, IrWideX = 11
  } kind = Naive;
  int rows = 0;       // rows per lane (partitioned kernels)
  int lanes = 0;      // lanes per system (x)
  int sys_per_wg = 0; // systems per workgroup (y/z)
  bool split = false; // PartYZ / IrWideYZ: each line shared by two workgroups
};

struct TuneKey {
  int elem_bytes;
  int dir;
  int64_t len;
  bool operator<(const TuneKey &o) const {
    if (elem_bytes != o.elem_bytes) return elem_bytes < o.elem_bytes;
    if (dir != o.dir) return dir < o.dir;
    return len < o.len;
  }
};
using TuneVal = Plan;
constexpr int TUNE_IR = 12; // TuneKey::elem_bytes of the FP32+FP64 solve

struct Context {
  std::mutex mu;
  int device = 0;
  hipStream_t stream = nullptr;
  pentadStatus_t last_status = PENTAD_STATUS_SUCCESS;
  std::string algo[3];     // per-direction override ("" = none)
  bool env_frozen = false; // use env_algo[] instead of reading getenv
  std::string env_algo[3] = {"auto", "auto", "auto"};
  const char *ran[3] = {"none", "none", "none"};
  std::string ran_cfg[3] = {"none", "none", "none"};
  std::map<TuneKey, TuneVal> tune; // autotuner results
  int force_rows[3] = {0, 0, 0};   // pentadsolver_set_kernel_params
  int force_width[3] = {0, 0, 0};
};

Context *ctx_of(pentadsolver_handle_t h);
void destroy_context_resources(Context *c);
void note_kernel(Context *c, int dir, const char *name, const std::string &cfg);
const char *ran_global(int dir);
void set_status(Context *c, pentadStatus_t s);
std::string requested_algo(Context *c, int dir);
std::string env_algo(int dir);
bool debug_launch();

// allow_split: the plan may share the lines of the outermost direction
// between two workgroups (scratch for their trade in the caller's buffer);
// the batched solve asks again without it when the caller passes none.
Plan plan_solve(Context *c, const Lines &ln, int elem_bytes, const void *const ptrs[6],
                bool allow_split = false);
Plan plan_ir(Context *c, const Lines &ln, const void *const ptrs[6], bool allow_split = false);
size_t plan_scratch_bytes(const Plan &pl, const Lines &ln, int elem_bytes);
// scratch of the FP64 batched solve when the caller's buffer holds cap bytes
// (the allocation-free entry point's fixed array)
size_t noalloc_need(Context *c, const double *ds, const double *dl, const double *d,
                    const double *du, const double *dw, const double *x, const int *dims,
                    int ndim, int solvedim, size_t cap);
template <typename T>
hipError_t launch_plan(const Plan &pl, const Diags<T> &a, T *x, const Lines &ln,
                       void *buffer, hipStream_t st);
hipError_t launch_plan_ir(const Plan &pl, const Diags<float> &a, double *x,
                          const Lines &ln, void *buffer, int max_steps,
                          double tol, int *steps, hipStream_t st);
std::string plan_config(const Plan &pl);
// the reference's name of a plain-solve plan: "naive", "transpose",
// "thomas-pcr" or "shared-fact" (what pentadsolver_kernel_that_ran reports)
const char *plan_algo_name(const Plan &pl);
// The batched solve along one axis (pads == nullptr: the dense layout).
// buffer holds buffer_cap bytes (SIZE_MAX: sized by the caller with the
// matching _buffer_size query); a kernel whose scratch exceeds it is not
// launched and the call returns PENTAD_STATUS_INVALID_VALUE.
template <typename T>
pentadStatus_t gpsv_batch_pitched(Context *c, const T *ds, const T *dl, const T *d,
                                  const T *du, const T *dw, T *x, const int *dims,
                                  const int *pads, int ndim, int solvedim, void *buffer,
                                  size_t buffer_cap);
// PENTA_AUTOTUNE=1: tune a (precision, direction, length) the first time it
// is solved with the automatic choice (pentad_tune.cpp); true when it tuned
bool autotune_env();
template <typename T>
bool autotune_on_first_use(Context *c, const T *ds, const T *dl, const T *d, const T *du,
                           const T *dw, T *x, const int *dims, int ndim, int solvedim);
// Device address of the fixed scratch array of the allocation-free solve on
// device dev (looked up once per device; pentad_noalloc.hip).
hipError_t noalloc_scratch(int dev, void **out);
// x-like partition (rows per lane, lanes per system) of the structured-
// coefficient kernels for a line length; false when none fits
bool stencil_plan_x(int elem_bytes, int64_t len, Plan &pl);

} // namespace pentad
